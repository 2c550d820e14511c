// C ABI of libpbf.so (include/pbf.h): context, parameters and entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "pbf.h"
#include "pbf_internal.h"

using namespace pbf;

// Scratch of pbf_pbd_constrain_noalloc: the iterate positions (float4) and the
// displacement sums (3 floats) of up to PBF_NOALLOC_MAX_PARTICLES particles,
// 28 MiB on every device the library runs on, held until the process ends.
// In host code the name is no device address: each device's copy is looked
// up once (scratch_address) and only that pointer reaches kernels.
static __device__ __align__(256) unsigned char
    scratch[(size_t)PBF_NOALLOC_MAX_PARTICLES * (sizeof(float4) + 3 * sizeof(float))];

struct pbf_ctx {
  pbf_params p{};
  GridDev g{};
  PhysConst pc{};
  long long full_cells = 0;  // sort-grid cells of one scene: 2nx*ny*nz (half cells along x)
  float m_factor = 0.f, mass = 0.f;
  int device = 0;

  // scenes
  int scene_cap = 0;
  int* off1 = nullptr;          // offsets {0, n} of a one-scene call
  int* scene_minmax = nullptr;  // 6 per scene
  SceneWin* win = nullptr;
  int* total_cells = nullptr;
  int* scan_part = nullptr;     // kScanDevBlocks + 1

  // compact cell arrays
  long long cell_cap = 0;
  int* cell_cnt = nullptr;    // cell_cap + 1
  int* cell_start = nullptr;  // cell_cap + 2

  // particle arrays (capacity cap_n)
  int cap_n = 0;
  int *key = nullptr, *slot = nullptr, *tmp = nullptr, *perm = nullptr, *inv = nullptr;
  int *sid = nullptr, *ssid = nullptr, *bpart = nullptr;
  int *tid = nullptr, *pid = nullptr;  // original indices: of tmp's entries, of the state kept in cell order
  int *nbr_cnt = nullptr, *cnt_orig = nullptr;
  float4 *xn = nullptr, *x0 = nullptr, *xa = nullptr, *xb = nullptr, *vs = nullptr, *vv = nullptr;
  float* xc = nullptr;  // sorted x^n as three coordinate planes of cap + kXnPad floats
  float* dpos = nullptr;
  // neighbor pass output: per-wave metadata, bit masks (scratch), lists
  long long* mask_base = nullptr;
  long long* list_base = nullptr;
  int* list_jbase = nullptr;
  int* list_mode = nullptr;
  int* stage = nullptr;         // per wave: kStageInts ints (staged lists)
  uint16_t* cnt_a = nullptr;    // per sorted particle: stage-slot entries of its list (staged lists)
  int* oog_rcnt = nullptr;      // per sorted particle: out-of-grid particles that list it
  int* oog_rev = nullptr;       // kOogRev per sorted particle: those particles
  int* oog_flag = nullptr;      // per original index: out of the grid (cap + 1, zero between sorts)
  int* oog_pre = nullptr;       // its exclusive scan (cap + 1): a particle's slot in its scene's bucket
  uint32_t* mask = nullptr;
  long long mask_cap = 0;  // words
  uint16_t* list = nullptr;
  long long list_cap = 0;  // 16-bit slots
  long long auto_list_per = 96;  // automatic list reservation, 16-bit slots per particle
  int4* wave_stat = nullptr;     // per wave of the neighbor pass: pairs, candidates, need
  int stats_waves = 0;           // waves of the last search whose statistics are not yet reduced

  int* scan_scratch = nullptr;
  size_t scan_scratch_n = 0;

  // viscosity constants and the sampled box field
  VisConst vc{};
  BoxField bf{};
  size_t bf_cap = 0;     // allocated samples
  bool bf_valid = false; // samples match the parameters
  float keep_normal = 0.f, damp = 1.f;

  // rollout with checkpoints (reverse mode): states every `every` steps,
  // the states of one segment, and the tape of one step
  struct Rollout {
    bool valid = false;
    int n = 0, nscenes = 0, nsteps = 0, every = 0, iters = 0;
    pbf_params p{};
    int* off = nullptr;  // nscenes + 1
    int off_cap = 0;
    float* ckpt = nullptr;  // per checkpoint: x (3n) then v (3n)
    size_t ckpt_cap = 0;    // floats
    float* seg = nullptr;   // per step of a segment: x (3n) then v (3n)
    size_t seg_cap = 0;
    float4* adj = nullptr;   // q1, q2, c, vsb (cap_n each)
    int adj_n = 0;
    float* out = nullptr;    // x, v written by a taped step (discarded)
    int out_n = 0;
    int* flag = nullptr;
  } roll;

  // Tape of one step of a segment: the arrays its neighbor search wrote
  // (swapped in for the step), copies of its cell starts, windows and
  // counters, and its PBD iterates and per-iteration sums.
  struct TapeSlot {
    int *perm = nullptr, *ssid = nullptr, *nbr_cnt = nullptr;
    float4 *xn = nullptr, *x0 = nullptr;
    long long *list_base = nullptr, *mask_base = nullptr;
    int *list_jbase = nullptr, *list_mode = nullptr, *stage = nullptr;
    uint16_t *cnt_a = nullptr, *list = nullptr;
    uint32_t* mask = nullptr;
    int* cell_start = nullptr;
    SceneWin* win = nullptr;
    unsigned long long* counters = nullptr;
    float4* it = nullptr;  // iterates 1..K-1, then the K aux arrays (cap_n each)
    double4* it64 = nullptr;  // double-precision rollout: P_0..P_{K-1}, then the K aux arrays (cap_n each)
    size_t it64_n = 0;        // double4 entries allocated
  };
  // rollout in double precision: checkpoints and segment states (x then v,
  // 3n doubles each), the rounded positions of the search, the iterates of an
  // untaped step (P_0, two ping-pong iterates, the integrated velocities) and
  // the adjoint arrays (e, e2, f, c, q1, q2, v*_bar)
  struct Rollout64 {
    bool valid = false;
    int n = 0, nscenes = 0, nsteps = 0, every = 0, iters = 0;
    pbf_params p{};
    int* off = nullptr;
    int off_cap = 0;
    double* ckpt = nullptr;
    size_t ckpt_cap = 0;
    double* seg = nullptr;
    size_t seg_cap = 0;
    float* xf = nullptr;
    size_t xf_cap = 0;
    double4* w4 = nullptr;  // 4 x cap_n
    size_t w4_cap = 0;
    double4* adj = nullptr;  // 7 x cap_n
    size_t adj_cap = 0;
    double* out = nullptr;  // 6 x cap_n: x, v written by a taped step (discarded)
    size_t out_cap = 0;
    int* flag = nullptr;
  } roll64;
  PhysConst64 pc64{};

  std::vector<TapeSlot> slots;
  size_t slot_n = 0, slot_list = 0, slot_mask = 0;
  long long slot_cells = 0;
  int slot_scenes = 0, slot_K = -1;

  // Captured middle step of a multi-step call (hipGraph), replayed for every
  // step between the first and the last: valid while gen, n, the scenes and
  // the iteration count are those of the capture. gen is bumped whenever a
  // device buffer of the context is reallocated or swapped or a kernel
  // constant changes (a graph holds pointers and constants by value).
  hipStream_t cap_stream = nullptr;
  unsigned long long gen = 0;
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    unsigned long long gen = ~0ull;
    int n = -1, nscenes = -1, iters = -1;
    const int* off = nullptr;
  } sg;
  long long graph_replays = 0, graph_captures = 0;

  unsigned long long* counters = nullptr;     // device, C_COUNT
  unsigned long long* h_counters = nullptr;   // pinned host copy
  hipEvent_t counters_ev = nullptr;
  bool counters_pending = false;

  int last_n = -1;  // particle count of the last neighbor search (-1: none)

  int pass_block = kBlock;  // threads per workgroup of the PBD passes (pbf_tuning)
  bool pass_persistent = false;  // persistent launches of the PBD passes (pbf_tuning)
  int pass_lanes = 0;            // lanes per particle in the PBD passes, 0 automatic (pbf_tuning)
  int last_lanes = 0;            // lanes per particle of the last step's PBD passes

  void* noalloc_scratch = nullptr;  // this device's copy of `scratch` (null: lookup failed, noalloc_err says why)
  std::string noalloc_err;

  // per-pass timing
  bool profiling = false;
  struct Mark {
    int pass;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;        // recorded, not yet summed
  std::vector<hipEvent_t> spare;  // reusable events
  double prof_ms[PBF_PASS_COUNT] = {0, 0, 0, 0};
  long long prof_n[PBF_PASS_COUNT] = {0, 0, 0, 0};

  std::string err;
};

namespace {

int fail(pbf_ctx* c, int status, const std::string& msg) {
  if (c) c->err = msg;
  return status;
}

// A failed call's error is also HIP's last error: it is cleared here, so that
// the launch check of a later call (hipGetLastError) does not report it again.
#define PBF_HIP(ctx, call)                                                                           \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) {                                                                          \
      (void)hipGetLastError();                                                                       \
      return fail(ctx, e_ == hipErrorOutOfMemory ? PBF_ERR_ALLOC : PBF_ERR_HIP,                      \
                  std::string(#call) + ": " + hipGetErrorString(e_));                               \
    }                                                                                                \
  } while (0)

#define PBF_LAUNCHED(ctx)                                                                            \
  do {                                                                                               \
    hipError_t e_ = hipGetLastError();                                                               \
    if (e_ != hipSuccess)                                                                            \
      return fail(ctx, PBF_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_));         \
  } while (0)

template <class T> void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

constexpr double kPi = 3.14159265358979323846;  // M_PI, Core/Utility/cuda_utilities.h:31

// SpikyKernel::Weight (Dynamics/ParticleSystem/Kernel.h:36-45), host side.
float spiky_w_host(float r, float h) {
#pragma clang fp contract(off)
  const float q = r / h;
  if (q > 1.0f) return 0.0f;
  const float d = (float)(1.0 - (double)q);
  const float hh = h * h;
  return 15.0f / ((float)kPi * hh * h) * d * d * d * 1.0f;
}

// SummationDensity::calculateScalingFactor (SummationDensity.cu:113-138): the
// lattice sum that makes a perfect lattice of spacing d have the rest density.
float scaling_factor_host(float d, float H) {
#pragma clang fp contract(off)
  const float V = d * d * d;
  float total = 0.0f;
  const int half_res = (int)(H / d + 1.0f);
  for (int i = -half_res; i <= half_res; i++)
    for (int j = -half_res; j <= half_res; j++)
      for (int k = -half_res; k <= half_res; k++) {
        const float x = (float)i * d, y = (float)j * d, z = (float)k * d;
        const float r = std::sqrt(x * x + y * y + z * z);
        total += V * spiky_w_host(r, H);
      }
  return (float)(1.0 / (double)total);
}

// Least float s >= 0 with sqrtf(s) >= h (IEEE square root is monotone), so
// that the strict test sqrtf(s) < h of NeighborQuery.cu:267 is s < T.
float sqrt_threshold(float h) {
  uint32_t lo = 0, hi = 0x7f800000u;  // [+0, +inf]
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    float s;
    std::memcpy(&s, &mid, 4);
    if (std::sqrt(s) >= h)
      hi = mid;
    else
      lo = mid + 1;
  }
  float t;
  std::memcpy(&t, &lo, 4);
  return t;
}

int validate(const pbf_params* p, std::string& why) {
  auto pos = [](float v) { return std::isfinite(v) && v > 0.0f; };
  if (!pos(p->smoothing_length)) return why = "smoothing_length must be finite and > 0", PBF_ERR_INVALID;
  if (!pos(p->sampling_distance)) return why = "sampling_distance must be finite and > 0", PBF_ERR_INVALID;
  if (!pos(p->dt)) return why = "dt must be finite and > 0", PBF_ERR_INVALID;
  if (!std::isfinite(p->rest_density)) return why = "rest_density must be finite", PBF_ERR_INVALID;
  if (p->iterations < 0 || p->iterations > 1000) return why = "iterations must be in [0, 1000]", PBF_ERR_INVALID;
  if (p->mask_words < 0) return why = "mask_words must be >= 0", PBF_ERR_INVALID;
  if (p->list_slots < 0) return why = "list_slots must be >= 0", PBF_ERR_INVALID;
  if (p->flags & ~(PBF_FLAG_VISCOSITY | PBF_FLAG_BOX | PBF_FLAG_EAGER)) return why = "unknown flag bits", PBF_ERR_INVALID;
  if (!std::isfinite(p->viscosity) || p->viscosity < 0.0f) return why = "viscosity must be finite and >= 0", PBF_ERR_INVALID;
  if (p->viscosity_iterations < 0 || p->viscosity_iterations > 1000)
    return why = "viscosity_iterations must be in [0, 1000]", PBF_ERR_INVALID;
  if (!pos(p->box_spacing)) return why = "box_spacing must be finite and > 0", PBF_ERR_INVALID;
  if (!std::isfinite(p->normal_friction) || !std::isfinite(p->tangent_friction))
    return why = "frictions must be finite", PBF_ERR_INVALID;
  for (int c = 0; c < 3; ++c)
    if (!std::isfinite(p->box_lo[c]) || !std::isfinite(p->box_hi[c]) || !(p->box_hi[c] > p->box_lo[c]))
      return why = "box bounds must be finite with hi > lo", PBF_ERR_INVALID;
  for (int c = 0; c < 3; ++c) {
    if (!std::isfinite(p->scene_lo[c]) || !std::isfinite(p->scene_hi[c]) || !(p->scene_hi[c] > p->scene_lo[c]))
      return why = "scene bounds must be finite with hi > lo", PBF_ERR_INVALID;
    if (!std::isfinite(p->gravity[c])) return why = "gravity must be finite", PBF_ERR_INVALID;
  }
  if (!std::isfinite(p->lambda_rest) || !std::isfinite(p->lambda_eps) || !std::isfinite(p->corr_scale) ||
      !std::isfinite(p->pair_eps))
    return why = "lambda/correction constants must be finite", PBF_ERR_INVALID;
  return PBF_OK;
}

// SpikyKernel::Weight and SummationDensity::calculateScalingFactor as above,
// in double (the parameters promoted to double).
double spiky_w_host64(double r, double h) {
#pragma clang fp contract(off)
  const double q = r / h;
  if (q > 1.0) return 0.0;
  const double d = 1.0 - q;
  const double hh = h * h;
  return 15.0 / (kPi * hh * h) * d * d * d * 1.0;
}

double scaling_factor_host64(double d, double H) {
#pragma clang fp contract(off)
  const double V = d * d * d;
  double total = 0.0;
  const int half_res = (int)(H / d + 1.0);
  for (int i = -half_res; i <= half_res; i++)
    for (int j = -half_res; j <= half_res; j++)
      for (int k = -half_res; k <= half_res; k++) {
        const double x = (double)i * d, y = (double)j * d, z = (double)k * d;
        const double r = std::sqrt(x * x + y * y + z * z);
        total += V * spiky_w_host64(r, H);
      }
  return 1.0 / total;
}

PhysConst64 derive64(const pbf_params* p) {
#pragma clang fp contract(off)
  PhysConst64 pc{};
  const double h = p->smoothing_length, d = p->sampling_distance;
  pc.h = h;
  pc.inv_h = 1.0 / h;
  const double mass = scaling_factor_host64(d, h) * ((double)p->rest_density * (d * d * d));
  pc.cW = mass * (15.0 / (kPi * (h * h) * h));
  pc.cG = -45.0 / (kPi * (h * h) * h);
  pc.lam_rest = p->lambda_rest;
  pc.lam_eps = p->lambda_eps;
  pc.corr = p->corr_scale;
  pc.pair_eps = p->pair_eps;
  pc.dt = p->dt;
  for (int a = 0; a < 3; ++a) pc.g[a] = p->gravity[a];
  return pc;
}

// GridHash::setSpace (Framework/Topology/GridHash.cu:47-75) and the constants
// of the passes.
int derive(pbf_ctx* c, const pbf_params* p) {
#pragma clang fp contract(off)
  GridDev g{};
  const float ds = p->smoothing_length;
  const float pad = 2.0f * ds;
  long long n[3];
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = p->scene_lo[a] - pad;
    const float nseg = (p->scene_hi[a] - p->scene_lo[a]) / ds;
    const float nf = std::ceil(nseg) + 1.0f + 4.0f;
    if (!(nf < 2.0e6f)) return fail(c, PBF_ERR_UNSUPPORTED, "grid too large along one axis");
    n[a] = (long long)nf;
  }
  const long long ncell = 2 * n[0] * n[1] * n[2];  // half cells along x
  if (ncell > (1ll << 30)) return fail(c, PBF_ERR_UNSUPPORTED, "grid has more than 2^29 cells");
  g.ds = ds;
  g.nx = (int)n[0];
  g.ny = (int)n[1];
  g.nz = (int)n[2];
  // Half-cell distance of a neighbor pair beyond 2 that rounding can produce:
  // s < thr2 bounds the true distance by h (1 + 2.1e-7); each q = (p - lo)/ds
  // carries two roundings, at most 1.2e-7 q_max (q_max = nx + 2 after the
  // clamp); in half cells 2 (2.1e-7 + 2 * 1.2e-7 q_max); four times that.
  g.ext_margin = 4.0f * (4.2e-7f + 4.8e-7f * (float)(n[0] + 2));

  PhysConst pc{};
  const float h = p->smoothing_length;
  pc.h = h;
  pc.inv_h = 1.0f / h;
  pc.thr2 = sqrt_threshold(h);
  const float m_factor = scaling_factor_host(p->sampling_distance, h);
  const float d = p->sampling_distance;
  const float mass = p->rest_density * (d * d * d);
  const float hh = h * h;
  pc.cW = (m_factor * mass) * (15.0f / ((float)kPi * hh * h));
  pc.cG = -45.0f / ((float)kPi * hh * h);
  pc.lam_rest = p->lambda_rest;
  pc.lam_eps = p->lambda_eps;
  pc.corr = p->corr_scale;
  pc.pair_eps = p->pair_eps;
  pc.dt = p->dt;
  for (int a = 0; a < 3; ++a) {
    pc.g[a] = p->gravity[a];
    pc.gdt[a] = (0.0f + p->gravity[a]) * p->dt;
  }
  // ImplicitViscosity (ImplicitViscosity.cu:9-18,52): w = 45/(13 pi h^3)(1 - r/h)
  VisConst vc{};
  vc.inv_h = pc.inv_h;
  vc.cw = 45.0f / (13.0f * (float)kPi * hh * h);
  vc.b = p->dt * p->viscosity / h;
  vc.eps = p->pair_eps;
  // box field of BoundaryConstraint::setCube (BoundaryConstraint.cu:109-117)
  // and DistanceField3D::setSpace (DistanceField3D.cu:194-201)
  BoxField bf{};
  long long nsamp = 1;
  for (int a = 0; a < 3; ++a) {
    const float nseg = std::floor((p->box_hi[a] - p->box_lo[a]) / p->box_spacing);
    if (!(nseg < 1.0e5f)) return fail(c, PBF_ERR_UNSUPPORTED, "box field too fine along one axis");
    const int nb = (int)nseg + 10;
    const float pad = 5.0f * p->box_spacing;
    const float p0 = p->box_lo[a] - pad, p1 = p->box_hi[a] + pad;
    bf.left[a] = p0;
    bf.h[a] = (p1 - p0) * (float)(1.0 / (double)(float)(nb + 1));
    bf.inv_h[a] = (float)(1.0 / (double)bf.h[a]);
    (a == 0 ? bf.nx : a == 1 ? bf.ny : bf.nz) = nb + 1;
    nsamp *= nb + 1;
  }
  if ((p->flags & PBF_FLAG_BOX) && nsamp > (1ll << 31))
    return fail(c, PBF_ERR_UNSUPPORTED, "box field has more than 2^31 samples");
  const bool box_same = c->bf_valid && std::memcmp(c->p.box_lo, p->box_lo, sizeof(p->box_lo)) == 0 &&
                        std::memcmp(c->p.box_hi, p->box_hi, sizeof(p->box_hi)) == 0 &&
                        c->p.box_spacing == p->box_spacing;
  bf.d = c->bf.d;
  c->bf = bf;
  c->bf_valid = box_same;
  c->vc = vc;
  c->keep_normal = 1.0f - p->normal_friction;
  c->damp = std::pow(2.71828182845904523536f, -p->dt * p->tangent_friction);
  c->g = g;
  c->full_cells = ncell;
  c->pc = pc;
  c->gen++;
  c->pc64 = derive64(p);
  c->m_factor = m_factor;
  c->mass = mass;
  c->p = *p;
  return PBF_OK;
}

// Scene arrays and the compact cell arrays for up to nscenes scenes (each
// scene's window is at most the whole grid, plus one out-of-grid bucket).
int ensure_scenes(pbf_ctx* c, int nscenes) {
  const long long need_cells = (long long)nscenes * (c->full_cells + 1);
  if (need_cells > (long long)INT32_MAX - 2)
    return fail(c, PBF_ERR_UNSUPPORTED, "scenes x grid cells exceed 2^31");
  if (nscenes > c->scene_cap) {
    c->gen++;
    dfree(c->scene_minmax);
    dfree(c->win);
    c->scene_cap = 0;
    PBF_HIP(c, hipMalloc(&c->scene_minmax, sizeof(int) * 6 * (size_t)nscenes));
    PBF_HIP(c, hipMalloc(&c->win, sizeof(SceneWin) * (size_t)nscenes));
    launch_scene_reset(c->scene_minmax, nscenes, nullptr);  // the sort keeps them empty between calls
    PBF_HIP(c, hipGetLastError());
    PBF_HIP(c, hipDeviceSynchronize());
    c->scene_cap = nscenes;
  }
  if (need_cells > c->cell_cap) {
    c->gen++;
    dfree(c->cell_cnt);
    dfree(c->cell_start);
    c->cell_cap = 0;
    PBF_HIP(c, hipMalloc(&c->cell_cnt, sizeof(int) * ((size_t)need_cells + 1)));
    PBF_HIP(c, hipMalloc(&c->cell_start, sizeof(int) * ((size_t)need_cells + 2)));
    PBF_HIP(c, hipMemset(c->cell_cnt, 0, sizeof(int) * ((size_t)need_cells + 1)));  // the scan re-zeroes what it reads
    PBF_HIP(c, hipDeviceSynchronize());
    c->cell_cap = need_cells;
    c->last_n = -1;
  }
  return PBF_OK;
}

int ensure_scan_scratch(pbf_ctx* c, long long n) {
  const size_t need = scan_scratch_ints(n);
  if (need <= c->scan_scratch_n) return PBF_OK;
  c->gen++;
  dfree(c->scan_scratch);
  c->scan_scratch_n = 0;
  PBF_HIP(c, hipMalloc(&c->scan_scratch, sizeof(int) * need));
  c->scan_scratch_n = need;
  return PBF_OK;
}

long long mask_words_for(const pbf_ctx* c, long long n) {
  const long long per = c->p.mask_words > 0 ? c->p.mask_words : 32;
  return (n + 63) / 64 * 64 * per;
}

long long list_slots_for(const pbf_ctx* c, long long n) {
  const long long per = c->p.list_slots > 0 ? c->p.list_slots : c->auto_list_per;
  return (n + 63) / 64 * 64 * per;
}

// 16-bit list slots per particle of every wave's fixed list region.
// Whether a search over n particles stages neighbor lists in LDS
// (stage_min_particles: 0 the library's threshold, < 0 never).
bool stages_lists(const pbf_ctx* c, int n) {
  const int smin = c->p.stage_min_particles == 0 ? kStageMinParticles : c->p.stage_min_particles;
  return smin > 0 && n >= smin;
}

int list_per_of(const pbf_ctx* c) {
  const long long waves = ((long long)c->cap_n + 63) / 64;
  return waves > 0 ? (int)std::min<long long>(c->list_cap / (64 * waves), INT_MAX) : 0;
}

int ensure_mask(pbf_ctx* c, long long words) {
  if (words <= c->mask_cap) return PBF_OK;
  c->gen++;
  dfree(c->mask);
  c->mask_cap = 0;
  PBF_HIP(c, hipMalloc(&c->mask, sizeof(uint32_t) * (size_t)words));
  c->mask_cap = words;
  return PBF_OK;
}

int ensure_list(pbf_ctx* c, long long slots) {
  if (slots <= c->list_cap) return PBF_OK;
  c->gen++;
  dfree(c->list);
  c->list_cap = 0;
  PBF_HIP(c, hipMalloc(&c->list, sizeof(uint16_t) * (size_t)slots));
  c->list_cap = slots;
  return PBF_OK;
}

int ensure_particles(pbf_ctx* c, int n, int nscenes) {
  int st = ensure_scenes(c, nscenes);
  if (st) return st;
  if (n >= kMaxParticles) return fail(c, PBF_ERR_UNSUPPORTED, "more than 2^27 - 1 particles in one call");
  st = ensure_scan_scratch(c, (long long)n + 1);
  if (st) return st;
  if (n < c->cap_n) return PBF_OK;  // index n holds the padding particle
  const int cap = std::max(((n + 1 + 255) / 256) * 256, 256);
  c->gen++;
  dfree(c->key); dfree(c->slot); dfree(c->tmp); dfree(c->perm); dfree(c->inv);
  dfree(c->sid); dfree(c->ssid); dfree(c->bpart); dfree(c->nbr_cnt); dfree(c->cnt_orig);
  dfree(c->tid); dfree(c->pid);
  dfree(c->xn); dfree(c->x0); dfree(c->xa); dfree(c->xb); dfree(c->vs); dfree(c->vv); dfree(c->xc);
  dfree(c->dpos); dfree(c->mask_base); dfree(c->list_base); dfree(c->list_jbase); dfree(c->list_mode);
  dfree(c->stage); dfree(c->cnt_a); dfree(c->wave_stat); dfree(c->oog_rcnt); dfree(c->oog_rev);
  dfree(c->oog_flag); dfree(c->oog_pre);
  c->cap_n = 0;
  const size_t ni = sizeof(int) * (size_t)cap, nf4 = sizeof(float4) * (size_t)cap;
  PBF_HIP(c, hipMalloc(&c->key, ni));
  PBF_HIP(c, hipMalloc(&c->slot, ni));
  PBF_HIP(c, hipMalloc(&c->tmp, ni));
  PBF_HIP(c, hipMalloc(&c->perm, ni));
  PBF_HIP(c, hipMalloc(&c->inv, ni));
  PBF_HIP(c, hipMalloc(&c->sid, ni));
  PBF_HIP(c, hipMalloc(&c->ssid, ni));
  PBF_HIP(c, hipMalloc(&c->tid, ni));
  PBF_HIP(c, hipMalloc(&c->pid, ni));
  PBF_HIP(c, hipMalloc(&c->bpart, sizeof(int) * bounds_part_ints(cap)));
  PBF_HIP(c, hipMalloc(&c->nbr_cnt, ni));
  PBF_HIP(c, hipMalloc(&c->cnt_orig, ni));
  PBF_HIP(c, hipMalloc(&c->xn, nf4 + sizeof(float4) * kXnPad));
  PBF_HIP(c, hipMalloc(&c->xc, 3 * sizeof(float) * ((size_t)cap + kXnPad)));
  PBF_HIP(c, hipMalloc(&c->x0, nf4));
  PBF_HIP(c, hipMalloc(&c->xa, nf4));
  PBF_HIP(c, hipMalloc(&c->xb, nf4));
  PBF_HIP(c, hipMalloc(&c->vs, nf4));
  PBF_HIP(c, hipMalloc(&c->vv, nf4));
  PBF_HIP(c, hipMalloc(&c->dpos, 3 * sizeof(float) * (size_t)cap));
  PBF_HIP(c, hipMemset(c->dpos, 0, 3 * sizeof(float) * (size_t)cap));
  const size_t nwv = (size_t)(cap / 64 + 4);
  PBF_HIP(c, hipMalloc(&c->mask_base, sizeof(long long) * nwv));
  PBF_HIP(c, hipMalloc(&c->list_base, sizeof(long long) * nwv));
  PBF_HIP(c, hipMalloc(&c->list_jbase, sizeof(int) * nwv));
  PBF_HIP(c, hipMalloc(&c->list_mode, sizeof(int) * nwv));
  PBF_HIP(c, hipMalloc(&c->stage, sizeof(int) * kStageInts * nwv));
  PBF_HIP(c, hipMalloc(&c->cnt_a, sizeof(uint16_t) * (size_t)cap));
  PBF_HIP(c, hipMalloc(&c->oog_rcnt, ni));
  PBF_HIP(c, hipMalloc(&c->oog_rev, ni * kOogRev));
  PBF_HIP(c, hipMalloc(&c->oog_flag, ni + sizeof(int)));
  PBF_HIP(c, hipMemset(c->oog_flag, 0, ni + sizeof(int)));
  PBF_HIP(c, hipMalloc(&c->oog_pre, ni + sizeof(int)));
  PBF_HIP(c, hipMalloc(&c->wave_stat, sizeof(int4) * nwv));
  c->cap_n = cap;
  st = ensure_mask(c, mask_words_for(c, cap));
  if (st) return st;
  st = ensure_list(c, list_slots_for(c, cap));
  if (st) return st;
  c->last_n = -1;
  return PBF_OK;
}

// Grows the mask and list reservations when the last finished call wanted
// more; never blocks (the counters are copied asynchronously at the end of
// each call).
int adapt_reservations(pbf_ctx* c) {
  if (!c->counters_pending) return PBF_OK;
  if (hipEventQuery(c->counters_ev) != hipSuccess) return PBF_OK;
  c->counters_pending = false;
  const long long mw = (long long)c->h_counters[C_MASK_WORDS];
  const long long need = (long long)c->h_counters[C_LIST_NEED];  // slots per particle of the fullest wave
  int st = PBF_OK;
  // a reservation given explicitly (> 0 per particle) stays as given
  if (c->p.mask_words == 0 && mw > c->mask_cap) st = ensure_mask(c, mw + mw / 8);
  if (!st && c->p.list_slots == 0 && need > list_per_of(c)) {
    c->auto_list_per = std::max(c->auto_list_per, (need + need / 8 + 3) / 4 * 4);
    st = ensure_list(c, list_slots_for(c, c->cap_n));
  }
  return st;
}

NbrView view_of(const pbf_ctx* c, int n) {
  NbrView nv;
  nv.g = c->g;
  nv.win = c->win;
  nv.ssid = c->ssid;
  nv.cell_start = c->cell_start;
  nv.xn = c->xn;
  nv.cx = c->xc;
  nv.cy = c->xc + (c->cap_n + kXnPad);
  nv.cz = c->xc + 2 * (size_t)(c->cap_n + kXnPad);
  nv.nbr_cnt = c->nbr_cnt;
  nv.list_base = c->list_base;
  nv.list_jbase = c->list_jbase;
  nv.list_mode = c->list_mode;
  nv.stage = c->stage;
  nv.cnt_a = c->cnt_a;
  nv.oog_rcnt = c->oog_rcnt;
  nv.oog_rev = c->oog_rev;
  nv.list = c->list;
  nv.mask_base = c->mask_base;
  nv.mask = c->mask;
  nv.counters = c->counters;
  nv.thr2 = c->pc.thr2;
  nv.n = n;
  nv.pad = n;
  return nv;
}

NbrOut out_of(pbf_ctx* c) {
  NbrOut o;
  o.mask = c->mask;
  o.mask_cap = c->mask_cap;
  o.list = c->list;
  o.list_per = list_per_of(c);
  o.mask_base = c->mask_base;
  o.list_base = c->list_base;
  o.list_jbase = c->list_jbase;
  o.list_mode = c->list_mode;
  o.stage = c->stage;
  o.stage_lists = 0;
  o.oog_rcnt = c->oog_rcnt;
  o.oog_rev = c->oog_rev;
  o.cnt_a = c->cnt_a;
  o.nbr_cnt = c->nbr_cnt;
  o.wave_stat = c->wave_stat;
  return o;
}

SortBufs sort_of(pbf_ctx* c, const int* scene_off, int nscenes) {
  SortBufs sb;
  sb.scene_off = scene_off;
  sb.nscenes = nscenes;
  sb.sid = c->sid;
  sb.ssid = c->ssid;
  sb.scene_minmax = c->scene_minmax;
  sb.win = c->win;
  sb.total_cells = c->total_cells;
  sb.cell_cap = c->cell_cap;
  sb.cell_cnt = c->cell_cnt;
  sb.cell_start = c->cell_start;
  sb.key = c->key;
  sb.slot = c->slot;
  sb.tmp = c->tmp;
  sb.perm = c->perm;
  sb.scan_part = c->scan_part;
  sb.bpart = c->bpart;
  sb.xc = c->xc;
  sb.oog_rcnt = c->oog_rcnt;
  sb.tid = c->tid;
  sb.oog_flag = c->oog_flag;
  sb.oog_pre = c->oog_pre;
  sb.xc_stride = c->cap_n + kXnPad;
  return sb;
}

// HIP-event pair around one pass on the call's stream (profiling only).
hipEvent_t take_event(pbf_ctx* c) {
  if (!c->spare.empty()) {
    hipEvent_t e = c->spare.back();
    c->spare.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int drain_marks(pbf_ctx* c);

// A roctx range (rocprofv3 --marker-trace shows the entry points and passes);
// a push without an attached tool costs a flag test.
struct Marker {
  explicit Marker(const char* name) { roctxRangePush(name); }
  ~Marker() { roctxRangePop(); }
  Marker(const Marker&) = delete;
  Marker& operator=(const Marker&) = delete;
};

const char* pass_name(int pass) {
  switch (pass) {
    case PBF_PASS_SORT: return "pbf sort";
    case PBF_PASS_NEIGHBORS: return "pbf neighbors";
    case PBF_PASS_DENSITY_LAMBDA: return "pbf density+lambda";
    default: return "pbf displacement";
  }
}

// Times one pass while profiling. A pass of several kernels (the sort) is
// bracketed by events recorded on the stream; a one-kernel pass hands its
// events to the launch (events(): they time that dispatch itself).
struct PassTimer {
  Marker mark;
  pbf_ctx* c;
  int pass;
  hipStream_t st;
  bool attach;
  KernelEvents ev{nullptr, nullptr};
  PassTimer(pbf_ctx* c_, int pass_, hipStream_t st_, bool attach_ = false)
      : mark(pass_name(pass_)), c(c_), pass(pass_), st(st_), attach(attach_) {
    if (!c->profiling || (ev.start = take_event(c)) == nullptr) return;
    if (attach && (ev.stop = take_event(c)) == nullptr) {
      c->spare.push_back(ev.start);
      ev.start = nullptr;
      return;
    }
    if (!attach) (void)hipEventRecord(ev.start, st);
  }
  const KernelEvents* events() const { return attach && ev.start ? &ev : nullptr; }
  ~PassTimer() {
    if (!ev.start) return;
    if (!attach) {
      ev.stop = take_event(c);
      if (!ev.stop) {
        c->spare.push_back(ev.start);
        return;
      }
      (void)hipEventRecord(ev.stop, st);
    }
    c->marks.push_back({pass, ev.start, ev.stop});
    // bound the events in flight: sum the finished ones every 512 passes
    if (c->marks.size() >= 512) (void)drain_marks(c);
  }
};

int drain_marks(pbf_ctx* c) {
  for (auto& m : c->marks) {
    PBF_HIP(c, hipEventSynchronize(m.b));
    float ms = 0.f;
    PBF_HIP(c, hipEventElapsedTime(&ms, m.a, m.b));
    c->prof_ms[m.pass] += ms;
    c->prof_n[m.pass] += 1;
    c->spare.push_back(m.a);
    c->spare.push_back(m.b);
  }
  c->marks.clear();
  return PBF_OK;
}

// NeighborQuery::compute on the positions x for a batch of scenes
// (GridHash::construct + queryNeighborDynamic), leaving the cell-sorted state
// and the neighbor lists in the context; with `integrate` the integrator is
// fused into the reorder.
int search(pbf_ctx* c, const SortIn& in, int n, const int* scene_off, int nscenes, bool integrate, hipStream_t st) {
  {
    PassTimer t(c, PBF_PASS_SORT, st);
    launch_sort(sort_of(c, scene_off, nscenes), c->g, in, n, c->xn, c->x0, c->vs, integrate ? &c->pc : nullptr,
                c->counters, st);
  }
  {
    PassTimer t(c, PBF_PASS_NEIGHBORS, st, true);
    NbrOut o = out_of(c);
    o.stage_lists = stages_lists(c, n) ? 1 : 0;
    launch_neighbors(view_of(c, n), o, c->counters, st, t.events());
  }
  c->stats_waves = (n + 63) / 64;
  PBF_LAUNCHED(c);
  c->last_n = n;
  return PBF_OK;
}

int search(pbf_ctx* c, const float* x, const float* v, int n, const int* scene_off, int nscenes, bool integrate,
           hipStream_t st) {
  return search(c, SortIn{x, v, nullptr, nullptr, nullptr}, n, scene_off, nscenes, integrate, st);
}

int record_counters(pbf_ctx* c, hipStream_t st) {
  launch_wave_stats(c->wave_stat, c->stats_waves, c->counters, st);
  c->stats_waves = 0;
  PBF_HIP(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(unsigned long long) * C_COUNT,
                            hipMemcpyDeviceToHost, st));
  PBF_HIP(c, hipEventRecord(c->counters_ev, st));
  c->counters_pending = true;
  return PBF_OK;
}

// CSR of the last search in original ids, in the reference's NeighborList
// layout: d_index[0..n-1] exclusive row starts, *total elements.
int export_csr(pbf_ctx* c, int n, int* d_index, int* d_elems, long long cap, long long* total, hipStream_t st) {
  if (n == 0) {
    *total = 0;
    return PBF_OK;
  }
  launch_invert(c->perm, c->inv, n, st);
  launch_csr_count(c->nbr_cnt, c->inv, n, c->cnt_orig, st);
  // cnt_orig holds cap_n > n entries: the scan writes n + 1 in place
  scan_exclusive(c->cnt_orig, c->cnt_orig, n, c->scan_scratch, st);
  PBF_LAUNCHED(c);
  PBF_HIP(c, hipMemcpyAsync(d_index, c->cnt_orig, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st));
  int tot = 0;
  PBF_HIP(c, hipMemcpyAsync(&tot, c->cnt_orig + n, sizeof(int), hipMemcpyDeviceToHost, st));
  PBF_HIP(c, hipStreamSynchronize(st));
  *total = tot;
  if (d_elems == nullptr) return PBF_OK;
  if (cap < (long long)tot)
    return fail(c, PBF_ERR_CAPACITY, "neighbor element buffer holds " + std::to_string(cap) + ", need " +
                                         std::to_string(tot));
  launch_csr_fill(view_of(c, n), c->perm, c->cnt_orig, d_elems, cap, st);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

// A caller's list (index of n row starts, nelems elements) as the kernels
// read it.
int caller_csr(pbf_ctx* c, const int* d_index, const int* d_elems, long long nelems, int n, Csr* out) {
  if (nelems < 0 || nelems > INT32_MAX) return fail(c, PBF_ERR_INVALID, "nelems must be in [0, 2^31 - 1]");
  if (nelems > 0 && !d_elems) return fail(c, PBF_ERR_INVALID, "null neighbor elements");
  *out = Csr{d_index, d_elems, n, (int)nelems};
  return PBF_OK;
}

}  // namespace

int pbf::ctx_fail(pbf_ctx* c, int status, const std::string& msg) { return fail(c, status, msg); }

namespace {

// Device address of the current device's copy of `scratch`, looked up once per
// device (hipGetSymbolAddress with that device current) and kept.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// DensityPBD::takeOneIteration x K over a caller list (DensityPBD.cu:306-365),
// then DP_UpdateVelocity (:368-388): xa holds the caller's x on entry, dpos
// zeros; minv selects the massInv variant.
void pbd_csr_rounds(const pbf_ctx* c, float4* xa, float* dpos, const Csr& nb, float* d_x, float* d_v,
                    float* d_rho, int n, const float* minv, hipStream_t st) {
  for (int it = 0; it < c->p.iterations; ++it) {
    const bool last = it + 1 == c->p.iterations;
    launch_csr_lambda(xa, nb, last ? d_rho : nullptr, c->pc, st, minv);
    launch_csr_dx_atomic(xa, nb, dpos, c->pc, st, minv);
    launch_csr_apply(xa, dpos, n, st);
  }
  launch_csr_finish(xa, d_x, d_v, n, c->pc.dt, st);
}

// Samples of the box field (K_DistanceFieldToBox, computed on the device once
// per box).
int ensure_box(pbf_ctx* c, hipStream_t st) {
  if (c->bf_valid) return PBF_OK;
  const size_t need = (size_t)c->bf.nx * c->bf.ny * c->bf.nz;
  if (need > c->bf_cap) {
    dfree(c->bf.d);
    c->bf_cap = 0;
    PBF_HIP(c, hipMalloc(&c->bf.d, sizeof(float) * need));
    c->bf_cap = need;
  }
  launch_box_sdf(c->bf, c->p.box_lo, c->p.box_hi, st);
  PBF_LAUNCHED(c);
  c->bf_valid = true;
  return PBF_OK;
}

// Lanes per particle of the PBD passes for a search over n particles: the
// context's setting, or automatically a lane pair where the search keeps
// unstaged lists or covers at most PBF_TWO_LANE_MAX_PARTICLES (measured: both
// waves of a list copying its stage cost more than the pair saves above that),
// one lane otherwise and with persistent launches (no lane-pair form).
int pass_lanes_for(const pbf_ctx* c, int n) {
  if (c->pass_lanes != 0) return c->pass_lanes;
  if (c->pass_persistent) return 1;
  return (!stages_lists(c, n) || n <= PBF_TWO_LANE_MAX_PARTICLES) ? 2 : 1;
}

// One fused step on a batch of scenes (scene_off: device offsets): the state
// (x_in, v_in) advances into (x_out, v_out), which may alias the inputs. With
// a tape (P: K iterate arrays, P[0] = the integrated positions; aux: K arrays)
// every PBD iterate, its lambda and the adjoint's per-iteration sums are kept.
// The box boundary runs first and the viscosity last when their flags are set
// (never with a tape). A plain step (no box, no viscosity, K >= 1) may take its
// state from `from` and leave it in `to` instead of the caller's arrays (the
// cell order of the step, between the steps of one call).
struct CellState {
  const float4* x;  // an iterate buffer
  const float4* v;  // vv
  const int* id;    // pid
};

int step_once(pbf_ctx* c, const float* x_in, const float* v_in, float* x_out, float* v_out, float* d_rho, int n,
              const int* scene_off, int nscenes, float4* const* P, float4* const* aux, hipStream_t st,
              const CellState* from = nullptr, CellState* to = nullptr) {
  const int K = c->p.iterations;
  const bool box = (c->p.flags & PBF_FLAG_BOX) != 0;
  const bool visc = (c->p.flags & PBF_FLAG_VISCOSITY) != 0 && c->p.viscosity_iterations > 0;
  int s;
  if (box) {
    if ((s = ensure_box(c, st))) return s;
    launch_box_constrain(c->bf, x_in, v_in, x_out, v_out, n, scene_off, nscenes, c->keep_normal, c->damp, st);
    x_in = x_out;
    v_in = v_out;
  }
  SortIn in{x_in, v_in, nullptr, nullptr, nullptr};
  if (from) {
    in.x4 = from->x;
    in.v4 = from->v;
    in.id = from->id;
  }
  s = search(c, in, n, scene_off, nscenes, true, st);
  if (s) return s;
  NbrView nv = view_of(c, n);
  const int lanes = pass_lanes_for(c, n);
  c->last_lanes = lanes;
  float4* bufs[2] = {c->xa, c->xb};
  if (K == 0 && !visc) launch_write_back(c->x0, c->vs, c->perm, n, x_out, v_out, st);
  for (int it = 0; it < K; ++it) {
    const bool last = it + 1 == K;
    float4* src = P ? P[it] : (it == 0 ? c->x0 : bufs[(it - 1) & 1]);
    float4* dst = P ? (last ? nullptr : P[it + 1]) : bufs[it & 1];
    {
      PassTimer t(c, PBF_PASS_DENSITY_LAMBDA, st, true);
      launch_pass_density_lambda(nv, src, c->pc, last ? d_rho : nullptr, c->perm, st, aux ? aux[it] : nullptr,
                                 c->pass_block, c->pass_persistent, lanes, t.events());
    }
    {
      PassTimer t(c, PBF_PASS_DISPLACEMENT, st, true);
      if (last && to) {
        launch_pass_displacement(nv, src, dst, c->pc, true, c->x0, c->vs, c->perm, nullptr, nullptr, st, c->vv,
                                 c->pid, c->pass_block, c->pass_persistent, lanes, t.events());
        *to = CellState{dst, c->vv, c->pid};
      } else {
        launch_pass_displacement(nv, src, dst, c->pc, last, c->x0, c->vs, c->perm, x_out, v_out, st,
                                 (last && visc) ? c->vv : nullptr, nullptr, c->pass_block, c->pass_persistent,
                                 lanes, t.events());
      }
    }
  }
  if (visc) {
    // ImplicitViscosity::constrain on the constrained positions and velocities
    const float4* xk = K == 0 ? c->x0 : bufs[(K - 1) & 1];
    const float4* v_old = K == 0 ? c->vs : c->vv;
    float4* V[2] = {K == 0 ? c->xa : bufs[K & 1], K == 0 ? c->xb : c->x0};
    const int T = c->p.viscosity_iterations;
    for (int t = 0; t < T; ++t) {
      const bool last = t + 1 == T;
      const float4* vin = t == 0 ? v_old : V[(t - 1) & 1];
      launch_viscosity(nv, xk, v_old, vin, last ? nullptr : V[t & 1], c->vc, c->perm, last ? x_out : nullptr,
                       last ? v_out : nullptr, st, lanes);
    }
  }
  PBF_LAUNCHED(c);
  return PBF_OK;
}

// nsteps fused steps in place; d_rho receives the last step's density.
// Plain steps keep the state in cell order between the steps (the next sort
// reads nearly in order); the last step writes the caller's arrays.
// A step between the first and the last of a plain multi-step call reads and
// writes only the context's cell-ordered state (the caller's arrays are not
// touched), so its launch sequence is the same every time: it is captured
// once into a hipGraph (kernel nodes only) and replayed on the call's stream.
int replay_middle_step(pbf_ctx* c, int n, const int* scene_off, int nscenes, CellState& cs, hipStream_t st) {
  auto& g = c->sg;
  if (!g.exec || g.gen != c->gen || g.n != n || g.nscenes != nscenes || g.off != scene_off ||
      g.iters != c->p.iterations) {
    if (g.exec) {
      (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
    }
    PBF_HIP(c, hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeRelaxed));
    CellState from = cs, to{};
    const int s = step_once(c, nullptr, nullptr, nullptr, nullptr, nullptr, n, scene_off, nscenes, nullptr, nullptr,
                            c->cap_stream, &from, &to);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->cap_stream, &graph);
    if (s || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (s) return s;
      return fail(c, PBF_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    if (to.x != cs.x || to.v != cs.v || to.id != cs.id) {
      (void)hipGraphDestroy(graph);
      return fail(c, PBF_ERR_HIP, "captured step left its state in other buffers");
    }
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      g.exec = nullptr;
      return fail(c, PBF_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    g.gen = c->gen;
    g.n = n;
    g.nscenes = nscenes;
    g.off = scene_off;
    g.iters = c->p.iterations;
    c->graph_captures++;
  }
  PBF_HIP(c, hipGraphLaunch(g.exec, st));
  c->graph_replays++;
  return PBF_OK;
}

int run_steps(pbf_ctx* c, float* d_x, float* d_v, float* d_rho, int n, const int* scene_off, int nscenes, int nsteps,
              hipStream_t st) {
  const bool keep = nsteps > 1 && c->p.iterations > 0 && (c->p.flags & (PBF_FLAG_BOX | PBF_FLAG_VISCOSITY)) == 0;
  // the per-pass timers record events around each pass: no replay then
  const bool graph = keep && nsteps > 2 && !c->profiling && (c->p.flags & PBF_FLAG_EAGER) == 0;
  CellState cs{};
  for (int step = 0; step < nsteps; ++step) {
    const bool last = step + 1 == nsteps;
    int s;
    if (graph && step > 0 && !last)
      s = replay_middle_step(c, n, scene_off, nscenes, cs, st);
    else
      s = step_once(c, d_x, d_v, d_x, d_v, last ? d_rho : nullptr, n, scene_off, nscenes, nullptr, nullptr, st,
                    keep && step > 0 ? &cs : nullptr, keep && !last ? &cs : nullptr);
    if (s) return s;
  }
  return record_counters(c, st);
}

template <class T> int ensure_buf(pbf_ctx* c, T*& p, size_t& cap, size_t need) {
  if (need <= cap) return PBF_OK;
  dfree(p);
  cap = 0;
  PBF_HIP(c, hipMalloc(&p, sizeof(T) * need));
  cap = need;
  return PBF_OK;
}

void free_slot(pbf_ctx::TapeSlot& t) {
  dfree(t.it64);
  t.it64_n = 0;
  dfree(t.perm); dfree(t.ssid); dfree(t.nbr_cnt); dfree(t.xn); dfree(t.x0);
  dfree(t.list_base); dfree(t.mask_base); dfree(t.list_jbase); dfree(t.list_mode); dfree(t.stage); dfree(t.cnt_a);
  dfree(t.list); dfree(t.mask); dfree(t.cell_start); dfree(t.win); dfree(t.counters); dfree(t.it);
}

void free_slots(pbf_ctx* c) {
  for (auto& t : c->slots) free_slot(t);
  c->slots.clear();
}

// m tape slots and the adjoint arrays for the current capacities and
// iterations.
int ensure_tape(pbf_ctx* c, int m) {
  auto& r = c->roll;
  const int K = c->p.iterations;
  const size_t cap = (size_t)c->cap_n;
  if (c->slot_n != cap || c->slot_list != (size_t)c->list_cap || c->slot_mask != (size_t)c->mask_cap ||
      c->slot_cells != c->cell_cap || c->slot_scenes != c->scene_cap || c->slot_K != K) {
    free_slots(c);
    c->slot_n = cap;
    c->slot_list = (size_t)c->list_cap;
    c->slot_mask = (size_t)c->mask_cap;
    c->slot_cells = c->cell_cap;
    c->slot_scenes = c->scene_cap;
    c->slot_K = K;
  }
  const size_t nwv = cap / 64 + 4;
  while ((int)c->slots.size() < m) {
    c->slots.emplace_back();
    auto& t = c->slots.back();
    PBF_HIP(c, hipMalloc(&t.perm, sizeof(int) * cap));
    PBF_HIP(c, hipMalloc(&t.ssid, sizeof(int) * cap));
    PBF_HIP(c, hipMalloc(&t.nbr_cnt, sizeof(int) * cap));
    PBF_HIP(c, hipMalloc(&t.xn, sizeof(float4) * (cap + kXnPad)));
    PBF_HIP(c, hipMalloc(&t.x0, sizeof(float4) * cap));
    PBF_HIP(c, hipMalloc(&t.list_base, sizeof(long long) * nwv));
    PBF_HIP(c, hipMalloc(&t.mask_base, sizeof(long long) * nwv));
    PBF_HIP(c, hipMalloc(&t.list_jbase, sizeof(int) * nwv));
    PBF_HIP(c, hipMalloc(&t.list_mode, sizeof(int) * nwv));
    PBF_HIP(c, hipMalloc(&t.stage, sizeof(int) * kStageInts * nwv));
    PBF_HIP(c, hipMalloc(&t.cnt_a, sizeof(uint16_t) * cap));
    PBF_HIP(c, hipMalloc(&t.list, sizeof(uint16_t) * std::max<size_t>(c->slot_list, 1)));
    PBF_HIP(c, hipMalloc(&t.mask, sizeof(uint32_t) * std::max<size_t>(c->slot_mask, 1)));
    PBF_HIP(c, hipMalloc(&t.cell_start, sizeof(int) * ((size_t)c->cell_cap + 2)));
    PBF_HIP(c, hipMalloc(&t.win, sizeof(SceneWin) * (size_t)std::max(c->scene_cap, 1)));
    PBF_HIP(c, hipMalloc(&t.counters, sizeof(unsigned long long) * C_COUNT));
    PBF_HIP(c, hipMalloc(&t.it, sizeof(float4) * cap * (size_t)std::max(2 * K - 1, 1)));
  }
  if (r.adj_n < c->cap_n) {
    dfree(r.adj);
    dfree(r.out);
    r.adj_n = r.out_n = 0;
    PBF_HIP(c, hipMalloc(&r.adj, sizeof(float4) * 4 * cap));
    PBF_HIP(c, hipMalloc(&r.out, sizeof(float) * 6 * cap));
    r.adj_n = r.out_n = c->cap_n;
  }
  if (!r.flag) PBF_HIP(c, hipMalloc(&r.flag, sizeof(int)));
  return PBF_OK;
}

// Exchanges the context's search outputs with a slot's arrays: swapped in
// before a taped step, the step writes into the slot; swapped back after it.
void swap_slot(pbf_ctx* c, pbf_ctx::TapeSlot& t) {
  c->gen++;
  std::swap(c->perm, t.perm);
  std::swap(c->ssid, t.ssid);
  std::swap(c->nbr_cnt, t.nbr_cnt);
  std::swap(c->xn, t.xn);
  std::swap(c->x0, t.x0);
  std::swap(c->list_base, t.list_base);
  std::swap(c->mask_base, t.mask_base);
  std::swap(c->list_jbase, t.list_jbase);
  std::swap(c->list_mode, t.list_mode);
  std::swap(c->stage, t.stage);
  std::swap(c->cnt_a, t.cnt_a);
  std::swap(c->list, t.list);
  std::swap(c->mask, t.mask);
}

NbrView slot_view(const pbf_ctx* c, const pbf_ctx::TapeSlot& t, int n) {
  NbrView nv = view_of(c, n);
  nv.win = t.win;
  nv.ssid = t.ssid;
  nv.cell_start = t.cell_start;
  nv.xn = t.xn;
  nv.nbr_cnt = t.nbr_cnt;
  nv.list_base = t.list_base;
  nv.list_jbase = t.list_jbase;
  nv.list_mode = t.list_mode;
  nv.stage = t.stage;
  nv.cnt_a = t.cnt_a;
  nv.list = t.list;
  nv.mask_base = t.mask_base;
  nv.mask = t.mask;
  nv.counters = t.counters;
  return nv;
}

bool bad_n(int n) { return n < 0; }

// Buffers of the double-precision rollout for the current capacity.
int ensure_roll64(pbf_ctx* c, int nscenes) {
  auto& r = c->roll64;
  const size_t cap = (size_t)c->cap_n;
  int s;
  if ((s = ensure_buf(c, r.xf, r.xf_cap, 3 * cap))) return s;
  if ((s = ensure_buf(c, r.w4, r.w4_cap, 4 * cap))) return s;
  if (nscenes + 1 > r.off_cap) {
    dfree(r.off);
    r.off_cap = 0;
    PBF_HIP(c, hipMalloc(&r.off, sizeof(int) * (size_t)(nscenes + 1)));
    r.off_cap = nscenes + 1;
  }
  if (!r.flag) PBF_HIP(c, hipMalloc(&r.flag, sizeof(int)));
  return PBF_OK;
}

// One fused step in double (x_in, v_in -> x_out, v_out; may alias): the
// neighbor search on the positions rounded to float, then the integrator and
// the PBD iterations in double. With a tape, P and aux hold the K iterates
// and the K per-iteration sums.
int step64_once(pbf_ctx* c, const double* x_in, const double* v_in, double* x_out, double* v_out, double* rho,
                int n, const int* off, int nscenes, double4* const* P, double4* const* aux, hipStream_t st) {
  auto& r = c->roll64;
  const int K = c->p.iterations;
  launch64_round(x_in, r.xf, n, st);
  int s = search(c, SortIn{r.xf, nullptr, nullptr, nullptr, nullptr}, n, off, nscenes, false, st);
  if (s) return s;
  const size_t cap = (size_t)c->cap_n;
  double4* vs = r.w4 + 3 * cap;
  std::vector<double4*> Pu;
  if (!P) {  // untaped: P_0, then two ping-pong iterates
    Pu.resize(std::max(K, 1));
    Pu[0] = r.w4;
    for (int k = 1; k < K; ++k) Pu[k] = r.w4 + (size_t)(1 + ((k - 1) & 1)) * cap;
    P = Pu.data();
  }
  c->last_lanes = pass_lanes_for(c, n);
  launch64_step(view_of(c, n), c->perm, x_in, v_in, x_out, v_out, rho, P, aux, K, vs, c->pc64, c->counters, r.flag,
                st, c->last_lanes);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int read_flag64(pbf_ctx* c, hipStream_t st, const char* what) {
  int flag = 0;
  PBF_HIP(c, hipMemcpyAsync(&flag, c->roll64.flag, sizeof(int), hipMemcpyDeviceToHost, st));
  PBF_HIP(c, hipStreamSynchronize(st));
  if (flag) return fail(c, PBF_ERR_UNSUPPORTED, std::string(what) + ": a step had particles outside the grid");
  return PBF_OK;
}

}  // namespace

extern "C" {

int pbf_abi_version(void) { return PBF_ABI_VERSION; }

const char* pbf_status_string(int s) {
  switch (s) {
    case PBF_OK: return "ok";
    case PBF_ERR_INVALID: return "invalid argument";
    case PBF_ERR_HIP: return "HIP runtime error";
    case PBF_ERR_CAPACITY: return "buffer capacity exceeded";
    case PBF_ERR_ALLOC: return "device allocation failed";
    case PBF_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
  }
}

void pbf_params_default(pbf_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->smoothing_length = 0.0125f;
  p->sampling_distance = 0.005f;
  p->rest_density = 1000.0f;
  p->dt = 0.001f;
  p->gravity[0] = 0.0f;
  p->gravity[1] = -9.8f;
  p->gravity[2] = 0.0f;
  p->iterations = 5;
  p->lambda_rest = 1000.0f;
  p->lambda_eps = 0.1f;
  p->corr_scale = 10.0f;
  p->pair_eps = 1e-6f;
  for (int a = 0; a < 3; ++a) {
    p->scene_lo[a] = 0.0f;
    p->scene_hi[a] = 1.0f;
  }
  p->mask_words = 0;
  p->list_slots = 0;
  p->viscosity = 1.0f;
  p->viscosity_iterations = 5;
  for (int a = 0; a < 3; ++a) {
    p->box_lo[a] = 0.0f;
    p->box_hi[a] = 1.0f;
  }
  p->box_spacing = 0.005f;
  p->normal_friction = 0.95f;
  p->tangent_friction = 0.0f;
  p->flags = 0;
  p->stage_min_particles = 0;
}

int pbf_create(const pbf_params* p, pbf_ctx** out) {
  if (!out) return PBF_ERR_INVALID;
  *out = nullptr;
  pbf_params def;
  if (!p) {
    pbf_params_default(&def);
    p = &def;
  }
  pbf_ctx* c = new (std::nothrow) pbf_ctx();
  if (!c) return PBF_ERR_ALLOC;
  std::string why;
  int st = validate(p, why);
  if (st) {
    delete c;
    return st;
  }
  st = derive(c, p);
  if (st) {
    delete c;
    return st;
  }
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipMalloc(&c->counters, sizeof(unsigned long long) * C_COUNT) != hipSuccess ||
      hipHostMalloc(&c->h_counters, sizeof(unsigned long long) * C_COUNT) != hipSuccess ||
      hipEventCreateWithFlags(&c->counters_ev, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->off1, sizeof(int) * 2) != hipSuccess || hipMalloc(&c->total_cells, sizeof(int)) != hipSuccess ||
      hipMalloc(&c->scan_part, sizeof(int) * (kScanDevBlocks + 1)) != hipSuccess) {
    pbf_destroy(c);
    return PBF_ERR_HIP;
  }
  std::memset(c->h_counters, 0, sizeof(unsigned long long) * C_COUNT);
  (void)hipMemset(c->counters, 0, sizeof(unsigned long long) * C_COUNT);
  {
    const hipError_t e = scratch_address(c->device, &c->noalloc_scratch);
    if (e != hipSuccess)
      c->noalloc_err = std::string("scratch lookup on device ") + std::to_string(c->device) + ": " + hipGetErrorString(e);
  }
  st = ensure_scenes(c, 1);
  if (st) {
    pbf_destroy(c);
    return st;
  }
  *out = c;
  return PBF_OK;
}

int pbf_destroy(pbf_ctx* c) {
  if (!c) return PBF_ERR_INVALID;
  (void)hipDeviceSynchronize();
  dfree(c->off1); dfree(c->scene_minmax); dfree(c->win); dfree(c->total_cells); dfree(c->scan_part);
  dfree(c->cell_cnt); dfree(c->cell_start);
  dfree(c->key); dfree(c->slot); dfree(c->tmp); dfree(c->perm); dfree(c->inv);
  dfree(c->sid); dfree(c->ssid); dfree(c->bpart); dfree(c->nbr_cnt); dfree(c->cnt_orig);
  dfree(c->tid); dfree(c->pid);
  dfree(c->xn); dfree(c->x0); dfree(c->xa); dfree(c->xb); dfree(c->vs); dfree(c->vv); dfree(c->xc);
  dfree(c->dpos); dfree(c->mask_base); dfree(c->list_base); dfree(c->list_jbase); dfree(c->list_mode);
  dfree(c->stage); dfree(c->cnt_a); dfree(c->wave_stat); dfree(c->oog_rcnt); dfree(c->oog_rev);
  dfree(c->oog_flag); dfree(c->oog_pre);
  dfree(c->mask); dfree(c->list); dfree(c->scan_scratch); dfree(c->bf.d);
  free_slots(c);
  dfree(c->roll.off); dfree(c->roll.ckpt); dfree(c->roll.seg); dfree(c->roll.adj);
  dfree(c->roll.out); dfree(c->roll.flag);
  dfree(c->roll64.off); dfree(c->roll64.ckpt); dfree(c->roll64.seg); dfree(c->roll64.xf); dfree(c->roll64.w4);
  dfree(c->roll64.adj); dfree(c->roll64.out); dfree(c->roll64.flag);
  dfree(c->counters);
  if (c->h_counters) (void)hipHostFree(c->h_counters);
  if (c->counters_ev) (void)hipEventDestroy(c->counters_ev);
  if (c->sg.exec) (void)hipGraphExecDestroy(c->sg.exec);
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  for (auto& m : c->marks) {
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  for (hipEvent_t e : c->spare) (void)hipEventDestroy(e);
  delete c;
  return PBF_OK;
}

int pbf_set_params(pbf_ctx* c, const pbf_params* p) {
  if (!c || !p) return fail(c, PBF_ERR_INVALID, "null argument");
  std::string why;
  int st = validate(p, why);
  if (st) return fail(c, st, why);
  PBF_HIP(c, hipDeviceSynchronize());
  const pbf_params old = c->p;
  st = derive(c, p);
  if (st) return st;
  st = ensure_scenes(c, std::max(c->scene_cap, 1));
  if (st) {
    derive(c, &old);
    return st;
  }
  c->last_n = -1;
  if (c->cap_n > 0) {
    st = ensure_mask(c, mask_words_for(c, c->cap_n));
    if (st) return st;
    return ensure_list(c, list_slots_for(c, c->cap_n));
  }
  return PBF_OK;
}

int pbf_get_params(const pbf_ctx* c, pbf_params* out) {
  if (!c || !out) return PBF_ERR_INVALID;
  *out = c->p;
  return PBF_OK;
}

int pbf_reserve(pbf_ctx* c, int n) {
  if (!c || bad_n(n)) return fail(c, PBF_ERR_INVALID, "bad particle count");
  return ensure_particles(c, n, 1);
}

int pbf_reserve_batch(pbf_ctx* c, int n, int nscenes) {
  if (!c || bad_n(n) || nscenes < 1) return fail(c, PBF_ERR_INVALID, "bad particle or scene count");
  return ensure_particles(c, n, nscenes);
}

int pbf_neighbor_query(pbf_ctx* c, const float* d_x, int n, int* d_index, int* d_elems, long long cap,
                       long long* total, void* stream) {
  Marker mark("pbf_neighbor_query");
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || !total || (n > 0 && (!d_x || !d_index)) || cap < 0) return fail(c, PBF_ERR_INVALID, "bad argument");
  hipStream_t st = (hipStream_t)stream;
  int s = ensure_particles(c, n, 1);
  if (s) return s;
  if (n == 0) {
    c->last_n = 0;
    return export_csr(c, 0, d_index, d_elems, cap, total, st);
  }
  s = adapt_reservations(c);
  if (s) return s;
  launch_set_offsets(c->off1, n, st);
  s = search(c, d_x, nullptr, n, c->off1, 1, false, st);
  if (s) return s;
  s = record_counters(c, st);
  if (s) return s;
  return export_csr(c, n, d_index, d_elems, cap, total, st);
}

int pbf_step_neighbors(pbf_ctx* c, int* d_index, int* d_elems, long long cap, long long* total, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (!total || cap < 0) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (c->last_n < 0) return fail(c, PBF_ERR_INVALID, "no neighbor search has run on this context");
  if (c->last_n > 0 && !d_index) return fail(c, PBF_ERR_INVALID, "null neighbor index");
  return export_csr(c, c->last_n, d_index, d_elems, cap, total, (hipStream_t)stream);
}

int pbf_integrate(pbf_ctx* c, float* d_x, float* d_v, const float* d_f, int n, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_v))) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n == 0) return PBF_OK;
  launch_integrate(d_x, d_v, d_f, n, c->pc, (hipStream_t)stream);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int pbf_density(pbf_ctx* c, const float* d_x, const int* d_index, const int* d_elems, long long nelems, float* d_rho,
                int n, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_index || !d_rho))) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n == 0) return PBF_OK;
  Csr nb;
  int s = caller_csr(c, d_index, d_elems, nelems, n, &nb);
  if (s) return s;
  launch_csr_density(d_x, nb, d_rho, c->pc, (hipStream_t)stream);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int pbf_pbd_constrain(pbf_ctx* c, float* d_x, float* d_v, const int* d_index, const int* d_elems, long long nelems,
                      float* d_rho, int n, void* stream) {
  Marker mark("pbf_pbd_constrain");
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_v || !d_index))) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n == 0) return PBF_OK;
  Csr nb;
  int s = caller_csr(c, d_index, d_elems, nelems, n, &nb);
  if (s) return s;
  s = ensure_particles(c, n, 1);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  launch_aos_to_f4(d_x, c->xa, n, st);
  pbd_csr_rounds(c, c->xa, c->dpos, nb, d_x, d_v, d_rho, n, nullptr, st);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int pbf_pbd_constrain_noalloc(pbf_ctx* c, float* d_x, float* d_v, const int* d_index, const int* d_elems,
                              long long nelems, float* d_rho, int n, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_v || !d_index))) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n > PBF_NOALLOC_MAX_PARTICLES)
    return fail(c, PBF_ERR_UNSUPPORTED,
                "pbf_pbd_constrain_noalloc: " + std::to_string(n) + " particles exceed its fixed scratch of " +
                    std::to_string(PBF_NOALLOC_MAX_PARTICLES) + " (use pbf_pbd_constrain)");
  if (n == 0) return PBF_OK;
  if (!c->noalloc_scratch) return fail(c, PBF_ERR_HIP, "pbf_pbd_constrain_noalloc: " + c->noalloc_err);
  Csr nb;
  int s = caller_csr(c, d_index, d_elems, nelems, n, &nb);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  float4* xa = static_cast<float4*>(c->noalloc_scratch);
  float* dpos = reinterpret_cast<float*>(xa + PBF_NOALLOC_MAX_PARTICLES);
  launch_aos_to_f4_clear(d_x, xa, dpos, n, st);
  pbd_csr_rounds(c, xa, dpos, nb, d_x, d_v, d_rho, n, nullptr, st);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int pbf_step(pbf_ctx* c, float* d_x, float* d_v, float* d_rho, int n, int nsteps, void* stream) {
  Marker mark("pbf_step");
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || nsteps < 0 || (n > 0 && (!d_x || !d_v))) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n == 0 || nsteps == 0) return PBF_OK;
  hipStream_t st = (hipStream_t)stream;
  int s = ensure_particles(c, n, 1);
  if (s) return s;
  s = adapt_reservations(c);
  if (s) return s;
  launch_set_offsets(c->off1, n, st);
  return run_steps(c, d_x, d_v, d_rho, n, c->off1, 1, nsteps, st);
}

int pbf_batch_step(pbf_ctx* c, float* d_x, float* d_v, float* d_rho, const int* d_scene_offsets, int nscenes, int n,
                   int nsteps, void* stream) {
  Marker mark("pbf_batch_step");
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || nsteps < 0 || nscenes < 1 || !d_scene_offsets || (n > 0 && (!d_x || !d_v)))
    return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n == 0 || nsteps == 0) return PBF_OK;
  hipStream_t st = (hipStream_t)stream;
  int s = ensure_particles(c, n, nscenes);
  if (s) return s;
  s = adapt_reservations(c);
  if (s) return s;
  return run_steps(c, d_x, d_v, d_rho, n, d_scene_offsets, nscenes, nsteps, st);
}

int pbf_rollout_forward(pbf_ctx* c, float* d_x, float* d_v, float* d_rho, const int* d_scene_offsets, int nscenes,
                        int n, int nsteps, int checkpoint_every, void* stream) {
  Marker mark("pbf_rollout_forward");
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || n == 0 || nsteps < 1 || checkpoint_every < 1 || nscenes < 1 || !d_x || !d_v ||
      (!d_scene_offsets && nscenes != 1))
    return fail(c, PBF_ERR_INVALID, "bad argument");
  if (c->p.flags & (PBF_FLAG_VISCOSITY | PBF_FLAG_BOX))
    return fail(c, PBF_ERR_UNSUPPORTED, "reverse mode covers the plain PBD step (no viscosity, no box)");
  hipStream_t st = (hipStream_t)stream;
  auto& r = c->roll;
  r.valid = false;
  int s = ensure_particles(c, n, nscenes);
  if (s) return s;
  s = adapt_reservations(c);
  if (s) return s;
  const size_t state = 6 * (size_t)n;
  const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
  if ((s = ensure_buf(c, r.ckpt, r.ckpt_cap, (size_t)nck * state))) return s;
  if ((s = ensure_buf(c, r.seg, r.seg_cap, 2 * state))) return s;
  if (nscenes + 1 > r.off_cap) {
    dfree(r.off);
    r.off_cap = 0;
    PBF_HIP(c, hipMalloc(&r.off, sizeof(int) * (size_t)(nscenes + 1)));
    r.off_cap = nscenes + 1;
  }
  if (d_scene_offsets)
    PBF_HIP(c, hipMemcpyAsync(r.off, d_scene_offsets, sizeof(int) * (size_t)(nscenes + 1), hipMemcpyDeviceToDevice,
                              st));
  else
    launch_set_offsets(r.off, n, st);
  for (int k = 0; k < nsteps; ++k) {
    if (k % checkpoint_every == 0) {
      float* ck = r.ckpt + (size_t)(k / checkpoint_every) * state;
      PBF_HIP(c, hipMemcpyAsync(ck, d_x, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st));
      PBF_HIP(c, hipMemcpyAsync(ck + 3 * (size_t)n, d_v, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice,
                                st));
    }
    s = step_once(c, d_x, d_v, d_x, d_v, k + 1 == nsteps ? d_rho : nullptr, n, r.off, nscenes, nullptr, nullptr, st);
    if (s) return s;
  }
  s = record_counters(c, st);
  if (s) return s;
  r.n = n;
  r.nscenes = nscenes;
  r.nsteps = nsteps;
  r.every = checkpoint_every;
  r.iters = c->p.iterations;
  r.p = c->p;
  r.valid = true;
  return PBF_OK;
}

int pbf_rollout_backward(pbf_ctx* c, float* d_xbar, float* d_vbar, void* stream) {
  Marker mark("pbf_rollout_backward");
  if (!c) return PBF_ERR_INVALID;
  if (!d_xbar || !d_vbar) return fail(c, PBF_ERR_INVALID, "bad argument");
  auto& r = c->roll;
  if (!r.valid) return fail(c, PBF_ERR_INVALID, "no rollout has run on this context");
  if (std::memcmp(&r.p, &c->p, sizeof(pbf_params)) != 0)
    return fail(c, PBF_ERR_INVALID, "parameters changed since the rollout");
  hipStream_t st = (hipStream_t)stream;
  const int n = r.n, K = r.iters;
  int s = ensure_particles(c, n, r.nscenes);
  if (s) return s;
  if ((s = adapt_reservations(c))) return s;
  if ((s = ensure_tape(c, std::min(r.every, r.nsteps)))) return s;
  const size_t cap = (size_t)c->cap_n, state = 6 * (size_t)n;
  AdjBufs ab;
  ab.e = c->xa;
  ab.e2 = c->xb;
  ab.f = c->vv;
  ab.q1 = r.adj;
  ab.q2 = r.adj + cap;
  ab.c = r.adj + 2 * cap;
  ab.vsb = r.adj + 3 * cap;
  ab.flag = r.flag;
  PBF_HIP(c, hipMemsetAsync(r.flag, 0, sizeof(int), st));
  std::vector<float4*> P(std::max(K, 1)), aux(std::max(K, 1));
  auto slot_arrays = [&](pbf_ctx::TapeSlot& t, float4* x0) {
    P[0] = x0;
    for (int k = 1; k < K; ++k) P[k] = t.it + (size_t)(k - 1) * cap;
    for (int k = 0; k < K; ++k) aux[k] = t.it + (size_t)(K - 1 + k) * cap;
  };
  float* sa = r.seg;
  float* sb = r.seg + state;
  const int nck = (r.nsteps + r.every - 1) / r.every;
  for (int ck = nck - 1; ck >= 0; --ck) {
    const int s0 = ck * r.every, m = std::min(s0 + r.every, r.nsteps) - s0;
    // the segment from its checkpoint, every step taped into its slot
    PBF_HIP(c, hipMemcpyAsync(sa, r.ckpt + (size_t)ck * state, sizeof(float) * state, hipMemcpyDeviceToDevice, st));
    for (int q = 0; q < m; ++q) {
      const float* in = (q & 1) ? sb : sa;
      float* xout = q + 1 == m ? r.out : ((q & 1) ? sa : sb);
      float* vout = q + 1 == m ? r.out + 3 * cap : xout + 3 * (size_t)n;
      auto& t = c->slots[q];
      swap_slot(c, t);
      slot_arrays(t, c->x0);
      s = step_once(c, in, in + 3 * (size_t)n, xout, vout, nullptr, n, r.off, r.nscenes, P.data(), aux.data(), st);
      if (!s) {
        launch_copy_cells(c->cell_start, t.cell_start, c->total_cells, c->cell_cap, st);
        if (hipMemcpyAsync(t.win, c->win, sizeof(SceneWin) * (size_t)r.nscenes, hipMemcpyDeviceToDevice, st) !=
                hipSuccess ||
            hipMemcpyAsync(t.counters, c->counters, sizeof(unsigned long long) * C_COUNT, hipMemcpyDeviceToDevice,
                           st) != hipSuccess)
          s = fail(c, PBF_ERR_HIP, "tape copy failed");
      }
      swap_slot(c, t);
      if (s) return s;
    }
    // reverse over the segment
    for (int q = m - 1; q >= 0; --q) {
      auto& t = c->slots[q];
      slot_arrays(t, t.x0);
      launch_adjoint_step(slot_view(c, t, n), P.data(), aux.data(), K, c->pc, t.perm, d_xbar, d_vbar, ab, st,
                          pass_lanes_for(c, n));
      PBF_LAUNCHED(c);
    }
  }
  // the context's own search arrays no longer describe one search
  c->last_n = -1;
  s = record_counters(c, st);
  if (s) return s;
  int flag = 0;
  PBF_HIP(c, hipMemcpyAsync(&flag, r.flag, sizeof(int), hipMemcpyDeviceToHost, st));
  PBF_HIP(c, hipStreamSynchronize(st));
  if (flag)
    return fail(c, PBF_ERR_UNSUPPORTED,
                "a step of the rollout had particles outside the grid; its adjoint is not implemented");
  return PBF_OK;
}

int pbf_rollout_forward_f64(pbf_ctx* c, double* d_x, double* d_v, double* d_rho, const int* d_scene_offsets,
                            int nscenes, int n, int nsteps, int checkpoint_every, void* stream) {
  Marker mark("pbf_rollout_forward_f64");
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || n == 0 || nsteps < 1 || checkpoint_every < 1 || nscenes < 1 || !d_x || !d_v ||
      (!d_scene_offsets && nscenes != 1))
    return fail(c, PBF_ERR_INVALID, "bad argument");
  if (c->p.flags & (PBF_FLAG_VISCOSITY | PBF_FLAG_BOX))
    return fail(c, PBF_ERR_UNSUPPORTED, "reverse mode covers the plain PBD step (no viscosity, no box)");
  hipStream_t st = (hipStream_t)stream;
  auto& r = c->roll64;
  r.valid = false;
  int s = ensure_particles(c, n, nscenes);
  if (s) return s;
  if ((s = adapt_reservations(c))) return s;
  if ((s = ensure_roll64(c, nscenes))) return s;
  const size_t state = 6 * (size_t)n;
  const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
  if ((s = ensure_buf(c, r.ckpt, r.ckpt_cap, (size_t)nck * state))) return s;
  if (d_scene_offsets)
    PBF_HIP(c, hipMemcpyAsync(r.off, d_scene_offsets, sizeof(int) * (size_t)(nscenes + 1), hipMemcpyDeviceToDevice,
                              st));
  else
    launch_set_offsets(r.off, n, st);
  PBF_HIP(c, hipMemsetAsync(r.flag, 0, sizeof(int), st));
  for (int k = 0; k < nsteps; ++k) {
    if (k % checkpoint_every == 0) {
      double* ck = r.ckpt + (size_t)(k / checkpoint_every) * state;
      PBF_HIP(c, hipMemcpyAsync(ck, d_x, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st));
      PBF_HIP(c, hipMemcpyAsync(ck + 3 * (size_t)n, d_v, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToDevice,
                                st));
    }
    s = step64_once(c, d_x, d_v, d_x, d_v, k + 1 == nsteps ? d_rho : nullptr, n, r.off, nscenes, nullptr, nullptr,
                    st);
    if (s) return s;
  }
  if ((s = record_counters(c, st))) return s;
  if ((s = read_flag64(c, st, "pbf_rollout_forward_f64"))) return s;
  r.n = n;
  r.nscenes = nscenes;
  r.nsteps = nsteps;
  r.every = checkpoint_every;
  r.iters = c->p.iterations;
  r.p = c->p;
  r.valid = true;
  return PBF_OK;
}

int pbf_rollout_backward_f64(pbf_ctx* c, double* d_xbar, double* d_vbar, void* stream) {
  Marker mark("pbf_rollout_backward_f64");
  if (!c) return PBF_ERR_INVALID;
  if (!d_xbar || !d_vbar) return fail(c, PBF_ERR_INVALID, "bad argument");
  auto& r = c->roll64;
  if (!r.valid) return fail(c, PBF_ERR_INVALID, "no double-precision rollout has run on this context");
  if (std::memcmp(&r.p, &c->p, sizeof(pbf_params)) != 0)
    return fail(c, PBF_ERR_INVALID, "parameters changed since the rollout");
  hipStream_t st = (hipStream_t)stream;
  const int n = r.n, K = r.iters;
  int s = ensure_particles(c, n, r.nscenes);
  if (s) return s;
  if ((s = adapt_reservations(c))) return s;
  if ((s = ensure_roll64(c, r.nscenes))) return s;
  const int m_max = std::min(r.every, r.nsteps);
  if ((s = ensure_tape(c, m_max))) return s;
  const size_t cap = (size_t)c->cap_n, state = 6 * (size_t)n;
  if ((s = ensure_buf(c, r.seg, r.seg_cap, 2 * state))) return s;
  if ((s = ensure_buf(c, r.adj, r.adj_cap, 7 * cap))) return s;
  if ((s = ensure_buf(c, r.out, r.out_cap, 6 * cap))) return s;
  const size_t per_slot = 2 * (size_t)std::max(K, 1) * cap;
  for (int q = 0; q < m_max; ++q) {
    auto& t = c->slots[q];
    if (t.it64_n < per_slot) {
      dfree(t.it64);
      t.it64_n = 0;
      PBF_HIP(c, hipMalloc(&t.it64, sizeof(double4) * per_slot));
      t.it64_n = per_slot;
    }
  }
  AdjBufs64 ab;
  ab.e = r.adj;
  ab.e2 = r.adj + cap;
  ab.f = r.adj + 2 * cap;
  ab.c = r.adj + 3 * cap;
  ab.q1 = r.adj + 4 * cap;
  ab.q2 = r.adj + 5 * cap;
  ab.vsb = r.adj + 6 * cap;
  std::vector<double4*> P(std::max(K, 1)), aux(std::max(K, 1));
  auto slot_arrays = [&](pbf_ctx::TapeSlot& t) {
    for (int k = 0; k < K; ++k) {
      P[k] = t.it64 + (size_t)k * cap;
      aux[k] = t.it64 + (size_t)(K + k) * cap;
    }
  };
  double* sa = r.seg;
  double* sb = r.seg + state;
  const int nck = (r.nsteps + r.every - 1) / r.every;
  PBF_HIP(c, hipMemsetAsync(r.flag, 0, sizeof(int), st));
  for (int ck = nck - 1; ck >= 0; --ck) {
    const int s0 = ck * r.every, m = std::min(s0 + r.every, r.nsteps) - s0;
    PBF_HIP(c, hipMemcpyAsync(sa, r.ckpt + (size_t)ck * state, sizeof(double) * state, hipMemcpyDeviceToDevice, st));
    for (int q = 0; q < m; ++q) {
      const double* in = (q & 1) ? sb : sa;
      double* xout = q + 1 == m ? r.out : ((q & 1) ? sa : sb);
      double* vout = q + 1 == m ? r.out + 3 * cap : xout + 3 * (size_t)n;
      auto& t = c->slots[q];
      swap_slot(c, t);
      slot_arrays(t);
      s = step64_once(c, in, in + 3 * (size_t)n, xout, vout, nullptr, n, r.off, r.nscenes, P.data(), aux.data(), st);
      if (!s) {
        launch_copy_cells(c->cell_start, t.cell_start, c->total_cells, c->cell_cap, st);
        if (hipMemcpyAsync(t.win, c->win, sizeof(SceneWin) * (size_t)r.nscenes, hipMemcpyDeviceToDevice, st) !=
                hipSuccess ||
            hipMemcpyAsync(t.counters, c->counters, sizeof(unsigned long long) * C_COUNT, hipMemcpyDeviceToDevice,
                           st) != hipSuccess)
          s = fail(c, PBF_ERR_HIP, "tape copy failed");
      }
      swap_slot(c, t);
      if (s) return s;
    }
    for (int q = m - 1; q >= 0; --q) {
      auto& t = c->slots[q];
      slot_arrays(t);
      launch64_adjoint_step(slot_view(c, t, n), P.data(), aux.data(), K, c->pc64, t.perm, d_xbar, d_vbar, ab, st,
                            pass_lanes_for(c, n));
      PBF_LAUNCHED(c);
    }
  }
  c->last_n = -1;
  if ((s = record_counters(c, st))) return s;
  return read_flag64(c, st, "pbf_rollout_backward_f64");
}

int pbf_neighbor_query_fixed(pbf_ctx* c, const float* d_x, int n, int limit, int* d_counts, int* d_elems,
                             void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_counts || !d_elems))) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (limit < 1 || limit > kFixedMax)
    return fail(c, PBF_ERR_UNSUPPORTED, "neighbor limit must be in [1, " + std::to_string(kFixedMax) + "]");
  if (n == 0) return PBF_OK;
  hipStream_t st = (hipStream_t)stream;
  int s = ensure_particles(c, n, 1);
  if (s) return s;
  s = adapt_reservations(c);
  if (s) return s;
  launch_set_offsets(c->off1, n, st);
  s = search(c, d_x, nullptr, n, c->off1, 1, false, st);
  if (s) return s;
  launch_nbr_fixed(view_of(c, n), c->perm, limit, d_counts, d_elems, st);
  PBF_LAUNCHED(c);
  return record_counters(c, st);
}

int pbf_pbd_constrain_mass(pbf_ctx* c, float* d_x, float* d_v, const int* d_index, const int* d_elems,
                           long long nelems, const float* d_mass_inv, float* d_rho, int n, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_v || !d_index || !d_mass_inv)))
    return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n == 0) return PBF_OK;
  Csr nb;
  int s = caller_csr(c, d_index, d_elems, nelems, n, &nb);
  if (s) return s;
  s = ensure_particles(c, n, 1);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  launch_aos_to_f4(d_x, c->xa, n, st);
  pbd_csr_rounds(c, c->xa, c->dpos, nb, d_x, d_v, d_rho, n, d_mass_inv, st);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int pbf_viscosity(pbf_ctx* c, const float* d_x, float* d_v, const int* d_index, const int* d_elems,
                  long long nelems, int n, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_v || !d_index))) return fail(c, PBF_ERR_INVALID, "bad argument");
  const int T = c->p.viscosity_iterations;
  if (n == 0 || T == 0) return PBF_OK;
  Csr nb;
  int s = caller_csr(c, d_index, d_elems, nelems, n, &nb);
  if (s) return s;
  s = ensure_particles(c, n, 1);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  launch_aos_to_f4(d_x, c->xa, n, st);
  launch_aos_to_f4(d_v, c->vv, n, st);
  float4* V[2] = {c->xb, c->x0};
  for (int t = 0; t < T; ++t) {
    const bool last = t + 1 == T;
    const float4* vin = t == 0 ? c->vv : V[(t - 1) & 1];
    launch_csr_viscosity(c->xa, nb, c->vv, vin, last ? nullptr : V[t & 1], last ? d_v : nullptr, c->vc, st);
  }
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int pbf_box_constrain(pbf_ctx* c, float* d_x, float* d_v, int n, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (bad_n(n) || (n > 0 && (!d_x || !d_v))) return fail(c, PBF_ERR_INVALID, "bad argument");
  if (n == 0) return PBF_OK;
  hipStream_t st = (hipStream_t)stream;
  int s = ensure_particles(c, n, 1);
  if (s) return s;
  if ((s = ensure_box(c, st))) return s;
  launch_set_offsets(c->off1, n, st);
  launch_box_constrain(c->bf, d_x, d_v, d_x, d_v, n, c->off1, 1, c->keep_normal, c->damp, st);
  PBF_LAUNCHED(c);
  return PBF_OK;
}

int pbf_get_stats(pbf_ctx* c, pbf_stats* out) {
  if (!c || !out) return fail(c, PBF_ERR_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  // counters_pending stays set: the next call still sees the completed copy
  // and grows the reservations if needed.
  if (c->counters_pending) PBF_HIP(c, hipEventSynchronize(c->counters_ev));
  out->neighbor_pairs = (long long)c->h_counters[C_PAIRS];
  out->out_of_grid = (long long)c->h_counters[C_OOG];
  out->nonfinite = (long long)c->h_counters[C_NONFINITE];
  out->mask_words_needed = (long long)c->h_counters[C_MASK_WORDS];
  out->mask_words_reserved = c->mask_cap;
  out->mask_fallback_waves = (long long)c->h_counters[C_FALLBACK];
  out->list_slots_needed = (long long)c->h_counters[C_LIST_SLOTS];
  out->list_slots_reserved = c->list_cap;
  out->cells_used = (long long)c->h_counters[C_CELLS];
  out->candidates = (long long)c->h_counters[C_CANDIDATES];
  out->staged_waves = (long long)c->h_counters[C_STAGED];
  out->grid_dims[0] = c->g.nx;
  out->grid_dims[1] = c->g.ny;
  out->grid_dims[2] = c->g.nz;
  for (int a = 0; a < 3; ++a) out->grid_lo[a] = c->g.lo[a];
  out->cell_size = c->g.ds;
  out->density_factor = c->m_factor;
  out->particle_mass = c->mass;
  out->neighbor_threshold = c->pc.thr2;
  out->graph_captures = c->graph_captures;
  out->graph_replays = c->graph_replays;
  out->pass_lanes = c->last_lanes;
  return PBF_OK;
}

int pbf_set_profiling(pbf_ctx* c, int enable) {
  if (!c) return PBF_ERR_INVALID;
  int s = drain_marks(c);
  if (s) return s;
  if (enable) {
    for (int i = 0; i < PBF_PASS_COUNT; ++i) {
      c->prof_ms[i] = 0.0;
      c->prof_n[i] = 0;
    }
  }
  c->profiling = enable != 0;
  return PBF_OK;
}

int pbf_get_profile(pbf_ctx* c, pbf_profile* out) {
  if (!c || !out) return fail(c, PBF_ERR_INVALID, "null argument");
  int s = drain_marks(c);
  if (s) return s;
  for (int i = 0; i < PBF_PASS_COUNT; ++i) {
    out->ms[i] = c->prof_ms[i];
    out->launches[i] = c->prof_n[i];
  }
  return PBF_OK;
}

int pbf_set_tuning(pbf_ctx* c, const pbf_tuning* t) {
  if (!c || !t) return fail(c, PBF_ERR_INVALID, "null argument");
  if (t->pass_block != 64 && t->pass_block != 128 && t->pass_block != 256 && t->pass_block != 512)
    return fail(c, PBF_ERR_INVALID, "pass_block must be 64, 128, 256 or 512");
  if (t->persistent != 0 && t->persistent != 1) return fail(c, PBF_ERR_INVALID, "persistent must be 0 or 1");
  if (t->pass_lanes < 0 || t->pass_lanes > 2) return fail(c, PBF_ERR_INVALID, "pass_lanes must be 0, 1 or 2");
  if (t->pass_lanes == 2 && t->persistent != 0)
    return fail(c, PBF_ERR_UNSUPPORTED, "pass_lanes 2 has no persistent form");
  if (t->pass_block != c->pass_block || t->stage_min_particles != c->p.stage_min_particles ||
      (t->persistent != 0) != c->pass_persistent || t->pass_lanes != c->pass_lanes) {
    c->pass_block = t->pass_block;
    c->p.stage_min_particles = t->stage_min_particles;
    c->pass_persistent = t->persistent != 0;
    c->pass_lanes = t->pass_lanes;
    c->gen++;  // a captured step holds the launch shapes and the staging choice
  }
  return PBF_OK;
}

int pbf_get_tuning(const pbf_ctx* c, pbf_tuning* out) {
  if (!c || !out) return PBF_ERR_INVALID;
  out->pass_block = c->pass_block;
  out->stage_min_particles = c->p.stage_min_particles;
  out->persistent = c->pass_persistent ? 1 : 0;
  out->pass_lanes = c->pass_lanes;
  return PBF_OK;
}

const char* pbf_last_error(const pbf_ctx* c) {
  if (!c) return "null context";
  return c->err.c_str();
}

}  // extern "C"
