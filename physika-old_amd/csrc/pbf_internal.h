// Internal types and kernel launchers of libpbf.so (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

struct pbf_ctx;

namespace pbf {

// Start and stop events attached to one kernel dispatch (hipExtLaunchKernel:
// the events time the kernel itself, with nothing recorded between kernels).
struct KernelEvents {
  hipEvent_t start, stop;
};

// Uniform grid of GridHash::setSpace (Framework/Topology/GridHash.cu:47-75):
// cell = h, origin = scene_lo - 2 cells, n = ceil((hi - lo)/h) + 1 + 4 per
// axis. Every scene of a batch uses this geometry, so a particle's cell
// coordinates are exactly those of a one-scene run.
// The reference grid (GridHash::setSpace): nx x ny x nz cells of size ds from
// lo. The sort grid refines x into half cells (index fi = floor(2q); the
// reference cell is fi >> 1); ext_margin (in half cells) bounds how far float
// rounding can move a neighbor pair's half-cell distance past 2.
struct GridDev {
  float lo[3];
  float ds;
  int nx, ny, nz;
  float ext_margin;
};

// A scene's cells are stored in a window: the bounding box of the sort-grid
// cells (half cells along x) of its in-grid particles (cells outside it are
// empty), indexed cell_base + (i - lo0) + dim0*((j - lo1) + dim1*(k - lo2)).
// The window is followed by the scene's bucket of out-of-grid particles at
// cell_base + ncell.
struct SceneWin {
  int lo[3];
  int dim[3];
  int cell_base;
  int ncell;
};

// Per-step constants, derived on the host from pbf_params.
struct PhysConst {
  float h, inv_h;
  float thr2;      // squared-distance bound: sqrtf(s) < h  <=>  s < thr2
  float cW;        // m_factor * mass * 15/(pi h^3): rho_i = cW * sum (1-q)^3
  float cG;        // -45/(pi h^3): Gradient = cG (1-q)^2
  float lam_rest;  // 1000
  float lam_eps;   // 0.1
  float corr;      // 10
  float pair_eps;  // 1e-6
  float dt;
  float g[3];      // gravity
  float gdt[3];    // (0 + g) * dt: velocity increment of the integrator at zero force
};

// Sampled box field (DistanceField3D, Framework/Topology/DistanceField3D.h):
// nx*ny*nz samples at left + (i,j,k)*h, x fastest; inverted (positive
// inside the box).
struct BoxField {
  float* d;
  float left[3], h[3], inv_h[3];
  int nx, ny, nz;
};

// ImplicitViscosity constants: w = cw (1 - r/h), b = dt mu / h.
struct VisConst {
  float inv_h, cw, b, eps;
};

// A caller's neighbor list in the layout of the reference's NeighborList
// (Framework/Topology/NeighborList.h:27-66): index[0..n-1] exclusive row
// starts and elems[0..nelems-1]; row i ends at index[i+1], the last row at
// nelems (NeighborList::getNeighborSize).
struct Csr {
  const int* index;
  const int* elems;
  int n;
  int nelems;
  __device__ __forceinline__ int begin(int i) const { return index[i]; }
  __device__ __forceinline__ int end(int i) const { return i + 1 < n ? index[i + 1] : nelems; }
};

enum Counter {
  C_PAIRS = 0, C_OOG, C_NONFINITE, C_MASK_WORDS, C_FALLBACK, C_LIST_SLOTS, C_CELLS, C_CANDIDATES, C_LIST_NEED,
  C_STAGED, C_COUNT
};

// Runs are visited in the order q = 0..8 of run_of(q): first the three rows of
// the particle's own z-plane (dk = 0), which hold most of its neighbors, then
// the other six.
__host__ __device__ constexpr int run_of(int q) { return q < 3 ? q + 3 : (q < 6 ? q - 3 : q); }

// List encodings (list_mode bits 0-1). Particles are processed in wavefronts
// of 64 consecutive sorted indices; a wave's neighbor list holds, in groups of
// four per lane:
//   kListDelta16: 16-bit j - list_jbase;
//   kListIndex32: 32-bit j (the wave's candidates span more than 65536 indices);
//   kListStaged:  a lane's first neighbors in the first three runs (q < 3), at
//                 most kA of them, as 16-bit slots of the wave's stage, padded
//                 with kStagePad to kA entries, then its other neighbors as
//                 16-bit j - list_jbase; kA is the wave's mean count in those
//                 runs rounded up to 8. The stage concatenates the union over
//                 the wave's lanes of run q, q = 0..2: slot P_q + (j - A_q) for
//                 j in interval q.
// A staged wave's PBD passes copy the stage (at most kStageCap records) into
// LDS once and read those neighbors from there.
constexpr int kListDelta16 = 0, kListIndex32 = 1, kListStaged = 2;
constexpr int kStageCap = 344;
constexpr uint32_t kStagePad = kStageCap;  // the stage slot of the padding particle
// Per wave, stage[kStageInts * wave + ...]: A_0, A_1, A_2, P_0, P_1, P_2, P_3 (the
// stage's length), kA (a multiple of 8).
constexpr int kStageInts = 8;
// Default of pbf_params.stage_min_particles: searches over fewer particles
// keep unstaged lists (their passes run about one round of waves,
// latency-bound, and the copy into LDS lengthens each wave).
constexpr int kStageMinParticles = 1 << 18;

// Out-of-grid particles that list one hashed particle, kept per particle for
// the displacement pass (a particle listed by more takes a scan of the scene's
// out-of-grid particles instead).
constexpr int kOogRev = 16;

// Cell-sorted view used by the neighbor pass and every pass over its lists.
struct NbrView {
  GridDev g;
  const SceneWin* win;         // per scene
  const int* ssid;             // scene of each sorted particle
  const int* cell_start;       // cell_start[c]..cell_start[c+1]: sorted range of compact cell c
  const float4* xn;            // sorted positions the neighbor sets are defined on
  const float* cx;             // sorted x^n coordinate planes (the neighbor pass reads
  const float* cy;             // candidates from these: one 16-byte load brings the
  const float* cz;             // same coordinate of four consecutive candidates)
  const int* nbr_cnt;          // neighbors per sorted particle
  const uint16_t* cnt_a;       // staged waves: stage-slot entries per particle (at most kA)
  const long long* list_base;  // per wave: first 16-bit slot of its list, -1 = no list
  const int* list_jbase;       // per wave: index base of 16-bit offsets
  const int* list_mode;        // per wave: bits 0-1: the encoding (kList*); bits 2..: groups of
                               // four that every lane of the wave fills (kListStaged: of the
                               // part after the stage slots)
  const int* stage;            // per wave: the stage (kStageInts ints; kListStaged only)
  const uint16_t* list;
  const long long* mask_base;  // per wave: first word of its global bit masks, -1 = none
  const uint32_t* mask;        // wave-interleaved bit masks over the cell runs
  const int* oog_rcnt;         // per sorted particle: out-of-grid particles that list it
  const int* oog_rev;          // their sorted indices, kOogRev slots per particle (in no order)
  const unsigned long long* counters;
  float thr2;
  int n;
  int pad;            // padding particle of the position buffers (index n)
};

// Position of the padding particle (index n of every sorted position buffer):
// far outside every stencil, so a pass's terms for it are exact zeros, and
// small enough that its squared distances stay finite.
constexpr float kPadCoord = 1.0e9f;

// Output arrays of the neighbor pass.
struct NbrOut {
  uint32_t* mask;
  long long mask_cap;   // words
  uint16_t* list;
  int list_per;         // 16-bit slots per particle: wave w owns [64 w list_per, 64 (w+1) list_per)
  long long* mask_base;
  long long* list_base;
  int* list_jbase;
  int* list_mode;
  int* stage;
  int stage_lists;      // 1: waves that allow it write kListStaged lists
  int* nbr_cnt;
  uint16_t* cnt_a;
  int* oog_rcnt;        // zeroed by the sort; out-of-grid lanes count themselves into their neighbors'
  int* oog_rev;
  int4* wave_stat;      // per wave: neighbor pairs, candidates, list slots per lane it needs, staged (0/1)
};

// Sorting state of a step (device arrays).
struct SortBufs {
  const int* scene_off;   // nscenes + 1 particle offsets
  int nscenes;
  int* sid;               // scene of each input particle (in the input's order)
  int* ssid;              // scene of each sorted particle
  int* scene_minmax;      // 6 per scene: min i,j,k then max i,j,k of in-grid cells
  SceneWin* win;          // per scene
  int* total_cells;       // device scalar: sum over scenes of ncell + 1
  long long cell_cap;     // allocated compact cells
  int* cell_cnt;          // cell_cap + 1
  int* cell_start;        // cell_cap + 2
  int* key;               // compact cell of each input particle
  int* slot;              // its arrival slot in the cell
  int* tmp;               // input indices scattered into cell order
  int* perm;              // original index of each sorted particle
  int* scan_part;         // partial sums of the device-length scan
  int* bpart;             // per block of the bounds pass: scene and cell box (bounds_part_ints)
  float* xc;              // sorted x^n as three planes (x, y, z), plane stride xc_stride
  int xc_stride;
  int* oog_rcnt;          // per sorted particle, zeroed by the reorder
  int* tid;               // original index of each particle of tmp (input in cell order only)
  int* oog_flag;          // per original index: 1 when out of the grid (n + 1; zero between sorts)
  int* oog_pre;           // exclusive scan of oog_flag (n + 1)
};

// Input of a sort: the caller's state in original order (x, v: xyz float
// triples), or, with x4 set, the state the previous step of the same call
// left in its cell order (x4, v4: one float4 per particle; id: the original
// index of each).
struct SortIn {
  const float* x;
  const float* v;
  const float4* x4;
  const float4* v4;
  const int* id;
};

constexpr int kBlock = 256;
// Particles per call: sorted arrays are indexed with 32-bit byte offsets
// (16 bytes per particle, plus the padding entries).
constexpr int kMaxParticles = 1 << 27;
// Entries allocated past the end of the sorted position array xn, so that a
// candidate group of 8 may read past a run's end.
constexpr int kXnPad = 32;
// Blocks of the device-length scan.
constexpr int kScanDevBlocks = 128;

size_t scan_scratch_ints(long long n);
size_t bounds_part_ints(long long n);
// out[0..n] = exclusive scan of in[0..n-1] (out[n] = total); in may equal out.
void scan_exclusive(const int* in, int* out, int n, int* scratch, hipStream_t st);

// Counting sort of a batch by (scene, window cell); integrate != nullptr fuses
// the integrator into the reorder.
void launch_sort(const SortBufs& sb, const GridDev& g, const SortIn& in, int n, float4* xn, float4* x0,
                 float4* vs, const PhysConst* integrate, unsigned long long* counters, hipStream_t st);
void launch_scene_reset(int* minmax, int nscenes, hipStream_t st);
void launch_set_offsets(int* off, int n, hipStream_t st);
// cell_start[0..*total_cells] from src to dst (cap: allocated cells).
void launch_copy_cells(const int* src, int* dst, const int* total_cells, long long cap, hipStream_t st);
void launch_neighbors(const NbrView& nv, const NbrOut& out, unsigned long long* counters, hipStream_t st, const KernelEvents* ev = nullptr);
// Sums the per-wave statistics of the last neighbor pass (nwaves waves) into
// counters[C_PAIRS], [C_CANDIDATES], [C_LIST_SLOTS] and [C_LIST_NEED].
void launch_wave_stats(const int4* wave_stat, int nwaves, unsigned long long* counters, hipStream_t st);
// block: threads per workgroup, 64, 128, 256 or 512; persistent: as many
// blocks as stay resident, each wave walking 64-particle tiles; lanes: 1, or 2
// for a lane pair per particle (not persistent). The results depend on none of
// them; each wave of a block has its own stage.
void launch_pass_density_lambda(const NbrView& nv, float4* xs, const PhysConst& pc,
                                float* rho_out, const int* perm, hipStream_t st, float4* aux = nullptr,
                                int block = kBlock, bool persistent = false, int lanes = 1,
                                const KernelEvents* ev = nullptr);
// last: write x/v to original order, or with v_sorted the positions to xd and
// the velocities to v_sorted (sorted order, for the viscosity passes or the
// next step of the call; id_out then receives each particle's original index).
void launch_pass_displacement(const NbrView& nv, const float4* xs, float4* xd, const PhysConst& pc,
                              bool last, const float4* x0, const float4* vs, const int* perm,
                              float* x_out, float* v_out, hipStream_t st, float4* v_sorted = nullptr,
                              int* id_out = nullptr, int block = kBlock, bool persistent = false,
                              int lanes = 1, const KernelEvents* ev = nullptr);
void launch_write_back(const float4* xs, const float4* vs, const int* perm, int n, float* x_out,
                       float* v_out, hipStream_t st);
void launch_invert(const int* perm, int* inv, int n, hipStream_t st);
void launch_csr_count(const int* nbr_cnt, const int* inv, int n, int* cnt_orig, hipStream_t st);
void launch_csr_fill(const NbrView& nv, const int* perm, const int* offsets, int* elems,
                     long long cap, hipStream_t st);

// Adjoint of one taped step (sorted order inside, original order at the
// boundary). P[k] (k < K): positions and lambda of PBD iteration k; aux[k]:
// (sum_j t^2/r d_ij, 1/(S + eps)) of that iteration.
struct AdjBufs {
  float4* e;    // adjoint of the current iterate minus c
  float4* e2;   // next e (ping-pong)
  float4* f;    // e + c, gathered by the pair terms
  float4* c;    // v_bar' / dt
  float4* q1;   // (f, rho_bar)
  float4* q2;   // (S_bar * G, S_bar)
  float4* vsb;  // adjoint of the integrated velocities
  int* flag;    // set when a taped step had out-of-grid particles
};
void launch_adjoint_step(const NbrView& nv, float4* const* P, float4* const* aux, int K, const PhysConst& pc,
                         const int* perm, float* xbar, float* vbar, const AdjBufs& ab, hipStream_t st,
                         int lanes = 1);

// Double-precision step and adjoint (reverse mode against FP64 references):
// the constants of PhysConst in double, computed from the parameters
// promoted to double.
struct PhysConst64 {
  double h, inv_h, cW, cG, lam_rest, lam_eps, corr, pair_eps, dt;
  double g[3];
};
struct AdjBufs64 {
  double4 *e, *e2, *f, *c, *q1, *q2, *vsb;
};
// xf[0..3n) = (float) x[0..3n)
void launch64_round(const double* x, float* xf, int n, hipStream_t st);
// One step on the sorted view of the search just run on the rounded
// positions: P[0] receives the integrated positions (w: lambda of iteration
// 0), P[k] iterate k (w: its lambda), aux[k] (taped steps, else nullptr)
// (G, 1/(S + eps)) of iteration k; vs the integrated velocities. Sets *flag
// when the search had particles outside the grid.
void launch64_step(const NbrView& nv, const int* perm, const double* x_in, const double* v_in, double* x_out,
                   double* v_out, double* rho, double4* const* P, double4* const* aux, int K, double4* vs,
                   const PhysConst64& pc, const unsigned long long* counters, int* flag, hipStream_t st,
                   int lanes = 1);
void launch64_adjoint_step(const NbrView& nv, double4* const* P, double4* const* aux, int K, const PhysConst64& pc,
                           const int* perm, double* xbar, double* vbar, const AdjBufs64& ab, hipStream_t st,
                           int lanes = 1);

// Box boundary and viscosity.
void launch_box_sdf(const BoxField& f, const float lo[3], const float hi[3], hipStream_t st);
void launch_box_constrain(const BoxField& f, const float* x_in, const float* v_in, float* x_out, float* v_out, int n,
                          const int* scene_off, int nscenes, float keep_normal, float damp, hipStream_t st);
// lanes: 1, or 2 for a lane pair per particle (same results)
void launch_viscosity(const NbrView& nv, const float4* xs, const float4* v_old, const float4* v_in, float4* v_next,
                      const VisConst& vc, const int* perm, float* x_out, float* v_out, hipStream_t st, int lanes = 1);
void launch_csr_viscosity(const float4* xa, const Csr& nb, const float4* v_old, const float4* v_in, float4* v_next,
                          float* v_out, const VisConst& vc, hipStream_t st);

// Module-granular kernels, original particle order, caller CSR.
void launch_integrate(float* x, float* v, const float* f, int n, const PhysConst& pc, hipStream_t st);
void launch_csr_density(const float* x, const Csr& nb, float* rho, const PhysConst& pc, hipStream_t st);
void launch_aos_to_f4(const float* x, float4* xa, int n, hipStream_t st);
// as launch_aos_to_f4, and dpos[0 .. 3n) := 0
void launch_aos_to_f4_clear(const float* x, float4* xa, float* dpos, int n, hipStream_t st);
void launch_csr_lambda(float4* xa, const Csr& nb, float* rho, const PhysConst& pc, hipStream_t st,
                       const float* minv = nullptr);
void launch_csr_dx_atomic(const float4* xa, const Csr& nb, float* dpos, const PhysConst& pc, hipStream_t st,
                          const float* minv = nullptr);
// Fixed-capacity rows (the `limit` nearest of each neighbor set, by distance
// then id) in original ids; limit <= kFixedMax.
constexpr int kFixedMax = 32;
void launch_nbr_fixed(const NbrView& nv, const int* perm, int limit, int* counts, int* elems, hipStream_t st);
void launch_csr_apply(float4* xa, float* dpos, int n, hipStream_t st);
void launch_csr_finish(const float4* xa, float* x, float* v, int n, float dt, hipStream_t st);

// Records msg as the context's last error and returns status (host code of
// the library's other translation units).
int ctx_fail(pbf_ctx* c, int status, const std::string& msg);

}  // namespace pbf
