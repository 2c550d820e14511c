// Internal types and kernel launchers of libpbf.so (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pbf {

// Uniform grid of GridHash::setSpace (Framework/Topology/GridHash.cu:47-75):
// cell = h, origin = scene_lo - 2 cells, n = ceil((hi - lo)/h) + 1 + 4 per
// axis, linear id i + nx*(j + ny*k). Cell `ncell` is the bucket that holds the
// particles outside the grid.
struct GridDev {
  float lo[3];
  float ds;
  int nx, ny, nz;
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

enum Counter { C_PAIRS = 0, C_OOG, C_NONFINITE, C_MASK_WORDS, C_FALLBACK, C_LIST_SLOTS, C_COUNT };

// Cell-sorted view used by the neighbor pass and every pass over its lists.
// Particles are processed in wavefronts of 64 consecutive sorted indices; a
// wave's neighbor list holds entry k of lane l at list_base + 64*k + l (16-bit
// j - list_jbase, or a 32-bit j in two slots when the wave's candidates span
// more than 65536 indices), so the lanes of a wave read consecutive slots.
struct NbrView {
  GridDev g;
  const int* cell_start;       // ncell + 2 entries; cell_start[ncell] = in-grid count
  const float4* xn;            // sorted positions the neighbor sets are defined on
  const int* nbr_cnt;          // neighbors per sorted particle
  const long long* list_base;  // per wave: first 16-bit slot of its list, -1 = no list
  const int* list_jbase;       // per wave: index base of 16-bit entries
  const int* list_mode;        // per wave: 0 = 16-bit deltas, 1 = 32-bit indices
  const uint16_t* list;
  const long long* mask_base;  // per wave: first word of its bit masks, -1 = none
  const uint32_t* mask;        // wave-interleaved bit masks over the cell runs
  const uint8_t* oog_flag;     // per cell: an out-of-grid particle lists this cell
  float thr2;
  int n;
};

// Output arrays of the neighbor pass.
struct NbrOut {
  uint32_t* mask;
  long long mask_cap;   // words
  uint16_t* list;
  long long list_cap;   // 16-bit slots
  long long* mask_base;
  long long* list_base;
  int* list_jbase;
  int* list_mode;
  int* nbr_cnt;
};

constexpr int kBlock = 256;
// Entries allocated past the end of the sorted position array xn, so that a
// candidate group of 8 (or a 32-candidate word) may read past a run's end.
constexpr int kXnPad = 32;

size_t scan_scratch_ints(long long n);
// out[0..n] = exclusive scan of in[0..n-1] (out[n] = total); in may equal out.
void scan_exclusive(const int* in, int* out, int n, int* scratch, hipStream_t st);

void launch_clear(int* cell_cnt, uint8_t* oog_flag, int ncell, unsigned long long* counters,
                  hipStream_t st);
void launch_bin(const float* x, int n, GridDev g, int* key, int* slot, int* cell_cnt,
                unsigned long long* counters, hipStream_t st);
void launch_scatter(const int* key, const int* slot, const int* cell_start, int* tmp, int n,
                    hipStream_t st);
void launch_reorder(const int* key, const int* cell_start, const int* tmp, const float* x,
                    const float* v, int n, int* perm, int* inv, float4* xn, float4* x0, float4* vs,
                    const PhysConst* integrate, hipStream_t st);
void launch_neighbors(const NbrView& nv, const NbrOut& out, unsigned long long* counters, hipStream_t st);
void launch_oog_mark(const NbrView& nv, uint8_t* oog_flag, hipStream_t st);
void launch_pass_density_lambda(const NbrView& nv, float4* xs, const PhysConst& pc,
                                float* rho_out, const int* perm, hipStream_t st);
void launch_pass_displacement(const NbrView& nv, const float4* xs, float4* xd, const PhysConst& pc,
                              bool last, const float4* x0, const float4* vs, const int* perm,
                              float* x_out, float* v_out, hipStream_t st);
void launch_write_back(const float4* xs, const float4* vs, const int* perm, int n, float* x_out,
                       float* v_out, hipStream_t st);
void launch_csr_count(const int* nbr_cnt, const int* inv, int n, int* cnt_orig, hipStream_t st);
void launch_csr_fill(const NbrView& nv, const int* perm, const int* offsets, int* elems,
                     long long cap, hipStream_t st);

// Module-granular kernels, original particle order, caller CSR.
void launch_integrate(float* x, float* v, const float* f, int n, const PhysConst& pc, hipStream_t st);
void launch_csr_density(const float* x, const int* off, const int* el, float* rho, int n,
                        const PhysConst& pc, hipStream_t st);
void launch_aos_to_f4(const float* x, float4* xa, int n, hipStream_t st);
void launch_csr_lambda(float4* xa, const int* off, const int* el, float* rho, int n,
                       const PhysConst& pc, hipStream_t st);
void launch_csr_dx_atomic(const float4* xa, const int* off, const int* el, float* dpos, int n,
                          const PhysConst& pc, hipStream_t st);
void launch_csr_apply(float4* xa, float* dpos, int n, hipStream_t st);
void launch_csr_finish(const float4* xa, float* x, float* v, int n, float dt, hipStream_t st);

}  // namespace pbf
