// HIP kernels of the PBD fluid step for gfx950 (CDNA4, wave64).
//
// Data layout (HBM):
//   caller arrays  x, v: packed xyz float triples in original particle order
//   sorted arrays  xn, x0, xa, xb, vs: float4 per particle in cell order
//                  (w of the position buffers carries lambda between passes)
//   cell arrays    cell_cnt[ncell+1], cell_start[ncell+2] (int), oog_flag[ncell]
//   neighbor masks one 32-bit word per (particle, 32 candidates of a cell run),
//                  interleaved by wavefront: word w of lane l of a wave sits at
//                  wave_base + 64*w + l, so the 64 lanes of a wave read and write
//                  consecutive words.
//
// A particle's 27 stencil cells form 9 runs of 3 x-consecutive cells, and in
// cell order each run is one contiguous index range [a, a+len). The neighbor
// set of the reference (every j in a valid stencil cell with
// sqrtf(|x_i - x_j|^2) < h, NeighborQuery.cu:245-276) is stored as one bit per
// candidate of those runs.
#include <hip/hip_runtime.h>

#include "pbf_internal.h"

namespace pbf {

// ---------------------------------------------------------------------------
// Exactly rounded helpers (bit-identical to the CPU restatement)

// GridHash::getIndex3 (GridHash.h:55-62): floor((p - lo)/ds) with a correctly
// rounded division, clamped to [-2, n+1] (same valid stencil cells; NaN -> -2).
__device__ __forceinline__ int cell_coord(float p, float lo, float ds, int n) {
#pragma clang fp contract(off)
  float f = floorf((p - lo) / ds);
  if (!(f >= -2.0f)) return -2;
  if (f > (float)(n + 1)) return n + 1;
  return (int)f;
}

__device__ __forceinline__ void cell_of(const GridDev& g, float x, float y, float z, int& ci, int& cj,
                                        int& ck) {
  ci = cell_coord(x, g.lo[0], g.ds, g.nx);
  cj = cell_coord(y, g.lo[1], g.ds, g.ny);
  ck = cell_coord(z, g.lo[2], g.ds, g.nz);
}

__device__ __forceinline__ bool in_grid(const GridDev& g, int ci, int cj, int ck) {
  return ci >= 0 && ci < g.nx && cj >= 0 && cj < g.ny && ck >= 0 && ck < g.nz;
}

// |x_i - x_j| < h exactly as NeighborQuery.cu:266-267 evaluates it: the
// squared norm in glm's order ((dx*dx + dy*dy) + dz*dz, no fused multiply-add)
// compared against thr2, the least float whose correctly rounded square root
// reaches h (the square root is monotone, so the two tests agree everywhere).
__device__ __forceinline__ bool within(float4 a, float4 b, float thr2) {
#pragma clang fp contract(off)
  float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  float s = (dx * dx + dy * dy) + dz * dz;
  return s < thr2;
}

// Blocks are dispatched round-robin over the 8 XCDs of the chip; remap so that
// each XCD (own L2) works on one contiguous range of cell-sorted particles.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int per = nb >> 3, rem = nb & 7, x = b & 7, idx = b >> 3;
  return x * per + min(x, rem) + idx;
}

// ---------------------------------------------------------------------------
// Cell runs and the neighbor iterator

// Run r (0..8) of cell (ci,cj,ck): cells (ci-1..ci+1, cj+dj, ck+dk) with
// dj = r/3-1, dk = r%3-1, clipped to the grid; [a, a+len) in sorted order.
__device__ __forceinline__ void run_range(const int* __restrict__ cell_start, const GridDev& g, int ci,
                                          int cj, int ck, int r, int& a, int& len) {
  int j = cj + (r / 3) - 1, k = ck + (r % 3) - 1;
  a = 0;
  len = 0;
  if (j < 0 || j >= g.ny || k < 0 || k >= g.nz) return;
  int i0 = max(ci - 1, 0), i1 = min(ci + 1, g.nx - 1);
  if (i0 > i1) return;
  int row = (j + k * g.ny) * g.nx;
  a = cell_start[row + i0];
  len = cell_start[row + i1 + 1] - a;
}

// Bits of the candidates [j0, j0+cnt) (cnt <= 32) within h of xi.
__device__ __forceinline__ uint32_t test_word(const float4* __restrict__ xn, float4 xi, int j0, int cnt,
                                              float thr2) {
  uint32_t bits = 0;
  for (int k = 0; k < cnt; ++k)
    if (within(xi, xn[j0 + k], thr2)) bits |= 1u << k;
  return bits;
}

// Calls body(j) for every neighbor j of sorted particle s, in run order. The
// loop is flat: one trip per neighbor, refilling the current mask word when it
// is exhausted, so lanes with different bit patterns stay converged on the
// body. With wave_base == -1 (mask reservation exhausted) the words are
// recomputed from xn.
template <class Body>
__device__ __forceinline__ void for_each_neighbor(const NbrView& nv, int s, float4 xni, int ci, int cj,
                                                  int ck, Body&& body) {
  const long long mb = nv.wave_base[s >> 6];
  const uint32_t* mw = nv.mask + (mb >= 0 ? mb + (s & 63) : 0);
  int r = 0, a = 0, len = 0, nw = 0, w = 0, wofs = 0, jb = 0;
  uint32_t bits = 0;
  bool more = true;
  while (more) {
    while (bits == 0) {
      if (w < nw) {
        jb = a + (w << 5);
        bits = mb >= 0 ? mw[(long long)(wofs + w) << 6] : test_word(nv.xn, xni, jb, min(32, len - (w << 5)), nv.thr2);
        ++w;
      } else if (r < 9) {
        wofs += nw;
        run_range(nv.cell_start, nv.g, ci, cj, ck, r, a, len);
        nw = (len + 31) >> 5;
        w = 0;
        ++r;
      } else {
        more = false;
        break;
      }
    }
    if (!more) break;
    int b = __builtin_ctz(bits);
    bits &= bits - 1u;
    body(jb + b);
  }
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d));
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}
__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int t = __shfl_up(v, d);
    if (lane >= d) v += t;
  }
  return v;
}

// ---------------------------------------------------------------------------
// Exclusive scan (Blelloch's recursive decomposition with wave-level
// shuffles): blocks of 1024 threads x 4 elements, recursion over block sums.

constexpr int kScanPerBlock = 4096;

__global__ __launch_bounds__(1024) void k_scan_block(const int* in, int* out, int n, int* block_sums) {
  __shared__ int wsum[16];
  const int base = blockIdx.x * kScanPerBlock + threadIdx.x * 4;
  int a0 = base < n ? in[base] : 0;
  int a1 = base + 1 < n ? in[base + 1] : 0;
  int a2 = base + 2 < n ? in[base + 2] : 0;
  int a3 = base + 3 < n ? in[base + 3] : 0;
  const int tsum = a0 + a1 + a2 + a3;
  const int incl = wave_inclusive_scan(tsum);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    int t = lane < 16 ? wsum[lane] : 0;
    t = wave_inclusive_scan(t);
    if (lane < 16) wsum[lane] = t;
  }
  __syncthreads();
  int e = (wave > 0 ? wsum[wave - 1] : 0) + incl - tsum;
  if (base < n) out[base] = e;
  e += a0;
  if (base + 1 < n) out[base + 1] = e;
  e += a1;
  if (base + 2 < n) out[base + 2] = e;
  e += a2;
  if (base + 3 < n) out[base + 3] = e;
  if (threadIdx.x == 1023) {
    block_sums[blockIdx.x] = wsum[15];
    if (gridDim.x == 1) out[n] = wsum[15];
  }
}

__global__ __launch_bounds__(1024) void k_scan_add(int* out, int n, const int* block_offsets, int nb) {
  const int add = block_offsets[blockIdx.x];
  const int base = blockIdx.x * kScanPerBlock + threadIdx.x * 4;
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (base + t < n) out[base + t] += add;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = block_offsets[nb];
}

size_t scan_scratch_ints(long long n) {
  long long nb = (n + kScanPerBlock - 1) / kScanPerBlock;
  if (nb <= 1) return 4;
  size_t here = (size_t)((nb + 3) / 4 * 4) + (size_t)((nb + 1 + 3) / 4 * 4);
  return here + scan_scratch_ints(nb);
}

void scan_exclusive(const int* in, int* out, int n, int* scratch, hipStream_t st) {
  const int nb = (n + kScanPerBlock - 1) / kScanPerBlock;
  if (nb <= 1) {
    k_scan_block<<<1, 1024, 0, st>>>(in, out, n, scratch);
    return;
  }
  int* sums = scratch;
  int* offs = scratch + (nb + 3) / 4 * 4;
  int* rest = offs + (nb + 1 + 3) / 4 * 4;
  k_scan_block<<<nb, 1024, 0, st>>>(in, out, n, sums);
  scan_exclusive(sums, offs, nb, rest, st);
  k_scan_add<<<nb, 1024, 0, st>>>(out, n, offs, nb);
}

// ---------------------------------------------------------------------------
// Hashing: GridHash::construct (GridHash.cu:298-323) as a counting sort that
// also reorders the particle data.

__global__ void k_clear(int* cell_cnt, uint8_t* oog_flag, int ncell, unsigned long long* counters) {
  const int stride = gridDim.x * blockDim.x;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= ncell; c += stride) {
    cell_cnt[c] = 0;
    if (c < ncell) oog_flag[c] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < C_COUNT) counters[threadIdx.x] = 0;
}

void launch_clear(int* cell_cnt, uint8_t* oog_flag, int ncell, unsigned long long* counters,
                  hipStream_t st) {
  int nb = min((ncell + 1 + kBlock - 1) / kBlock, 4096);
  k_clear<<<nb, kBlock, 0, st>>>(cell_cnt, oog_flag, ncell, counters);
}

// K_CalculateParticleNumber (GridHash.cu:78-87): key = cell id or the
// out-of-grid bucket; slot = arrival order inside the cell.
__global__ __launch_bounds__(kBlock) void k_bin(const float* __restrict__ x, int n, GridDev g,
                                                int* __restrict__ key, int* __restrict__ slot,
                                                int* __restrict__ cell_cnt,
                                                unsigned long long* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool oog = false, nonfinite = false;
  if (i < n) {
    const float px = x[3 * (size_t)i], py = x[3 * (size_t)i + 1], pz = x[3 * (size_t)i + 2];
    nonfinite = !(isfinite(px) && isfinite(py) && isfinite(pz));
    int ci, cj, ck;
    cell_of(g, px, py, pz, ci, cj, ck);
    int k = g.ncell;
    if (!nonfinite && in_grid(g, ci, cj, ck))
      k = ci + g.nx * (cj + g.ny * ck);
    else
      oog = true;
    key[i] = k;
    slot[i] = atomicAdd(&cell_cnt[k], 1);
  }
  const unsigned long long b_oog = __ballot(oog), b_nf = __ballot(nonfinite);
  if ((threadIdx.x & 63) == 0) {
    if (b_oog) atomicAdd(&counters[C_OOG], (unsigned long long)__popcll(b_oog));
    if (b_nf) atomicAdd(&counters[C_NONFINITE], (unsigned long long)__popcll(b_nf));
  }
}

void launch_bin(const float* x, int n, GridDev g, int* key, int* slot, int* cell_cnt,
                unsigned long long* counters, hipStream_t st) {
  k_bin<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, n, g, key, slot, cell_cnt, counters);
}

// K_ConstructHashTable (GridHash.cu:187-198): ids into their cell's slots.
__global__ __launch_bounds__(kBlock) void k_scatter(const int* __restrict__ key, const int* __restrict__ slot,
                                                    const int* __restrict__ cell_start, int* __restrict__ tmp,
                                                    int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tmp[cell_start[key[i]] + slot[i]] = i;
}

void launch_scatter(const int* key, const int* slot, const int* cell_start, int* tmp, int n,
                    hipStream_t st) {
  k_scatter<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(key, slot, cell_start, tmp, n);
}

// Final position inside the cell = rank of the particle id among the cell's
// ids (deterministic order, independent of atomic arrival). Cells holding more
// than kRankMax particles keep arrival order (still a valid cell sort).
// The integrator (ParticleIntegrator.cu:60-70,103-112, force zeroed by begin)
// is fused here: x^n goes to xn (the neighbor sets are built on it,
// PositionBasedFluidModel.cpp:93-94), the integrated state to x0 / vs.
constexpr int kRankMax = 1024;

__global__ __launch_bounds__(kBlock) void k_reorder(const int* __restrict__ key, const int* __restrict__ cell_start,
                                                    const int* __restrict__ tmp, const float* __restrict__ x,
                                                    const float* __restrict__ v, int n, int* __restrict__ perm,
                                                    int* __restrict__ inv, float4* __restrict__ xn,
                                                    float4* __restrict__ x0, float4* __restrict__ vs,
                                                    PhysConst pc, int integrate) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = key[i];
  const int a = cell_start[k], b = cell_start[k + 1];
  int s;
  if (b - a <= kRankMax) {
    int rank = 0;
    for (int t = a; t < b; ++t) rank += tmp[t] < i;
    s = a + rank;
  } else {
    s = -1;
    for (int t = a; t < b; ++t)
      if (tmp[t] == i) s = t;
  }
  perm[s] = i;
  inv[i] = s;
  const float px = x[3 * (size_t)i], py = x[3 * (size_t)i + 1], pz = x[3 * (size_t)i + 2];
  xn[s] = make_float4(px, py, pz, 0.0f);
  if (integrate) {
    const float vx = v[3 * (size_t)i] + pc.gdt[0];
    const float vy = v[3 * (size_t)i + 1] + pc.gdt[1];
    const float vz = v[3 * (size_t)i + 2] + pc.gdt[2];
    x0[s] = make_float4(px + vx * pc.dt, py + vy * pc.dt, pz + vz * pc.dt, 0.0f);
    vs[s] = make_float4(vx, vy, vz, 0.0f);
  } else {
    x0[s] = make_float4(px, py, pz, 0.0f);
  }
}

void launch_reorder(const int* key, const int* cell_start, const int* tmp, const float* x, const float* v,
                    int n, int* perm, int* inv, float4* xn, float4* x0, float4* vs, const PhysConst* integrate,
                    hipStream_t st) {
  PhysConst pc{};
  if (integrate) pc = *integrate;
  k_reorder<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(key, cell_start, tmp, x, v, n, perm, inv, xn, x0, vs,
                                                          pc, integrate ? 1 : 0);
}

// ---------------------------------------------------------------------------
// Neighbor masks: K_CalNeighborSize + K_GetNeighborElements
// (NeighborQuery.cu:245-276,337-367) in one traversal of the candidates.

__global__ __launch_bounds__(kBlock) void k_neighbors(NbrView nv, uint32_t* __restrict__ mask,
                                                      long long* __restrict__ wave_base,
                                                      int* __restrict__ nbr_cnt, long long mask_cap,
                                                      unsigned long long* __restrict__ counters) {
  const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool active = s < nv.n;
  float4 xi = make_float4(0.f, 0.f, 0.f, 0.f);
  int ci = -2, cj = -2, ck = -2;
  int words = 0;
  if (active) {
    xi = nv.xn[s];
    cell_of(nv.g, xi.x, xi.y, xi.z, ci, cj, ck);
    for (int r = 0; r < 9; ++r) {
      int a, len;
      run_range(nv.cell_start, nv.g, ci, cj, ck, r, a, len);
      words += (len + 31) >> 5;
    }
  }
  const int wmax = wave_max(words);
  long long base = 0;
  if (lane == 0) {
    const unsigned long long need = 64ull * (unsigned long long)wmax;
    const unsigned long long at = atomicAdd(&counters[C_MASK_WORDS], need);
    base = (long long)at;
    if ((long long)(at + need) > mask_cap) {
      base = -1;
      atomicAdd(&counters[C_FALLBACK], 1ull);
    }
    wave_base[s >> 6] = base;
  }
  base = __shfl(base, 0);
  int cnt = 0;
  if (active) {
    uint32_t* mw = mask + (base >= 0 ? base + lane : 0);
    int wofs = 0;
    for (int r = 0; r < 9; ++r) {
      int a, len;
      run_range(nv.cell_start, nv.g, ci, cj, ck, r, a, len);
      for (int w = 0; (w << 5) < len; ++w) {
        const uint32_t bits = test_word(nv.xn, xi, a + (w << 5), min(32, len - (w << 5)), nv.thr2);
        if (base >= 0) mw[(long long)(wofs + w) << 6] = bits;
        cnt += __popc(bits);
      }
      wofs += (len + 31) >> 5;
    }
    nbr_cnt[s] = cnt;
  }
  const int tot = wave_sum(cnt);
  if (lane == 0 && tot) atomicAdd(&counters[C_PAIRS], (unsigned long long)tot);
}

void launch_neighbors(const NbrView& nv, uint32_t* mask, long long* wave_base, int* nbr_cnt, long long mask_cap,
                      unsigned long long* counters, hipStream_t st) {
  k_neighbors<<<(nv.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nv, mask, wave_base, nbr_cnt, mask_cap, counters);
}

// Cells inside the stencil of an out-of-grid particle. Such a particle is not
// hashed (GridHash.cu:85,194) yet queries, so an in-grid j it lists receives
// -dp_kj from it (DensityPBD.cu:162) without listing it back.
__global__ void k_oog_mark(NbrView nv, uint8_t* oog_flag) {
  const int n_in = nv.cell_start[nv.g.ncell];
  const int stride = gridDim.x * blockDim.x;
  for (int k = n_in + blockIdx.x * blockDim.x + threadIdx.x; k < nv.n; k += stride) {
    const float4 p = nv.xn[k];
    int ci, cj, ck;
    cell_of(nv.g, p.x, p.y, p.z, ci, cj, ck);
    for (int dk = -1; dk <= 1; ++dk)
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
          const int a = ci + di, b = cj + dj, c = ck + dk;
          if (in_grid(nv.g, a, b, c)) oog_flag[a + nv.g.nx * (b + nv.g.ny * c)] = 1;
        }
  }
}

void launch_oog_mark(const NbrView& nv, uint8_t* oog_flag, hipStream_t st) {
  k_oog_mark<<<64, kBlock, 0, st>>>(nv, oog_flag);
}

// ---------------------------------------------------------------------------
// PBD passes over the masks (DensityPBD::takeOneIteration, DensityPBD.cu:306-365).
//
// Pair terms with t = max(1 - r/h, 0):
//   W(r)   = 15/(pi h^3) t^3             (Kernel.h:36-45)
//   dW/dr  = -45/(pi h^3) t^2            (Kernel.h:47-57)
// r comes from v_rsq_f32 (within 2 ulp); r == 0 (self) gives t = 1.

struct PairGeom {
  float dx, dy, dz, r, rinv, t;
};

__device__ __forceinline__ PairGeom pair_geom(float4 xi, float4 xj, float inv_h) {
  PairGeom p;
  p.dx = xi.x - xj.x;
  p.dy = xi.y - xj.y;
  p.dz = xi.z - xj.z;
  const float r2 = p.dx * p.dx + p.dy * p.dy + p.dz * p.dz;
  p.rinv = __builtin_amdgcn_rsqf(fmaxf(r2, 1e-30f));
  p.r = r2 * p.rinv;
  p.t = fmaxf(1.0f - p.r * inv_h, 0.0f);
  return p;
}

// Pass A: K_ComputeDensity (SummationDensity.cu:13-36) and K_ComputeLambdas
// (DensityPBD.cu:46-90) fused: one traversal gives rho_i, sum g and sum g.g;
// lambda_i goes to xs[s].w for pass B.
__global__ __launch_bounds__(kBlock) void k_density_lambda(NbrView nv, float4* __restrict__ xs, PhysConst pc,
                                                           float* __restrict__ rho_out,
                                                           const int* __restrict__ perm) {
  const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (s >= nv.n) return;
  const float4 xi = xs[s];
  const float4 xni = nv.xn[s];
  int ci, cj, ck;
  cell_of(nv.g, xni.x, xni.y, xni.z, ci, cj, ck);
  float rho = 0.f, sgx = 0.f, sgy = 0.f, sgz = 0.f, sgg = 0.f;
  for_each_neighbor(nv, s, xni, ci, cj, ck, [&](int j) {
    const float4 xj = xs[j];
    const PairGeom p = pair_geom(xi, xj, pc.inv_h);
    const float t2 = p.t * p.t;
    rho += t2 * p.t;
    const float gq = p.r > pc.pair_eps ? t2 * p.rinv : 0.0f;
    const float gx = gq * p.dx, gy = gq * p.dy, gz = gq * p.dz;
    sgx += gx;
    sgy += gy;
    sgz += gz;
    sgg += gx * gx + gy * gy + gz * gz;
  });
  rho *= pc.cW;
  const float S = (pc.cG * pc.cG) * (sgg + (sgx * sgx + sgy * sgy + sgz * sgz));
  float lam = -(rho - pc.lam_rest) / (S + pc.lam_eps);
  lam = lam > 0.0f ? 0.0f : lam;
  reinterpret_cast<float*>(xs + s)[3] = lam;
  if (rho_out) rho_out[perm[s]] = rho;
}

void launch_pass_density_lambda(const NbrView& nv, float4* xs, const PhysConst& pc, float* rho_out,
                                const int* perm, hipStream_t st) {
  k_density_lambda<<<(nv.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nv, xs, pc, rho_out, perm);
}

// Pass B: K_ComputeDisplacement + K_UpdatePosition (DensityPBD.cu:135-179,
// 227-237) in gather form. The reference adds +dp_ij to i and -dp_ij to j from
// both endpoints of every listed pair and dp_ji == -dp_ij, so a hashed particle
// receives 2*sum_j dp_ij; an out-of-grid particle k is listed by nobody and
// receives sum_j dp_kj once, while each hashed j it lists receives dp_jk once
// more. The last iteration fuses DP_UpdateVelocity (DensityPBD.cu:368-378) and
// the scatter back to original order.
__global__ __launch_bounds__(kBlock) void k_displacement(NbrView nv, const float4* __restrict__ xs,
                                                         float4* __restrict__ xd, PhysConst pc, int last,
                                                         const float4* __restrict__ x0,
                                                         const float4* __restrict__ vs,
                                                         const int* __restrict__ perm, float* __restrict__ x_out,
                                                         float* __restrict__ v_out) {
  const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (s >= nv.n) return;
  const float4 xi = xs[s];
  const float4 xni = nv.xn[s];
  const float li = xi.w;
  int ci, cj, ck;
  cell_of(nv.g, xni.x, xni.y, xni.z, ci, cj, ck);
  // sum over the listed j of (lambda_i + lambda_j) t^2 / r (x_i - x_j)
  auto pair = [&](float4 xj, float& cx, float& cy, float& cz) {
    const PairGeom p = pair_geom(xi, xj, pc.inv_h);
    const float c = p.r > pc.pair_eps ? (li + xj.w) * (p.t * p.t) * p.rinv : 0.0f;
    cx += c * p.dx;
    cy += c * p.dy;
    cz += c * p.dz;
  };
  float ax = 0.f, ay = 0.f, az = 0.f;
  for_each_neighbor(nv, s, xni, ci, cj, ck, [&](int j) { pair(xs[j], ax, ay, az); });
  const int n_in = nv.cell_start[nv.g.ncell];
  float fa = 1.0f;  // out-of-grid particle: its own pairs once
  float bx = 0.f, by = 0.f, bz = 0.f;
  if (s < n_in) {
    fa = 2.0f;
    if (nv.oog_flag[ci + nv.g.nx * (cj + nv.g.ny * ck)]) {
      // out-of-grid particles k that list s add -dp_ks = dp_sk to s once
      for (int k = n_in; k < nv.n; ++k) {
        const float4 xnk = nv.xn[k];
        int ki, kj, kk;
        cell_of(nv.g, xnk.x, xnk.y, xnk.z, ki, kj, kk);
        if (abs(ki - ci) <= 1 && abs(kj - cj) <= 1 && abs(kk - ck) <= 1 && within(xnk, xni, nv.thr2))
          pair(xs[k], bx, by, bz);
      }
    }
  }
  const float scale = pc.corr * pc.cG;
  const float nx = xi.x + scale * (fa * ax + bx);
  const float ny = xi.y + scale * (fa * ay + by);
  const float nz = xi.z + scale * (fa * az + bz);
  if (!last) {
    xd[s] = make_float4(nx, ny, nz, 0.0f);
    return;
  }
  const float4 xo = x0[s];
  const float4 vv = vs[s];
  const size_t i = (size_t)perm[s];
  x_out[3 * i] = nx;
  x_out[3 * i + 1] = ny;
  x_out[3 * i + 2] = nz;
  v_out[3 * i] = vv.x + (nx - xo.x) / pc.dt;
  v_out[3 * i + 1] = vv.y + (ny - xo.y) / pc.dt;
  v_out[3 * i + 2] = vv.z + (nz - xo.z) / pc.dt;
}

void launch_pass_displacement(const NbrView& nv, const float4* xs, float4* xd, const PhysConst& pc, bool last,
                              const float4* x0, const float4* vs, const int* perm, float* x_out, float* v_out,
                              hipStream_t st) {
  k_displacement<<<(nv.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nv, xs, xd, pc, last ? 1 : 0, x0, vs, perm,
                                                                 x_out, v_out);
}

__global__ __launch_bounds__(kBlock) void k_write_back(const float4* __restrict__ xs, const float4* __restrict__ vs,
                                                       const int* __restrict__ perm, int n, float* __restrict__ x_out,
                                                       float* __restrict__ v_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = (size_t)perm[s];
  const float4 a = xs[s], b = vs[s];
  x_out[3 * i] = a.x;
  x_out[3 * i + 1] = a.y;
  x_out[3 * i + 2] = a.z;
  v_out[3 * i] = b.x;
  v_out[3 * i + 1] = b.y;
  v_out[3 * i + 2] = b.z;
}

void launch_write_back(const float4* xs, const float4* vs, const int* perm, int n, float* x_out, float* v_out,
                       hipStream_t st) {
  k_write_back<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xs, vs, perm, n, x_out, v_out);
}

// ---------------------------------------------------------------------------
// CSR export (NeighborList, NeighborList.h:27-66) in original ids.

__global__ __launch_bounds__(kBlock) void k_csr_count(const int* __restrict__ nbr_cnt, const int* __restrict__ inv,
                                                      int n, int* __restrict__ cnt_orig) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt_orig[i] = nbr_cnt[inv[i]];
}

void launch_csr_count(const int* nbr_cnt, const int* inv, int n, int* cnt_orig, hipStream_t st) {
  k_csr_count<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nbr_cnt, inv, n, cnt_orig);
}

__global__ __launch_bounds__(kBlock) void k_csr_fill(NbrView nv, const int* __restrict__ perm,
                                                     const int* __restrict__ offsets, int* __restrict__ elems,
                                                     long long cap) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nv.n) return;
  const int i = perm[s];
  int o = offsets[i];
  if ((long long)offsets[i + 1] > cap) return;
  const float4 xni = nv.xn[s];
  int ci, cj, ck;
  cell_of(nv.g, xni.x, xni.y, xni.z, ci, cj, ck);
  for_each_neighbor(nv, s, xni, ci, cj, ck, [&](int j) { elems[o++] = perm[j]; });
}

void launch_csr_fill(const NbrView& nv, const int* perm, const int* offsets, int* elems, long long cap,
                     hipStream_t st) {
  k_csr_fill<<<(nv.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nv, perm, offsets, elems, cap);
}

// ---------------------------------------------------------------------------
// Module-granular kernels, original order, caller CSR.

// ParticleIntegrator::integrate: K_UpdateVelocity then K_UpdatePosition
// (ParticleIntegrator.cu:60-70,103-112), each multiply and add rounded
// separately as the source writes them.
__global__ __launch_bounds__(kBlock) void k_integrate(float* __restrict__ x, float* __restrict__ v,
                                                      const float* __restrict__ f, int n, PhysConst pc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = 3 * (size_t)i;
  float vx, vy, vz;
  if (f) {
    vx = v[b] + (f[b] + pc.g[0]) * pc.dt;
    vy = v[b + 1] + (f[b + 1] + pc.g[1]) * pc.dt;
    vz = v[b + 2] + (f[b + 2] + pc.g[2]) * pc.dt;
  } else {
    vx = v[b] + pc.gdt[0];
    vy = v[b + 1] + pc.gdt[1];
    vz = v[b + 2] + pc.gdt[2];
  }
  v[b] = vx;
  v[b + 1] = vy;
  v[b + 2] = vz;
  x[b] = x[b] + vx * pc.dt;
  x[b + 1] = x[b + 1] + vy * pc.dt;
  x[b + 2] = x[b + 2] + vz * pc.dt;
}

void launch_integrate(float* x, float* v, const float* f, int n, const PhysConst& pc, hipStream_t st) {
  k_integrate<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, v, f, n, pc);
}

__device__ __forceinline__ float4 load3(const float* x, int i) {
  const size_t b = 3 * (size_t)i;
  return make_float4(x[b], x[b + 1], x[b + 2], 0.0f);
}

// K_ComputeDensity (SummationDensity.cu:13-36) over a CSR list.
__global__ __launch_bounds__(kBlock) void k_csr_density(const float* __restrict__ x, const int* __restrict__ off,
                                                        const int* __restrict__ el, float* __restrict__ rho, int n,
                                                        PhysConst pc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 xi = load3(x, i);
  float acc = 0.f;
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const PairGeom p = pair_geom(xi, load3(x, el[k]), pc.inv_h);
    acc += p.t * p.t * p.t;
  }
  rho[i] = acc * pc.cW;
}

void launch_csr_density(const float* x, const int* off, const int* el, float* rho, int n, const PhysConst& pc,
                        hipStream_t st) {
  k_csr_density<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, off, el, rho, n, pc);
}

__global__ __launch_bounds__(kBlock) void k_aos_to_f4(const float* __restrict__ x, float4* __restrict__ xa, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xa[i] = load3(x, i);
}

void launch_aos_to_f4(const float* x, float4* xa, int n, hipStream_t st) {
  k_aos_to_f4<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, xa, n);
}

// Density and K_ComputeLambdas (DensityPBD.cu:46-90) over a CSR list.
__global__ __launch_bounds__(kBlock) void k_csr_lambda(float4* __restrict__ xa, const int* __restrict__ off,
                                                       const int* __restrict__ el, float* __restrict__ rho, int n,
                                                       PhysConst pc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 xi = xa[i];
  float acc = 0.f, sgx = 0.f, sgy = 0.f, sgz = 0.f, sgg = 0.f;
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const PairGeom p = pair_geom(xi, xa[el[k]], pc.inv_h);
    const float t2 = p.t * p.t;
    acc += t2 * p.t;
    const float gq = p.r > pc.pair_eps ? t2 * p.rinv : 0.0f;
    const float gx = gq * p.dx, gy = gq * p.dy, gz = gq * p.dz;
    sgx += gx;
    sgy += gy;
    sgz += gz;
    sgg += gx * gx + gy * gy + gz * gz;
  }
  acc *= pc.cW;
  const float S = (pc.cG * pc.cG) * (sgg + (sgx * sgx + sgy * sgy + sgz * sgz));
  float lam = -(acc - pc.lam_rest) / (S + pc.lam_eps);
  lam = lam > 0.0f ? 0.0f : lam;
  reinterpret_cast<float*>(xa + i)[3] = lam;
  if (rho) rho[i] = acc;
}

void launch_csr_lambda(float4* xa, const int* off, const int* el, float* rho, int n, const PhysConst& pc,
                       hipStream_t st) {
  k_csr_lambda<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, off, el, rho, n, pc);
}

// K_ComputeDisplacement (DensityPBD.cu:135-179): pairwise scatter with float
// atomics, valid for any list (symmetric or not), order nondeterministic as in
// the reference.
__global__ __launch_bounds__(kBlock) void k_csr_dx_atomic(const float4* __restrict__ xa, const int* __restrict__ off,
                                                          const int* __restrict__ el, float* __restrict__ dpos, int n,
                                                          PhysConst pc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 xi = xa[i];
  const float scale = pc.corr * pc.cG;
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const int j = el[k];
    const float4 xj = xa[j];
    const PairGeom p = pair_geom(xi, xj, pc.inv_h);
    if (!(p.r > pc.pair_eps)) continue;
    const float c = scale * (xi.w + xj.w) * (p.t * p.t) * p.rinv;
    const float px = c * p.dx, py = c * p.dy, pz = c * p.dz;
    ax += px;
    ay += py;
    az += pz;
    atomicAdd(&dpos[3 * (size_t)j], -px);
    atomicAdd(&dpos[3 * (size_t)j + 1], -py);
    atomicAdd(&dpos[3 * (size_t)j + 2], -pz);
  }
  atomicAdd(&dpos[3 * (size_t)i], ax);
  atomicAdd(&dpos[3 * (size_t)i + 1], ay);
  atomicAdd(&dpos[3 * (size_t)i + 2], az);
}

void launch_csr_dx_atomic(const float4* xa, const int* off, const int* el, float* dpos, int n, const PhysConst& pc,
                          hipStream_t st) {
  k_csr_dx_atomic<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, off, el, dpos, n, pc);
}

// K_UpdatePosition (DensityPBD.cu:227-237); clears dpos for the next round
// (m_deltaPos.reset(), DensityPBD.cu:315).
__global__ __launch_bounds__(kBlock) void k_csr_apply(float4* __restrict__ xa, float* __restrict__ dpos, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = 3 * (size_t)i;
  float4 p = xa[i];
  p.x += dpos[b];
  p.y += dpos[b + 1];
  p.z += dpos[b + 2];
  xa[i] = p;
  dpos[b] = 0.f;
  dpos[b + 1] = 0.f;
  dpos[b + 2] = 0.f;
}

void launch_csr_apply(float4* xa, float* dpos, int n, hipStream_t st) {
  k_csr_apply<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, dpos, n);
}

// DP_UpdateVelocity (DensityPBD.cu:368-378) with x_old = the caller's x, then
// x := the constrained positions.
__global__ __launch_bounds__(kBlock) void k_csr_finish(const float4* __restrict__ xa, float* __restrict__ x,
                                                       float* __restrict__ v, int n, float dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = 3 * (size_t)i;
  const float4 p = xa[i];
  v[b] += (p.x - x[b]) / dt;
  v[b + 1] += (p.y - x[b + 1]) / dt;
  v[b + 2] += (p.z - x[b + 2]) / dt;
  x[b] = p.x;
  x[b + 1] = p.y;
  x[b + 2] = p.z;
}

void launch_csr_finish(const float4* xa, float* x, float* v, int n, float dt, hipStream_t st) {
  k_csr_finish<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(xa, x, v, n, dt);
}

}  // namespace pbf
