// Counting sort of the particles by grid cell (GridHash::construct,
// Framework/Topology/GridHash.cu:298-323), the exclusive scan it needs
// (Core/Utility/Scan.cu), the fused integrator, and the scatter back to the
// caller's order.
//
// HBM layout: caller arrays x, v are packed xyz float triples in original
// order; the sorted arrays xn, x0, xa, xb, vs are float4 in cell order (w of the
// position buffers carries lambda between passes); cell arrays cell_cnt[ncell+1],
// cell_start[ncell+2].
#include "pbf_device.h"

namespace pbf {

// ---------------------------------------------------------------------------
// Exclusive scan: blocks of 1024 threads x 4 elements scanned with wave64
// shuffles, recursion over the block sums.

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
// Hashing

__global__ void k_clear(int* cell_cnt, uint8_t* oog_flag, int ncell, unsigned long long* counters) {
  const int stride = gridDim.x * blockDim.x;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= ncell; c += stride) {
    cell_cnt[c] = 0;
    if (c < ncell) oog_flag[c] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < C_COUNT) counters[threadIdx.x] = 0;
}

void launch_clear(int* cell_cnt, uint8_t* oog_flag, int ncell, unsigned long long* counters, hipStream_t st) {
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

void launch_bin(const float* x, int n, GridDev g, int* key, int* slot, int* cell_cnt, unsigned long long* counters,
                hipStream_t st) {
  k_bin<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(x, n, g, key, slot, cell_cnt, counters);
}

// K_ConstructHashTable (GridHash.cu:187-198): ids into their cell's slots.
__global__ __launch_bounds__(kBlock) void k_scatter(const int* __restrict__ key, const int* __restrict__ slot,
                                                    const int* __restrict__ cell_start, int* __restrict__ tmp,
                                                    int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tmp[cell_start[key[i]] + slot[i]] = i;
}

void launch_scatter(const int* key, const int* slot, const int* cell_start, int* tmp, int n, hipStream_t st) {
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
    s = a;
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

void launch_reorder(const int* key, const int* cell_start, const int* tmp, const float* x, const float* v, int n,
                    int* perm, int* inv, float4* xn, float4* x0, float4* vs, const PhysConst* integrate,
                    hipStream_t st) {
  PhysConst pc{};
  if (integrate) pc = *integrate;
  k_reorder<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(key, cell_start, tmp, x, v, n, perm, inv, xn, x0, vs,
                                                          pc, integrate ? 1 : 0);
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

__global__ __launch_bounds__(kBlock) void k_csr_count(const int* __restrict__ nbr_cnt, const int* __restrict__ inv,
                                                      int n, int* __restrict__ cnt_orig) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt_orig[i] = nbr_cnt[inv[i]];
}

void launch_csr_count(const int* nbr_cnt, const int* inv, int n, int* cnt_orig, hipStream_t st) {
  k_csr_count<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nbr_cnt, inv, n, cnt_orig);
}

}  // namespace pbf
