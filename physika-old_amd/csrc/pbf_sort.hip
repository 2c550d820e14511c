// Counting sort of the particles by (scene, grid cell) (GridHash::construct,
// Framework/Topology/GridHash.cu:298-323, for a batch of independent scenes),
// the exclusive scans it needs (Core/Utility/Scan.cu), the fused integrator and
// the scatter back to the caller's order.
//
// HBM layout: caller arrays x, v are packed xyz float triples in original order
// (scenes concatenated); the sorted arrays xn, x0, xa, xb, vs are float4 in
// (scene, cell) order (w of the position buffers carries lambda between
// passes). Cells are compact: each scene owns the window of cells spanned by
// its in-grid particles plus one bucket for its out-of-grid particles.
// Between the steps of one call the state stays in the previous step's cell
// order (float4 x and v plus each particle's original index): the sort then
// reads nearly in order, and the order within a cell is still by original
// index, so the result is the same as from the caller's order.
#include <climits>

#include "pbf_device.h"

namespace pbf {

// ---------------------------------------------------------------------------
// Exclusive scan with a host-known length: blocks of 1024 threads x 4 elements
// scanned with wave64 shuffles, recursion over the block sums.

constexpr int kScanPerBlock = 4096;

// Exclusive scan of a 4096-element tile held 4 per thread; returns this
// thread's exclusive prefix of its first element and the tile total.
__device__ __forceinline__ int block_scan4(int a0, int a1, int a2, int a3, int& total) {
  __shared__ int wsum[16];
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
  const int e = (wave > 0 ? wsum[wave - 1] : 0) + incl - tsum;
  total = wsum[15];
  __syncthreads();  // wsum is reused by the next call
  return e;
}

__device__ __forceinline__ void store4(int* out, int base, int n, int e, int a0, int a1, int a2) {
  if (base < n) out[base] = e;
  e += a0;
  if (base + 1 < n) out[base + 1] = e;
  e += a1;
  if (base + 2 < n) out[base + 2] = e;
  e += a2;
  if (base + 3 < n) out[base + 3] = e;
}

__global__ __launch_bounds__(1024) void k_scan_block(const int* in, int* out, int n, int* block_sums) {
  const int base = blockIdx.x * kScanPerBlock + threadIdx.x * 4;
  const int a0 = base < n ? in[base] : 0;
  const int a1 = base + 1 < n ? in[base + 1] : 0;
  const int a2 = base + 2 < n ? in[base + 2] : 0;
  const int a3 = base + 3 < n ? in[base + 3] : 0;
  int total;
  const int e = block_scan4(a0, a1, a2, a3, total);
  store4(out, base, n, e, a0, a1, a2);
  if (threadIdx.x == 0) {
    block_sums[blockIdx.x] = total;
    if (gridDim.x == 1) out[n] = total;
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
// Exclusive scan with a device-resident length (reduce, scan the block sums,
// rescan): kScanDevBlocks blocks, each owning a contiguous chunk of tiles.

__device__ __forceinline__ void dev_chunk(int n, int& lo, int& hi) {
  const int tiles = (n + kScanPerBlock - 1) / kScanPerBlock;
  const int per = (tiles + gridDim.x - 1) / gridDim.x;
  lo = min(n, (int)blockIdx.x * per * kScanPerBlock);
  hi = min(n, lo + per * kScanPerBlock);
}

// Length min(*n_ptr, n_max); gate (may be null): the scan runs only when
// *gate != 0.
__global__ __launch_bounds__(1024) void k_scan_dev_reduce(const int* in, const int* n_ptr, int n_max, int* part,
                                                          const unsigned long long* gate) {
  __shared__ int wsum[16];
  if (gate && *gate == 0) return;
  const int n = min(*n_ptr, n_max);
  int lo, hi;
  dev_chunk(n, lo, hi);
  int acc = 0;
  for (int t = lo + threadIdx.x; t < hi; t += 1024) acc += in[t];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < 16; ++w) s += wsum[w];
    part[blockIdx.x] = s;
  }
}

// Each block first sums the block totals before its own (gridDim.x <= 1024)
// and the last block writes the grand total. The counts are zeroed once read,
// so the next sort starts from zero counts without a clearing pass (the array
// is zeroed when allocated).
__global__ __launch_bounds__(1024) void k_scan_dev_down(int* in, int* out, const int* n_ptr, int n_max,
                                                        const int* part, const unsigned long long* gate) {
  if (gate && *gate == 0) return;
  const int n = min(*n_ptr, n_max);
  int lo, hi;
  dev_chunk(n, lo, hi);
  int before;
  {
    const int i = threadIdx.x;
    const int v = i < (int)gridDim.x ? part[i] : 0;
    int total;
    const int e = block_scan4(v, 0, 0, 0, total);
    __shared__ int off;
    if (i == (int)blockIdx.x) off = e;
    if (blockIdx.x == gridDim.x - 1 && i == 0) out[n] = total;
    __syncthreads();
    before = off;
  }
  int running = before;
  for (int t0 = lo; t0 < hi; t0 += kScanPerBlock) {
    const int base = t0 + threadIdx.x * 4;
    const int a0 = base < hi ? in[base] : 0;
    const int a1 = base + 1 < hi ? in[base + 1] : 0;
    const int a2 = base + 2 < hi ? in[base + 2] : 0;
    const int a3 = base + 3 < hi ? in[base + 3] : 0;
    int total;
    const int e = block_scan4(a0, a1, a2, a3, total);
    store4(out, base, hi, running + e, a0, a1, a2);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (base + t < hi) in[base + t] = 0;
    running += total;
  }
}

// ---------------------------------------------------------------------------
// Scene windows

// Empty boxes (min INT_MAX, max INT_MIN). Run once when the boxes are
// allocated; k_scene_windows empties each box again once it has read it.
__global__ void k_scene_reset(int* minmax, int nscenes) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 6 * nscenes) minmax[t] = (t % 6) < 3 ? INT_MAX : INT_MIN;
}

void launch_scene_reset(int* minmax, int nscenes, hipStream_t st) {
  if (nscenes > 0) k_scene_reset<<<(6 * nscenes + kBlock - 1) / kBlock, kBlock, 0, st>>>(minmax, nscenes);
}

__device__ __forceinline__ int scene_of(const int* __restrict__ off, int nscenes, int i) {
  int lo = 0, hi = nscenes;  // off[lo] <= i < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Position and velocity of input particle i (caller order or cell order).
template <bool kCellOrder>
__device__ __forceinline__ float3 in_pos(const SortIn& in, int i) {
  if (kCellOrder) {
    const float4 p = in.x4[i];
    return make_float3(p.x, p.y, p.z);
  }
  return make_float3(in.x[3 * (size_t)i], in.x[3 * (size_t)i + 1], in.x[3 * (size_t)i + 2]);
}
template <bool kCellOrder>
__device__ __forceinline__ float3 in_vel(const SortIn& in, int i) {
  if (kCellOrder) {
    const float4 p = in.v4[i];
    return make_float3(p.x, p.y, p.z);
  }
  return make_float3(in.v[3 * (size_t)i], in.v[3 * (size_t)i + 1], in.v[3 * (size_t)i + 2]);
}

__device__ __forceinline__ bool grid_cell(const GridDev& g, float3 p, int& ci, int& cj, int& ck, bool& nonfinite) {
  nonfinite = !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z));
  cell_of(g, p.x, p.y, p.z, ci, cj, ck);
  return !nonfinite && in_grid(g, ci, cj, ck);
}

// Scene of every particle and the cell bounding box of each scene's in-grid
// particles. Each block covers kBoundsSpan particles (kBoundsPer per thread,
// coalesced). A block whose particles all belong to one scene writes its
// box to bpart (8 ints: scene, min i j k, max i j k, and its out-of-grid and
// non-finite counts packed as o | f << 16); blocks that straddle
// scenes mark bpart -1 and update the scenes' boxes with atomics (per wave,
// per lane where a wave straddles). Same-address atomics from every block
// would serialize in L2, so the single-scene blocks are reduced per scene by
// k_scene_reduce (by k_scene_windows for one scene).
constexpr int kBoundsBlock = 256, kBoundsPer = 4, kBoundsSpan = kBoundsBlock * kBoundsPer;

size_t bounds_part_ints(long long n) { return 8 * (size_t)((n + kBoundsSpan - 1) / kBoundsSpan + 1); }

// Wave-wide sum with every lane active (DPP row shifts, then lane reads).
__device__ __forceinline__ int wave_sum_full(int v) {
  v += dpp_row_shr<0x111>(0, v);
  v += dpp_row_shr<0x112>(0, v);
  v += dpp_row_shr<0x114>(0, v);
  v += dpp_row_shr<0x118>(0, v);
  return __builtin_amdgcn_readlane(v, 15) + __builtin_amdgcn_readlane(v, 31) + __builtin_amdgcn_readlane(v, 47) +
         __builtin_amdgcn_readlane(v, 63);
}

template <bool kCellOrder>
__global__ __launch_bounds__(kBoundsBlock) void k_scene_bounds(SortIn in, int n, GridDev g,
                                                               const int* __restrict__ off, int nscenes,
                                                               int* __restrict__ sid, int* __restrict__ minmax,
                                                               int* __restrict__ bpart, int* __restrict__ oog_flag) {
  __shared__ int red[kBoundsBlock / 64][7];
  const int base = blockIdx.x * kBoundsSpan;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int sq[kBoundsPer], cq[kBoundsPer][3];
  bool inq[kBoundsPer];
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  int counts = 0;  // out-of-grid | non-finite << 16
#pragma unroll
  for (int q = 0; q < kBoundsPer; ++q) {
    const int i = base + q * kBoundsBlock + (int)threadIdx.x;
    int sc = -1, c[3] = {0, 0, 0};
    bool nonfinite = false, inside = false;
    if (i < n) {
      sc = scene_of(off, nscenes, i);
      sid[i] = sc;
      inside = grid_cell(g, in_pos<kCellOrder>(in, i), c[0], c[1], c[2], nonfinite);
      counts += (inside ? 0 : 1) + (nonfinite ? 1 << 16 : 0);
      if (!inside) oog_flag[kCellOrder ? in.id[i] : i] = 1;
    }
    sq[q] = sc;
    inq[q] = inside;
    for (int a = 0; a < 3; ++a) {
      cq[q][a] = c[a];
      if (inside) {
        lo[a] = min(lo[a], c[a]);
        hi[a] = max(hi[a], c[a]);
      }
    }
  }
  // every lane of the block is active (no early exit)
  for (int a = 0; a < 3; ++a) {
    lo[a] = wave_min_full(lo[a]);
    hi[a] = wave_max_full(hi[a]);
  }
  counts = wave_sum_full(counts);
  if (lane == 0) {
    for (int a = 0; a < 3; ++a) {
      red[wave][a] = lo[a];
      red[wave][3 + a] = hi[a];
    }
    red[wave][6] = counts;
  }
  __syncthreads();
  // one scene when the block's first and last particles share it (uniform)
  const int last = min(n, base + kBoundsSpan) - 1;
  const bool one_scene = scene_of(off, nscenes, base) == scene_of(off, nscenes, last);
  if (threadIdx.x == 0) {
    int r[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    int tot = 0;
    for (int w = 0; w < kBoundsBlock / 64; ++w) {
      for (int a = 0; a < 3; ++a) r[a] = min(r[a], red[w][a]);
      for (int a = 3; a < 6; ++a) r[a] = max(r[a], red[w][a]);
      tot += red[w][6];
    }
    int* bp = bpart + 8 * (size_t)blockIdx.x;
    bp[0] = one_scene ? scene_of(off, nscenes, base) : -1;
    for (int a = 0; a < 6; ++a) bp[1 + a] = r[a];
    bp[7] = tot;
  }
  if (one_scene) return;  // block-uniform
  // a block that straddles scenes: per particle slot q, per wave where the
  // wave's particles share a scene, else per lane
#pragma unroll
  for (int q = 0; q < kBoundsPer; ++q) {
    const int i = base + q * kBoundsBlock + (int)threadIdx.x;
    const bool valid = i < n;
    const int s0 = __shfl(sq[q], 0);  // lane 0 is valid whenever any lane is
    if (__all(!valid || sq[q] == s0)) {
      int m[6];
      for (int a = 0; a < 3; ++a) {
        m[a] = wave_min_full(inq[q] ? cq[q][a] : INT_MAX);
        m[3 + a] = wave_max_full(inq[q] ? cq[q][a] : INT_MIN);
      }
      if (lane == 0 && s0 >= 0 && m[3] >= m[0]) {
        int* box = minmax + 6 * s0;
        for (int a = 0; a < 3; ++a) atomicMin(box + a, m[a]);
        for (int a = 3; a < 6; ++a) atomicMax(box + a, m[a]);
      }
    } else if (inq[q]) {
      int* box = minmax + 6 * sq[q];
      for (int a = 0; a < 3; ++a) atomicMin(box + a, cq[q][a]);
      for (int a = 0; a < 3; ++a) atomicMax(box + 3 + a, cq[q][a]);
    }
  }
}

// One block per scene: reduce the boxes of the single-scene blocks of the
// scene's particle range and merge them into the scene's box. Each entry is
// read as two 16-byte loads, four entries per thread in flight.
constexpr int kReduceBlock = 1024;

__device__ __forceinline__ void merge_box(int r[6], const int* __restrict__ bpart, int b, int s) {
  const int4 u = reinterpret_cast<const int4*>(bpart)[2 * (size_t)b];
  const int4 w = reinterpret_cast<const int4*>(bpart)[2 * (size_t)b + 1];
  if (u.x != s) return;
  r[0] = min(r[0], u.y);
  r[1] = min(r[1], u.z);
  r[2] = min(r[2], u.w);
  r[3] = max(r[3], w.x);
  r[4] = max(r[4], w.y);
  r[5] = max(r[5], w.z);
}

__global__ __launch_bounds__(kReduceBlock) void k_scene_reduce(const int* __restrict__ off, int nscenes,
                                                               const int* __restrict__ bpart,
                                                               int* __restrict__ minmax) {
  __shared__ int red[kReduceBlock / 64][6];
  const int s = blockIdx.x;
  const int b0 = off[s] / kBoundsSpan;
  const int b1 = off[s + 1] > off[s] ? (off[s + 1] - 1) / kBoundsSpan + 1 : b0;
  int r[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
  int b = b0 + (int)threadIdx.x;
  for (; b + 3 * kReduceBlock < b1; b += 4 * kReduceBlock) {
#pragma unroll
    for (int q = 0; q < 4; ++q) merge_box(r, bpart, b + q * kReduceBlock, s);
  }
  for (; b < b1; b += kReduceBlock) merge_box(r, bpart, b, s);
  // every lane of the block is active
  for (int a = 0; a < 3; ++a) r[a] = wave_min_full(r[a]);
  for (int a = 3; a < 6; ++a) r[a] = wave_max_full(r[a]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int a = 0; a < 6; ++a) red[wave][a] = r[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kReduceBlock / 64; ++w) {
      for (int a = 0; a < 3; ++a) r[a] = min(r[a], red[w][a]);
      for (int a = 3; a < 6; ++a) r[a] = max(r[a], red[w][a]);
    }
    if (r[3] >= r[0]) {
      int* m = minmax + 6 * s;
      atomicMin(m + 0, r[0]);
      atomicMin(m + 1, r[1]);
      atomicMin(m + 2, r[2]);
      atomicMax(m + 3, r[3]);
      atomicMax(m + 4, r[4]);
      atomicMax(m + 5, r[5]);
    }
  }
}

// Windows and compact cell bases of all scenes (one block); each scene's box
// is emptied again for the next sort. With one scene the block also reduces
// the bounds blocks' boxes (k_scene_reduce is not launched). The block also
// sets the search's counters: the out-of-grid and non-finite totals of the
// bounds pass, the cell total, and zero for the counters the neighbor pass
// accumulates.
__global__ __launch_bounds__(1024) void k_scene_windows(int* __restrict__ minmax, int nscenes,
                                                        SceneWin* __restrict__ win, int* __restrict__ total_cells,
                                                        const int* __restrict__ bpart, int nbounds,
                                                        unsigned long long* __restrict__ counters) {
  __shared__ unsigned long long red[2][16];
  __shared__ int box[16][6];
  unsigned long long oog = 0, nonfinite = 0;
  const bool one = nscenes == 1;
  int r[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
#pragma unroll 4
  for (int b = threadIdx.x; b < nbounds; b += blockDim.x) {
    const unsigned v = (unsigned)bpart[8 * (size_t)b + 7];
    oog += v & 0xffffu;
    nonfinite += v >> 16;
    if (one) merge_box(r, bpart, b, 0);
  }
  if (one) {  // every lane of the block is active
    for (int a = 0; a < 3; ++a) r[a] = wave_min_full(r[a]);
    for (int a = 3; a < 6; ++a) r[a] = wave_max_full(r[a]);
    if ((threadIdx.x & 63) == 0)
      for (int a = 0; a < 6; ++a) box[threadIdx.x >> 6][a] = r[a];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
        for (int a = 0; a < 3; ++a) r[a] = min(r[a], box[w][a]);
        for (int a = 3; a < 6; ++a) r[a] = max(r[a], box[w][a]);
      }
      for (int a = 0; a < 3; ++a) minmax[a] = min(minmax[a], r[a]);
      for (int a = 3; a < 6; ++a) minmax[a] = max(minmax[a], r[a]);
    }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    oog += __shfl_xor(oog, d);
    nonfinite += __shfl_xor(nonfinite, d);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = oog;
    red[1][threadIdx.x >> 6] = nonfinite;
  }
  int running = 0;
  for (int s0 = 0; s0 < nscenes; s0 += 1024) {
    const int s = s0 + threadIdx.x;
    SceneWin w{};
    int cells = 0;
    if (s < nscenes) {
      int* m = minmax + 6 * s;
      if (m[3] >= m[0]) {
        for (int a = 0; a < 3; ++a) {
          w.lo[a] = m[a];
          w.dim[a] = m[a + 3] - m[a] + 1;
        }
        w.ncell = w.dim[0] * w.dim[1] * w.dim[2];
      } else {
        for (int a = 0; a < 3; ++a) {
          w.lo[a] = 0;
          w.dim[a] = 0;
        }
        w.ncell = 0;
      }
      cells = w.ncell + 1;  // + out-of-grid bucket
      for (int a = 0; a < 6; ++a) m[a] = a < 3 ? INT_MAX : INT_MIN;
    }
    int total;
    const int e = block_scan4(cells, 0, 0, 0, total);
    if (s < nscenes) {
      w.cell_base = running + e;
      win[s] = w;
    }
    running += total;
  }
  if (threadIdx.x == 0) {
    *total_cells = running;
    unsigned long long o = 0, f = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      o += red[0][w];
      f += red[1][w];
    }
    counters[C_CELLS] = (unsigned long long)running;
    counters[C_OOG] = o;
    counters[C_NONFINITE] = f;
    counters[C_MASK_WORDS] = 0;
    counters[C_FALLBACK] = 0;
  }
}

// K_CalculateParticleNumber (GridHash.cu:78-87): key = compact cell or the
// scene's out-of-grid bucket; slot = arrival order inside a cell, and inside a
// bucket the rank of the particle's original index among the scene's
// out-of-grid particles (oog_pre, the scan of the bounds pass's flags), so a
// bucket is in original-index order at any size without atomics. Block 0 sets
// each bucket's count.
template <bool kCellOrder>
__global__ __launch_bounds__(kBlock) void k_bin(SortIn in, int n, GridDev g,
                                                const int* __restrict__ sid, const SceneWin* __restrict__ win,
                                                int* __restrict__ key, int* __restrict__ slot,
                                                int* __restrict__ cell_cnt, const int* __restrict__ scene_off,
                                                int nscenes, const int* __restrict__ oog_pre,
                                                const unsigned long long* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && counters[C_OOG] != 0) {
    for (int s = threadIdx.x; s < nscenes; s += blockDim.x) {
      const SceneWin w = win[s];
      cell_cnt[w.cell_base + w.ncell] = oog_pre[min(scene_off[s + 1], n)] - oog_pre[min(scene_off[s], n)];
    }
  }
  if (i >= n) return;
  int ci, cj, ck;
  bool nonfinite;
  const bool inside = grid_cell(g, in_pos<kCellOrder>(in, i), ci, cj, ck, nonfinite);
  const SceneWin w = win[sid[i]];
  int k = w.cell_base + w.ncell;
  if (inside) k = w.cell_base + (ci - w.lo[0]) + w.dim[0] * ((cj - w.lo[1]) + w.dim[1] * (ck - w.lo[2]));
  key[i] = k;
  const int lane = threadIdx.x & 63;
  int sl = 0;
  if (kCellOrder) {
    // Input in the previous step's cell order: in-grid lanes of one cell are
    // mostly adjacent. One atomic per run of equal keys, at the run's first
    // lane (a key split over several runs of the wave counts once per run);
    // out-of-grid lanes end runs and count below. Active lanes are the low
    // lanes of the wave.
    const int prev = __shfl_up(k, 1);
    const bool cut = !inside || lane == 0 || prev != k;
    const unsigned long long cuts = __ballot(cut), act = __ballot(true);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
    const int hl = 63 - __clzll(cuts & upto);               // first lane of this run
    const unsigned long long after = cuts & ~upto;
    const int end = after ? __ffsll((long long)after) - 1 : 64 - __clzll(act);
    int base = 0;
    if (inside && cut) base = atomicAdd(&cell_cnt[k], end - lane);
    const int b = __shfl(base, hl);
    if (inside) sl = b + (lane - hl);
  } else if (inside) {
    sl = atomicAdd(&cell_cnt[k], 1);
  }
  if (!inside) sl = oog_pre[kCellOrder ? in.id[i] : i] - oog_pre[min(scene_off[sid[i]], n)];
  slot[i] = sl;
}

// K_ConstructHashTable (GridHash.cu:187-198): ids into their cell's slots.
// Input in cell order: tid receives the original index beside each input index.
template <bool kCellOrder>
__global__ __launch_bounds__(kBlock) void k_scatter(const int* __restrict__ key, const int* __restrict__ slot,
                                                    const int* __restrict__ cell_start, int* __restrict__ tmp,
                                                    const int* __restrict__ id, int* __restrict__ tid, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = cell_start[key[i]] + slot[i];
  tmp[p] = i;
  if (kCellOrder) tid[p] = id[i];
}

// Final position inside the cell = rank of the particle id among the cell's
// ids (deterministic order, independent of atomic arrival). Cells holding more
// than kRankMax particles keep arrival order (still a valid cell sort); a
// scene's out-of-grid bucket is in original-index order already (k_bin). One
// thread per slot of the scattered order: slot t's particle lands inside its
// own cell's range, so the sorted arrays are written nearly in order and only
// the inputs are read scattered.
// The integrator (ParticleIntegrator.cu:60-70,103-112, force zeroed by begin)
// is fused here: x^n goes to xn (the neighbor sets are built on it,
// PositionBasedFluidModel.cpp:93-94), the integrated state to x0 / vs.
constexpr int kRankMax = 1024;

template <bool kCellOrder>
__global__ __launch_bounds__(kBlock) void k_reorder(const int* __restrict__ key, const int* __restrict__ cell_start,
                                                    const int* __restrict__ tmp, const int* __restrict__ tid,
                                                    const int* __restrict__ sid, SortIn in, int n,
                                                    int* __restrict__ perm, int* __restrict__ ssid,
                                                    float4* __restrict__ xn,
                                                    float4* __restrict__ x0, float4* __restrict__ vs, PhysConst pc,
                                                    int integrate, float* __restrict__ xc, int xc_stride,
                                                    int* __restrict__ oog_rcnt, const SceneWin* __restrict__ win) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) x0[n] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);  // the padding particle
  if (t >= n) return;
  const int i = tmp[t];  // input index
  const int id = kCellOrder ? tid[t] : i;  // original index
  const int k = key[i];
  const int a = cell_start[k], b = cell_start[k + 1];
  int s = t;
  const SceneWin w = win[sid[i]];
  if (k != w.cell_base + w.ncell && b - a <= kRankMax) {
    int rank = 0;
    const int* ids = kCellOrder ? tid : tmp;
    int u = a;
    for (; u + 4 <= b; u += 4) {  // four loads in flight
      const int c0 = ids[u], c1 = ids[u + 1], c2 = ids[u + 2], c3 = ids[u + 3];
      rank += (c0 < id) + (c1 < id) + (c2 < id) + (c3 < id);
    }
    for (; u < b; ++u) rank += ids[u] < id;
    s = a + rank;
  }
  perm[s] = id;
  ssid[s] = sid[i];
  oog_rcnt[s] = 0;
  const float3 p = in_pos<kCellOrder>(in, i);
  const float px = p.x, py = p.y, pz = p.z;
  xn[s] = make_float4(px, py, pz, 0.0f);
  xc[s] = px;
  xc[xc_stride + s] = py;
  xc[2 * (size_t)xc_stride + s] = pz;
  if (integrate) {
    const float3 v = in_vel<kCellOrder>(in, i);
    const float vx = v.x + pc.gdt[0];
    const float vy = v.y + pc.gdt[1];
    const float vz = v.z + pc.gdt[2];
    x0[s] = make_float4(px + vx * pc.dt, py + vy * pc.dt, pz + vz * pc.dt, 0.0f);
    vs[s] = make_float4(vx, vy, vz, 0.0f);
  } else {
    x0[s] = make_float4(px, py, pz, 0.0f);
  }
}

template <bool kCellOrder>
void sort_in(const SortBufs& sb, const GridDev& g, const SortIn& in, int n, float4* xn, float4* x0, float4* vs,
             const PhysConst* integrate, unsigned long long* counters, hipStream_t st) {
  const int nb = (n + kBlock - 1) / kBlock;
  const int nbounds = (n + kBoundsSpan - 1) / kBoundsSpan;
  k_scene_bounds<kCellOrder><<<nbounds, kBoundsBlock, 0, st>>>(in, n, g, sb.scene_off, sb.nscenes, sb.sid,
                                                               sb.scene_minmax, sb.bpart, sb.oog_flag);
  if (sb.nscenes > 1)
    k_scene_reduce<<<sb.nscenes, kReduceBlock, 0, st>>>(sb.scene_off, sb.nscenes, sb.bpart, sb.scene_minmax);
  k_scene_windows<<<1, 1024, 0, st>>>(sb.scene_minmax, sb.nscenes, sb.win, sb.total_cells, sb.bpart, nbounds,
                                      counters);
  // out-of-grid ranks by original index (only when the bounds pass found any;
  // the last scene offset is n)
  const unsigned long long* any_oog = counters + C_OOG;
  const int* n_dev = sb.scene_off + sb.nscenes;
  k_scan_dev_reduce<<<kScanDevBlocks, 1024, 0, st>>>(sb.oog_flag, n_dev, n, sb.scan_part, any_oog);
  k_scan_dev_down<<<kScanDevBlocks, 1024, 0, st>>>(sb.oog_flag, sb.oog_pre, n_dev, n, sb.scan_part, any_oog);
  k_bin<kCellOrder><<<nb, kBlock, 0, st>>>(in, n, g, sb.sid, sb.win, sb.key, sb.slot, sb.cell_cnt, sb.scene_off,
                                           sb.nscenes, sb.oog_pre, counters);
  k_scan_dev_reduce<<<kScanDevBlocks, 1024, 0, st>>>(sb.cell_cnt, sb.total_cells, INT_MAX, sb.scan_part, nullptr);
  k_scan_dev_down<<<kScanDevBlocks, 1024, 0, st>>>(sb.cell_cnt, sb.cell_start, sb.total_cells, INT_MAX,
                                                   sb.scan_part, nullptr);
  k_scatter<kCellOrder><<<nb, kBlock, 0, st>>>(sb.key, sb.slot, sb.cell_start, sb.tmp, in.id, sb.tid, n);
  PhysConst pc{};
  if (integrate) pc = *integrate;
  k_reorder<kCellOrder><<<nb, kBlock, 0, st>>>(sb.key, sb.cell_start, sb.tmp, sb.tid, sb.sid, in, n, sb.perm,
                                               sb.ssid, xn, x0, vs, pc, integrate ? 1 : 0, sb.xc, sb.xc_stride,
                                               sb.oog_rcnt, sb.win);
}

void launch_sort(const SortBufs& sb, const GridDev& g, const SortIn& in, int n, float4* xn, float4* x0, float4* vs,
                 const PhysConst* integrate, unsigned long long* counters, hipStream_t st) {
  if (in.x4)
    sort_in<true>(sb, g, in, n, xn, x0, vs, integrate, counters, st);
  else
    sort_in<false>(sb, g, in, n, xn, x0, vs, integrate, counters, st);
}

// Copy of cell_start[0..total] (the device-side cell count of the last sort).
__global__ void k_copy_cells(const int* __restrict__ src, int* __restrict__ dst, const int* __restrict__ total_cells) {
  const int total = *total_cells + 1;
  const int stride = gridDim.x * blockDim.x;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < total; c += stride) dst[c] = src[c];
}

void launch_copy_cells(const int* src, int* dst, const int* total_cells, long long cap, hipStream_t st) {
  const int blocks = (int)std::min<long long>((cap + 2 + kBlock - 1) / kBlock, 1024);
  k_copy_cells<<<blocks, kBlock, 0, st>>>(src, dst, total_cells);
}

__global__ void k_set_offsets(int* off, int n) {
  if (threadIdx.x == 0) {
    off[0] = 0;
    off[1] = n;
  }
}

void launch_set_offsets(int* off, int n, hipStream_t st) { k_set_offsets<<<1, 64, 0, st>>>(off, n); }

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

// inv[perm[s]] = s (the CSR export needs the sorted slot of every particle).
__global__ __launch_bounds__(kBlock) void k_invert(const int* __restrict__ perm, int* __restrict__ inv, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) inv[perm[s]] = s;
}

void launch_invert(const int* perm, int* inv, int n, hipStream_t st) {
  if (n > 0) k_invert<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(perm, inv, n);
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
