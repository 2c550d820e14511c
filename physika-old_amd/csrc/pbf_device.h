// Device helpers shared by the kernel files of libpbf.so (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "pbf_internal.h"

namespace pbf {

// GridHash::getIndex3 (Framework/Topology/GridHash.h:55-62): floor((p - lo)/ds)
// with a correctly rounded division (the library is built with
// -fhip-fp32-correctly-rounded-divide-sqrt), clamped to [-2, n+1] (the set of
// valid +-1 stencil cells is unchanged; NaN maps to -2).
__device__ __forceinline__ int cell_coord(float p, float lo, float ds, int n) {
#pragma clang fp contract(off)
  float f = floorf((p - lo) / ds);
  if (!(f >= -2.0f)) return -2;
  if (f > (float)(n + 1)) return n + 1;
  return (int)f;
}

// Half-cell index along x: floor(2q) with q = (p - lo)/ds as above (doubling
// is exact, so floor(2q) >> 1 is the reference cell floor(q)), clamped to
// [-4, 2n+3] (the reference clamp in half cells); frac = 2q - floor(2q).
__device__ __forceinline__ int half_coord(float p, float lo, float ds, int n, float& frac) {
#pragma clang fp contract(off)
  const float q2 = 2.0f * ((p - lo) / ds);
  const float f = floorf(q2);
  frac = q2 - f;
  if (!(f >= -4.0f)) return -4;
  if (f > (float)(2 * n + 3)) return 2 * n + 3;
  return (int)f;
}

// Sort-grid cell of a point: half cells along x, reference cells along y and z.
__device__ __forceinline__ void cell_of(const GridDev& g, float x, float y, float z, int& fi, int& cj, int& ck,
                                        float& frac) {
  fi = half_coord(x, g.lo[0], g.ds, g.nx, frac);
  cj = cell_coord(y, g.lo[1], g.ds, g.ny);
  ck = cell_coord(z, g.lo[2], g.ds, g.nz);
}
__device__ __forceinline__ void cell_of(const GridDev& g, float x, float y, float z, int& fi, int& cj, int& ck) {
  float frac;
  cell_of(g, x, y, z, fi, cj, ck, frac);
}

// The reference's in-grid test (GridHash.h:55-62) on a sort-grid cell.
__device__ __forceinline__ bool in_grid(const GridDev& g, int fi, int cj, int ck) {
  return fi >= 0 && fi < 2 * g.nx && cj >= 0 && cj < g.ny && ck >= 0 && ck < g.nz;
}

// |x_i - x_j| < h exactly as NeighborQuery.cu:266-267 evaluates it: the squared
// norm in glm's order ((dx*dx + dy*dy) + dz*dz, no fused multiply-add) against
// thr2, the least float whose correctly rounded square root reaches h (the
// square root is monotone, so the two tests agree for every input).
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

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d));
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d));
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}
// Wave-wide minimum / maximum with every lane active, in scalar registers:
// four DPP row shifts (1, 2, 4, 8; lanes shifted in from outside the row keep
// the identity) leave each 16-lane row's result in its lane 15, and four lane
// reads combine the rows. Cheaper than the six exchanges of wave_min/wave_max.
template <int kCtrl>
__device__ __forceinline__ int dpp_row_shr(int identity, int v) {
  return __builtin_amdgcn_update_dpp(identity, v, kCtrl, 0xf, 0xf, false);
}
__device__ __forceinline__ int wave_min_full(int v) {
  v = min(v, dpp_row_shr<0x111>(INT_MAX, v));
  v = min(v, dpp_row_shr<0x112>(INT_MAX, v));
  v = min(v, dpp_row_shr<0x114>(INT_MAX, v));
  v = min(v, dpp_row_shr<0x118>(INT_MAX, v));
  return min(min(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
             min(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
}
__device__ __forceinline__ int wave_max_full(int v) {
  v = max(v, dpp_row_shr<0x111>(INT_MIN, v));
  v = max(v, dpp_row_shr<0x112>(INT_MIN, v));
  v = max(v, dpp_row_shr<0x114>(INT_MIN, v));
  v = max(v, dpp_row_shr<0x118>(INT_MIN, v));
  return max(max(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
             max(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
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

__device__ __forceinline__ float4 load3(const float* x, int i) {
  const size_t b = 3 * (size_t)i;
  return make_float4(x[b], x[b + 1], x[b + 2], 0.0f);
}

// Pair geometry with t = max(1 - r/h, 0):
//   W(r)  = 15/(pi h^3) t^3   (Dynamics/ParticleSystem/Kernel.h:36-45)
//   Grad  = -45/(pi h^3) t^2  (Kernel.h:47-57)
// r from v_rsq_f32 (within 2 ulp); a particle paired with itself gets t = 1.
struct PairGeom {
  float dx, dy, dz, r, rinv, t;
};

// Every multiply-add that is fused is written as fmaf and nothing else is
// contracted, so each copy of a pair body (list, mask and re-test traversals)
// rounds identically and all paths give bit-identical sums.
__device__ __forceinline__ PairGeom pair_geom(float4 xi, float4 xj, float inv_h) {
#pragma clang fp contract(off)
  PairGeom p;
  p.dx = xi.x - xj.x;
  p.dy = xi.y - xj.y;
  p.dz = xi.z - xj.z;
  // r2 >= 1e-30 (a particle paired with itself has r2 = 1e-30, r ~ 1e-15,
  // t = 1), so the reciprocal square root stays finite. For two distinct
  // particles the term changes r2 only when dx*dx lies exactly halfway between
  // two floats (that tie then rounds up).
  const float r2 = fmaf(p.dz, p.dz, fmaf(p.dy, p.dy, fmaf(p.dx, p.dx, 1e-30f)));
  p.rinv = __builtin_amdgcn_rsqf(r2);
  p.r = r2 * p.rinv;
  // t = max(1 - r/h, 0); r >= 0 keeps 1 - r/h <= 1, so the clamp to [0, 1]
  // (the multiply-add's clamp bit) is the same value
  p.t = __builtin_amdgcn_fmed3f(fmaf(-p.r, inv_h, 1.0f), 0.0f, 1.0f);
  return p;
}

}  // namespace pbf
