// Neighbor iteration over the cell-sorted view (shared by the forward PBD
// passes and their adjoint).
//
// A particle's 27 stencil cells form 9 runs of 3 x-consecutive cells; in cell
// order each run is one contiguous index range [a, a+len). The neighbor pass
// keeps every particle's neighbors as a wave-interleaved list (16-bit offsets
// from a per-wave base, or 32-bit indices), with bit masks over the runs or a
// re-test of the candidates as fallbacks when the reservations are full.
// Neighbor order is ascending sorted index on every path.
#pragma once
#include "pbf_device.h"

namespace pbf {

// Run r (0..8) of cell (ci,cj,ck): cells (ci-1..ci+1, cj+dj, ck+dk) with
// dk = r/3-1, dj = r%3-1 (so runs come in ascending sorted index), clipped to
// the scene's window (cells outside it hold no particle, and the window lies
// inside the grid, so this is the reference's in-grid stencil).
//
// Cells that cannot hold a neighbor of p are left out: the run is dropped when
// its row lies at least h away in y and z, and its left or right cell when that
// cell does. The bound is the distance from p to the cell's box (cell faces at
// lo + c*ds), shortened by `margin` (covers the rounding of the faces and of p's
// cell) and compared with cut2 >= the squared-distance threshold (covers the
// rounding of the squared norm), so no neighbor is ever left out and the
// neighbor lists are unchanged. A particle outside its clamped cell only makes
// the bounds on its own side smaller (clamped at 0).
__device__ __forceinline__ void run_range(const int* __restrict__ cell_start, const SceneWin& w, const GridDev& g,
                                          float4 p, float margin, float cut2, int ci, int cj, int ck, int r, int& a,
                                          int& len) {
  const int dj = (r % 3) - 1, dk = (r / 3) - 1;
  const int lj = cj + dj - w.lo[1], lk = ck + dk - w.lo[2];
  a = 0;
  len = 0;
  if (lj < 0 || lj >= w.dim[1] || lk < 0 || lk >= w.dim[2]) return;
  float by = 0.0f, bz = 0.0f;
  if (dj != 0) by = fmaxf((dj < 0 ? p.y - (g.lo[1] + (float)cj * g.ds) : (g.lo[1] + (float)(cj + 1) * g.ds) - p.y) - margin, 0.0f);
  if (dk != 0) bz = fmaxf((dk < 0 ? p.z - (g.lo[2] + (float)ck * g.ds) : (g.lo[2] + (float)(ck + 1) * g.ds) - p.z) - margin, 0.0f);
  const float base = by * by + bz * bz;
  if (base >= cut2) return;
  int i0 = max(ci - 1, w.lo[0]), i1 = min(ci + 1, w.lo[0] + w.dim[0] - 1);
  const float bl = fmaxf(p.x - (g.lo[0] + (float)ci * g.ds) - margin, 0.0f);
  const float br = fmaxf((g.lo[0] + (float)(ci + 1) * g.ds) - p.x - margin, 0.0f);
  if (base + bl * bl >= cut2) i0 = max(i0, ci);
  if (base + br * br >= cut2) i1 = min(i1, ci);
  if (i0 > i1) return;
  const int row = w.cell_base + w.dim[0] * (lj + w.dim[1] * lk) - w.lo[0];
  a = cell_start[row + i0];
  len = cell_start[row + i1 + 1] - a;
}

__device__ __forceinline__ bool in_window(const SceneWin& w, int ci, int cj, int ck) {
  return ci >= w.lo[0] && ci < w.lo[0] + w.dim[0] && cj >= w.lo[1] && cj < w.lo[1] + w.dim[1] && ck >= w.lo[2] &&
         ck < w.lo[2] + w.dim[2];
}

__device__ __forceinline__ int window_cell(const SceneWin& w, int ci, int cj, int ck) {
  return w.cell_base + (ci - w.lo[0]) + w.dim[0] * ((cj - w.lo[1]) + w.dim[1] * (ck - w.lo[2]));
}

// Bits of the candidates [j0, j0+cnt) (1 <= cnt <= 32) within h of xi. Loads
// go out eight at a time from one base address; the sorted position array
// carries kXnPad far-away entries past its end, so a group may read past the
// run (those bits are masked off) without index clamping.
__device__ __forceinline__ uint32_t test_word(const float4* __restrict__ xn, float4 xi, int j0, int cnt,
                                              float thr2) {
  uint32_t bits = 0;
  const float4* p = xn + j0;
  for (int k0 = 0; k0 < cnt; k0 += 8) {
    float4 c[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = p[k0 + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) bits |= (uint32_t)within(xi, c[u], thr2) << (k0 + u);
  }
  return cnt >= 32 ? bits : bits & ((1u << cnt) - 1u);
}

// Neighbor position load: all four components (pass B reads lambda from w),
// or xyz only (a float4 load whose w is dead lets the register allocator
// reuse that register while the load is in flight, forcing an early wait).
template <bool kW>
__device__ __forceinline__ float4 load_pos(const float4* __restrict__ xs, int j) {
  if (kW) return xs[j];
  const float* q = reinterpret_cast<const float*>(xs + j);
  return make_float4(q[0], q[1], q[2], 0.0f);
}

// List layout: a lane's entries come in groups of four that sit together
// (8 bytes as 16-bit deltas, 16 bytes as 32-bit indices), and group b of the
// wave's 64 lanes is contiguous: group b of lane l at base + (64*b + l) groups.
__device__ __forceinline__ long long list_slot(long long lb, int k, int lane, int mode) {
  return lb + ((((long long)(k >> 2) << 6) + lane) << (2 + mode)) + ((long long)(k & 3) << mode);
}

struct Group4 {
  int j[4];
};

// Indices of group b (entries 4b..4b+3); entries past cnt become s.
__device__ __forceinline__ Group4 decode16(uint2 e, int jb, int b, int cnt, int s) {
  const int k = b << 2;
  Group4 g;
  g.j[0] = k + 0 < cnt ? jb + (int)(e.x & 0xffffu) : s;
  g.j[1] = k + 1 < cnt ? jb + (int)(e.x >> 16) : s;
  g.j[2] = k + 2 < cnt ? jb + (int)(e.y & 0xffffu) : s;
  g.j[3] = k + 3 < cnt ? jb + (int)(e.y >> 16) : s;
  return g;
}
__device__ __forceinline__ Group4 decode32(int4 e, int b, int cnt, int s) {
  const int k = b << 2;
  Group4 g;
  g.j[0] = k + 0 < cnt ? e.x : s;
  g.j[1] = k + 1 < cnt ? e.y : s;
  g.j[2] = k + 2 < cnt ? e.z : s;
  g.j[3] = k + 3 < cnt ? e.w : s;
  return g;
}

// Counted loop over one lane's list, software-pipelined two groups deep:
// while the bodies of group b run, the gathered data of group b+1 and the
// indices of group b+2 are in flight (one wait per group). gather(j) loads
// what the body needs of neighbor j. The group load is issued every iteration
// (the last group again at the end) so that the wait counts stay static.
// kPadded: entries past the count were decoded to a padding particle whose
// terms are exactly zero, and the body runs on them unguarded.
template <bool kPadded, class GroupT, class LoadFn, class DecodeFn, class GatherFn, class Body>
__device__ __forceinline__ void list_loop(int cnt, LoadFn&& load_group, DecodeFn&& decode, GatherFn&& gather,
                                          Body&& body) {
  using T = decltype(gather(0));
  const int ng = (cnt + 3) >> 2;
  if (ng == 0) return;
  GroupT e = load_group(0);
  Group4 g = decode(e, 0);
  T a[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) a[u] = gather(g.j[u]);
  e = load_group(min(1, ng - 1));
  for (int b = 0; b < ng; ++b) {
    const Group4 gn = decode(e, b + 1);
    T an[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) an[u] = gather(gn.j[u]);
    e = load_group(min(b + 2, ng - 1));
    // keep the loads above in this iteration: without the fence the compiler
    // sinks them to their uses in the next one and every group waits on its
    // own gathers
    asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (kPadded || (b << 2) + u < cnt) body(g.j[u], a[u]);
    g = gn;
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = an[u];
  }
}

// Calls body(j, gather(j)) for every neighbor j of sorted particle s, in
// ascending j. The list is used when the wave has one; otherwise (reservation
// exhausted) the words come from the masks, or are recomputed from xn.
// kPadded (list path only): the body also runs on up to three entries of the
// padding particle nv.pad, whose position lies far outside every stencil; a
// body whose terms vanish with the kernel's t = 0 (and are finite) adds exact
// zeros for them.
template <bool kPadded = false, class GatherFn, class Body>
__device__ __forceinline__ void for_each_neighbor(const NbrView& nv, int s, GatherFn&& gather, Body&& body) {
  const int wave = s >> 6, lane = s & 63;
  const long long lb = nv.list_base[wave];
  if (lb >= 0) {
    const int cnt = nv.nbr_cnt[s];
    const int past = kPadded ? nv.pad : s;
    if (nv.list_mode[wave] == 0) {
      const uint2* L = reinterpret_cast<const uint2*>(nv.list + lb) + lane;
      const int jb = nv.list_jbase[wave];
      list_loop<kPadded, uint2>(
          cnt, [&](int b) { return L[(long long)b << 6]; },
          [&](uint2 e, int b) { return decode16(e, jb, b, cnt, past); }, gather, body);
    } else {
      const int4* L = reinterpret_cast<const int4*>(nv.list + lb) + lane;
      list_loop<kPadded, int4>(
          cnt, [&](int b) { return L[(long long)b << 6]; }, [&](int4 e, int b) { return decode32(e, b, cnt, past); },
          gather, body);
    }
    return;
  }
  const float4 xni = nv.xn[s];
  int ci, cj, ck;
  cell_of(nv.g, xni.x, xni.y, xni.z, ci, cj, ck);
  const SceneWin win = nv.win[nv.ssid[s]];
  const long long mb = nv.mask_base[wave];
  int wofs = 0;
  for (int r = 0; r < 9; ++r) {
    int ra, rl;
    run_range(nv.cell_start, win, nv.g, xni, nv.cull_margin, nv.cull_cut2, ci, cj, ck, r, ra, rl);
    const int nw = (rl + 31) >> 5;
    for (int w = 0; w < nw; ++w) {
      const int jb = ra + (w << 5);
      uint32_t bits = mb >= 0 ? nv.mask[mb + ((long long)(wofs + w) << 6) + lane]
                              : test_word(nv.xn, xni, jb, min(32, rl - (w << 5)), nv.thr2);
      while (bits) {
        const int j = jb + __builtin_ctz(bits);
        bits &= bits - 1u;
        body(j, gather(j));
      }
    }
    wofs += nw;
  }
}

// for_each_neighbor with the neighbor's position (kW: with its w component).
template <bool kW, bool kPadded = false, class Body>
__device__ __forceinline__ void for_each_neighbor_x(const NbrView& nv, int s, const float4* __restrict__ xs,
                                                    Body&& body) {
  for_each_neighbor<kPadded>(nv, s, [&](int j) { return load_pos<kW>(xs, j); }, body);
}

}  // namespace pbf
