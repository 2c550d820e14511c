// Neighbor iteration over the cell-sorted view (shared by the forward PBD
// passes and their adjoint).
//
// A particle's 27 stencil cells form 9 runs of 3 x-consecutive cells; in cell
// order each run is one contiguous index range [a, a+len). The neighbor pass
// keeps every particle's neighbors as a wave-interleaved list (16-bit offsets
// from a per-wave base, stage slots, or 32-bit indices), with bit masks over
// the runs or a re-test of the candidates as fallbacks when the reservations
// are full. Neighbor order on every path: runs in visit order (run_of), then
// ascending sorted index.
#pragma once
#include <type_traits>

#include "pbf_device.h"

namespace pbf {

// The half cells along x a particle in half cell fi visits: [fi + xlo, fi + xhi].
// The reference visits its cell and both x neighbors (GridHash cells,
// NeighborQuery.cu:14-41), i.e. half cells [2c-2, 2c+3] with c = fi >> 1; a
// neighbor (closer than h) lies at most 2 half cells away unless rounding
// stretched the pair, which can only add the 3rd half cell on the side that
// stays inside the reference cells (right when fi is even, left when odd), and
// only for a particle within ext_margin of the face on that side.
__device__ __forceinline__ void stencil_x(const GridDev& g, int fi, float frac, int& xlo, int& xhi) {
  xlo = -2;
  xhi = 2;
  if (fi & 1) {
    if (frac < g.ext_margin) xlo = -3;
  } else if (frac > 1.0f - g.ext_margin) {
    xhi = 3;
  }
}

// The q-th run visited (q = 0..8) of sort cell (fi,cj,ck): half cells
// fi+xlo..fi+xhi of row (cj+dj, ck+dk) with dk = r/3-1, dj = r%3-1 for
// r = run_of(q), clipped to the scene's window (cells outside it hold no
// particle, and the window lies inside the grid, so this is the reference's
// in-grid stencil). Branch-free: a run outside the window reads entry 0 of
// cell_start (always allocated) and gets a = len = 0, so the loads of all
// nine runs issue together and a wave waits on them once, not once per run.
__device__ __forceinline__ void run_range(const int* __restrict__ cell_start, const SceneWin& w, int fi, int cj,
                                          int ck, int xlo, int xhi, int q, int& a, int& len) {
  const int r = run_of(q);
  const int lj = cj + (r % 3) - 1 - w.lo[1], lk = ck + (r / 3) - 1 - w.lo[2];
  const int i0 = max(fi + xlo, w.lo[0]), i1 = min(fi + xhi, w.lo[0] + w.dim[0] - 1);
  const bool ok = lj >= 0 && lj < w.dim[1] && lk >= 0 && lk < w.dim[2] && i0 <= i1;
  const int row = w.cell_base + w.dim[0] * (lj + w.dim[1] * lk) - w.lo[0];
  const int a0 = cell_start[ok ? row + i0 : 0];
  const int a1 = cell_start[ok ? row + i1 + 1 : 0];
  a = ok ? a0 : 0;
  len = ok ? a1 - a0 : 0;
}

__device__ __forceinline__ bool in_window(const SceneWin& w, int ci, int cj, int ck) {
  return ci >= w.lo[0] && ci < w.lo[0] + w.dim[0] && cj >= w.lo[1] && cj < w.lo[1] + w.dim[1] && ck >= w.lo[2] &&
         ck < w.lo[2] + w.dim[2];
}

__device__ __forceinline__ int window_cell(const SceneWin& w, int ci, int cj, int ck) {
  return w.cell_base + (ci - w.lo[0]) + w.dim[0] * ((cj - w.lo[1]) + w.dim[1] * (ck - w.lo[2]));
}

// bits * 2 + (s < thr2): the squared distance of one candidate shifted into a
// word (v_cmp into VCC, then an add with carry-in; NaN compares false as in the
// reference's test). thr2 is uniform and read from a scalar register.
__device__ __forceinline__ uint32_t push_bit(uint32_t bits, float s, float thr2) {
  uint32_t out;
  asm("v_cmp_lt_f32_e64 vcc, %2, %3\n\tv_addc_co_u32_e32 %0, vcc, %1, %1, vcc"
      : "=v"(out)
      : "v"(bits), "v"(s), "s"(thr2)
      : "vcc");
  return out;
}

// The squared norm of xi - c as the reference evaluates it (see within()).
__device__ __forceinline__ float sq_dist(float4 a, float4 b) {
#pragma clang fp contract(off)
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return (dx * dx + dy * dy) + dz * dz;
}

// Word of the candidates p[0..cnt) (1 <= cnt <= 32), most significant bit
// first: bit 31-k is set when candidate k lies within h of xi. Candidates are
// read eight at a time from one base address (p may point into global memory
// or LDS); the arrays are padded so that a group may read past the run's end,
// and those bits are cleared.
__device__ __forceinline__ uint32_t test_word(const float4* p, float4 xi, int cnt, float thr2) {
  uint32_t bits = 0;
  int k0 = 0;
  for (; k0 < cnt; k0 += 8) {
    float4 c[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = p[k0 + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) bits = push_bit(bits, sq_dist(xi, c[u]), thr2);
  }
  return (bits << (32 - k0)) & (~0u << (32 - cnt));
}

// test_word over the coordinate planes: kG / 4 16-byte loads per coordinate
// bring kG consecutive candidates (the loads need only 4-byte alignment), six
// loads per eight where the float4 array needs eight. kG = 16 halves the load
// round trips of a word at the cost of testing up to 15 slots past the run
// (the planes are padded like xn, by kXnPad >= 16 floats).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
template <int kG = 8>
__device__ __forceinline__ uint32_t test_word_planes(const float* cx, const float* cy, const float* cz, int j0,
                                                     float4 xi, int cnt, float thr2) {
  static_assert(kG == 8 || kG == 16, "groups of 8 or 16 candidates");
  static_assert(kXnPad >= kG, "the planes' padding covers a group past the last particle");
  uint32_t bits = 0;
  int k0 = 0;
  // one 32-bit byte offset for the three planes (scalar bases): the loads take
  // the scalar-base form and the loop advances a single register
  const char *bx = reinterpret_cast<const char*>(cx), *by = reinterpret_cast<const char*>(cy),
             *bz = reinterpret_cast<const char*>(cz);
  for (; k0 < cnt; k0 += kG) {
    const uint32_t o = (uint32_t)(j0 + k0) << 2;
    f4u qx[kG / 4], qy[kG / 4], qz[kG / 4];
#pragma unroll
    for (int g = 0; g < kG / 4; ++g) qx[g] = *reinterpret_cast<const f4u*>(bx + o + 16 * g);
#pragma unroll
    for (int g = 0; g < kG / 4; ++g) qy[g] = *reinterpret_cast<const f4u*>(by + o + 16 * g);
#pragma unroll
    for (int g = 0; g < kG / 4; ++g) qz[g] = *reinterpret_cast<const f4u*>(bz + o + 16 * g);
#pragma unroll
    for (int u = 0; u < kG; ++u)
      bits = push_bit(bits, sq_dist(xi, make_float4(qx[u >> 2][u & 3], qy[u >> 2][u & 3], qz[u >> 2][u & 3], 0.0f)),
                      thr2);
  }
  return (bits << (32 - k0)) & (~0u << (32 - cnt));
}

// Pops the first candidate (most significant set bit) of a non-zero word.
__device__ __forceinline__ int pop_first(uint32_t& bits) {
  const int c = __builtin_clz(bits);
  bits ^= 0x80000000u >> c;
  return c;
}

// Neighbor position load: all four components (pass B reads lambda from w),
// or xyz only (a float4 load whose w is dead lets the register allocator
// reuse that register while the load is in flight, forcing an early wait).
// The byte offset is 32-bit (fewer than kMaxParticles entries), so the load
// takes the scalar-base + vector-offset form.
template <bool kW>
__device__ __forceinline__ float4 load_pos_bytes(const float4* __restrict__ xs, uint32_t o) {
  const char* b = reinterpret_cast<const char*>(xs) + o;
  if (kW) return *reinterpret_cast<const float4*>(b);
  const float* q = reinterpret_cast<const float*>(b);
  return make_float4(q[0], q[1], q[2], 0.0f);
}
template <bool kW>
__device__ __forceinline__ float4 load_pos(const float4* __restrict__ xs, int j) {
  return load_pos_bytes<kW>(xs, (uint32_t)j << 4);
}

// List layout: a lane's entries come in groups of four. 32-bit indices: a
// group is 16 bytes and group g of the wave's 64 lanes is contiguous (group g
// of lane l at base + (64*g + l) * 16 B). 16-bit deltas: a group is 8 bytes and
// groups 2p and 2p+1 of a lane sit together, so one 16-byte load fetches both
// (pair p of lane l at base + (64*p + l) * 16 B). Offsets in 16-bit slots.
__device__ __forceinline__ long long list_slot(long long lb, int k, int lane, int mode) {
  const int g = k >> 2;
  if (mode == 0)
    return lb + ((((long long)(g >> 1) << 6) + lane) << 3) + ((g & 1) << 2) + (k & 3);
  return lb + ((((long long)g << 6) + lane) << 3) + ((long long)(k & 3) << 1);
}

struct Group4 {
  int j[4];
};

// Indices of group b (entries 4b..4b+3). kFull: the group is complete;
// otherwise entries past cnt become `past`.
template <bool kFull>
__device__ __forceinline__ Group4 decode16(uint2 e, int jb, int b, int cnt, int past) {
  const int k = b << 2;
  Group4 g;
  g.j[0] = jb + (int)(e.x & 0xffffu);
  g.j[1] = jb + (int)(e.x >> 16);
  g.j[2] = jb + (int)(e.y & 0xffffu);
  g.j[3] = jb + (int)(e.y >> 16);
  if (!kFull) {
#pragma unroll
    for (int u = 0; u < 4; ++u) g.j[u] = k + u < cnt ? g.j[u] : past;
  }
  return g;
}
// The stage of a staged wave (kListStaged): slot t of interval q holds sorted
// index t + D_q, D_q = A_q - P_q; interval q covers slots [P_q, P_{q+1}).
struct StageMap {
  int d0, p1, p2, inc1, inc2;  // inc_q = D_q - D_{q-1}
  __device__ __forceinline__ void load(const int* __restrict__ st) {
    const int d1 = st[1] - st[4], d2 = st[2] - st[5];
    d0 = st[0] - st[3];
    p1 = st[4];
    p2 = st[5];
    inc1 = d1 - d0;
    inc2 = d2 - d1;
  }
  // the sorted index of slot t (t below the stage's length): D of the last
  // interval starting at or before t (an empty interval starts where the next
  // one does)
  __device__ __forceinline__ int index_of(int t) const {
    return t + d0 + (t >= p1 ? inc1 : 0) + (t >= p2 ? inc2 : 0);
  }
};

// Byte offset e * 16 + base of the 16-bit list entry e held in the low
// (kHigh false) or high half of w, in one v_mad_u32_u16 (op_sel selects the
// half of the first operand); base is wave-uniform.
template <bool kHigh>
__device__ __forceinline__ int entry_bytes(uint32_t w, uint32_t base) {
  uint32_t r;
  if (kHigh)
    asm("v_mad_u32_u16 %0, %1, 16, %2 op_sel:[1,0,0,0]" : "=v"(r) : "v"(w), "s"(base));
  else
    asm("v_mad_u32_u16 %0, %1, 16, %2" : "=v"(r) : "v"(w), "s"(base));
  return (int)r;
}

// decode16 into byte offsets: entry e becomes e * 16 + base, entries past cnt
// become past (a byte offset too).
template <bool kFull>
__device__ __forceinline__ Group4 decode16_bytes(uint2 e, uint32_t base, int b, int cnt, int past) {
  const int k = b << 2;
  Group4 g;
  g.j[0] = entry_bytes<false>(e.x, base);
  g.j[1] = entry_bytes<true>(e.x, base);
  g.j[2] = entry_bytes<false>(e.y, base);
  g.j[3] = entry_bytes<true>(e.y, base);
  if (!kFull) {
#pragma unroll
    for (int u = 0; u < 4; ++u) g.j[u] = k + u < cnt ? g.j[u] : past;
  }
  return g;
}

// decode16 of stage slots into sorted indices.
template <bool kFull>
__device__ __forceinline__ Group4 decode16_staged(uint2 e, const StageMap& m, int b, int cnt, int past) {
  const int k = b << 2;
  Group4 g;
  g.j[0] = m.index_of((int)(e.x & 0xffffu));
  g.j[1] = m.index_of((int)(e.x >> 16));
  g.j[2] = m.index_of((int)(e.y & 0xffffu));
  g.j[3] = m.index_of((int)(e.y >> 16));
  if (!kFull) {
#pragma unroll
    for (int u = 0; u < 4; ++u) g.j[u] = k + u < cnt ? g.j[u] : past;
  }
  return g;
}

template <bool kFull>
__device__ __forceinline__ Group4 decode32(int4 e, int b, int cnt, int past) {
  const int k = b << 2;
  Group4 g;
  g.j[0] = e.x;
  g.j[1] = e.y;
  g.j[2] = e.z;
  g.j[3] = e.w;
  if (!kFull) {
#pragma unroll
    for (int u = 0; u < 4; ++u) g.j[u] = k + u < cnt ? g.j[u] : past;
  }
  return g;
}

// Parity of an entry's index in the neighbor order (the canonical order:
// runs in visit order, ascending sorted index). The PBD passes keep one partial
// sum per parity and add the two at the end, so that a lane pair splitting a
// list by parity and a single lane walking all of it sum identically. A group
// of four starts at an even index on every path below except the part of a
// staged list after its stage slots, which starts at the particle's stage count;
// the loops tell the body about that shift with flip (flip(true): the parities
// trade places until the next flip(true)).
using EvenEntry = std::integral_constant<int, 0>;
using OddEntry = std::integral_constant<int, 1>;

// body(j, a, parity) on the four entries of group b (guarded unless full).
template <bool kPadded, class T, class Full, class Body>
__device__ __forceinline__ void group_bodies(int b, Full, int cnt, const Group4& g, const T (&a)[4], Body& body) {
  auto one = [&](int u, auto par) {
    if (Full::value || kPadded || (b << 2) + u < cnt) body(g.j[u], a[u], par);
  };
  one(0, EvenEntry{});
  one(1, OddEntry{});
  one(2, EvenEntry{});
  one(3, OddEntry{});
}

// Counted loop over one lane's list, software-pipelined two groups deep:
// while the bodies of group b run, the gathered data of group b+1 and the
// indices of group b+2 are in flight (one wait per group). gather(j) loads
// what the body needs of neighbor j. The group load is issued every iteration
// (the last group again at the end) so that the wait counts stay static.
// Groups below nfull are complete on every lane of the wave and run without
// per-entry checks (a loop with a uniform trip count); the rest check each
// entry against the count. kPadded: entries past the count were decoded to a
// padding particle whose terms are exactly zero, and the body runs on them
// unguarded. decode(e, b, full) returns the indices of group b.
template <bool kPadded, class GroupT, class LoadFn, class DecodeFn, class GatherFn, class Body>
__device__ __forceinline__ void list_loop(int cnt, int nfull, LoadFn&& load_group, DecodeFn&& decode,
                                          GatherFn&& gather, Body&& body) {
  using T = decltype(gather(0));
  const int ng = (cnt + 3) >> 2;
  if (ng == 0) return;
  // two group buffers that trade roles every step (no register copies)
  Group4 g0, g1;
  T a0[4], a1[4];
  GroupT e = load_group(0);
  g0 = nfull > 0 ? decode(e, 0, std::true_type{}) : decode(e, 0, std::false_type{});
#pragma unroll
  for (int u = 0; u < 4; ++u) a0[u] = gather(g0.j[u]);
  e = load_group(min(1, ng - 1));
  // runs the bodies of group b (gc, ac) while group b+1 is gathered into
  // (gx, ax) and the indices of group b+2 are loaded
  auto step = [&](int b, auto full, const Group4& gc, const T(&ac)[4], Group4& gx, T(&ax)[4]) {
    gx = decode(e, b + 1, full);
#pragma unroll
    for (int u = 0; u < 4; ++u) ax[u] = gather(gx.j[u]);
    e = load_group(min(b + 2, ng - 1));
    // keep the loads above in this step: without the fence the compiler sinks
    // them to their uses in the next one and every group waits on its own
    // gathers
    asm volatile("" ::: "memory");
    group_bodies<kPadded>(b, full, cnt, gc, ac, body);
  };
  int b = 0;
  for (; b + 2 < nfull; b += 2) {
    step(b, std::true_type{}, g0, a0, g1, a1);
    step(b + 1, std::true_type{}, g1, a1, g0, a0);
  }
  for (; b < ng; b += 2) {
    step(b, std::false_type{}, g0, a0, g1, a1);
    if (b + 1 >= ng) break;
    step(b + 1, std::false_type{}, g1, a1, g0, a0);
  }
}

// list_loop for 16-bit lists: one 16-byte load brings two groups. An even
// step runs group b while group b+1 (the pair's second half) is gathered and
// the next pair is loaded; an odd step runs group b+1 while group b+2 (the
// first half of that pair) is gathered. decode(e, b, full) takes the 8 bytes of
// group b.
template <bool kPadded, class DecodeFn, class GatherFn, class Body>
__device__ __forceinline__ void list_loop_pairs(int cnt, int nfull, const uint4* __restrict__ L, DecodeFn&& decode,
                                                GatherFn&& gather, Body&& body) {
  using T = decltype(gather(0));
  const int ng = (cnt + 3) >> 2;
  if (ng == 0) return;
  const int np = (ng + 1) >> 1;
  Group4 g0, g1;
  T a0[4], a1[4];
  uint4 pc = L[0], pn;
  g0 = nfull > 0 ? decode(make_uint2(pc.x, pc.y), 0, std::true_type{})
                 : decode(make_uint2(pc.x, pc.y), 0, std::false_type{});
#pragma unroll
  for (int u = 0; u < 4; ++u) a0[u] = gather(g0.j[u]);
  auto bodies = [&](int b, auto full, const Group4& g, const T(&a)[4]) {
    group_bodies<kPadded>(b, full, cnt, g, a, body);
  };
  auto even = [&](int b, auto full) {
    g1 = decode(make_uint2(pc.z, pc.w), b + 1, full);
#pragma unroll
    for (int u = 0; u < 4; ++u) a1[u] = gather(g1.j[u]);
    pn = L[(long long)min((b >> 1) + 1, np - 1) << 6];
    asm volatile("" ::: "memory");  // keep the loads above in this step
    bodies(b, full, g0, a0);
  };
  auto odd = [&](int b, auto full) {
    g0 = decode(make_uint2(pn.x, pn.y), b + 1, full);
#pragma unroll
    for (int u = 0; u < 4; ++u) a0[u] = gather(g0.j[u]);
    asm volatile("" ::: "memory");
    bodies(b, full, g1, a1);
    pc = pn;
  };
  int b = 0;
  for (; b + 2 < nfull; b += 2) {
    even(b, std::true_type{});
    odd(b + 1, std::true_type{});
  }
  for (; b < ng; b += 2) {
    even(b, std::false_type{});
    if (b + 1 >= ng) break;
    odd(b + 1, std::false_type{});
  }
}

// The neighbor pass's per-wave words that select a wave's list path, uniform
// over the wave (scalar registers). The neighbor pass writes all three for
// every wave, so they are read unconditionally and together: no load waits on
// another (a latency chain that shows when waves are few).
struct WaveList {
  long long lb;  // list base in 16-bit slots; negative: no list
  int lm;        // list mode | groups every lane fills << 2
  int jb;        // index base of 16-bit offsets
};
__device__ __forceinline__ WaveList wave_list(const NbrView& nv, int wave) {
  return WaveList{nv.list_base[wave], nv.list_mode[wave], nv.list_jbase[wave]};
}

// Calls body(j, gather(j), parity) for every neighbor j of sorted particle s,
// in the canonical order (parity: EvenEntry or OddEntry, the entry's index mod
// 2, with flip as described at EvenEntry). The list is used when the wave has
// one; otherwise (reservation exhausted) the words come from the masks, or are
// recomputed from xn. kPadded (list path only): the body also runs on up to
// three entries of the padding particle nv.pad, whose position lies far outside
// every stencil; a body whose terms vanish with the kernel's t = 0 (and are
// finite) adds exact zeros for them. wl: the words of s's wave (wave_list).
template <bool kPadded = false, class GatherFn, class Body, class Flip>
__device__ __forceinline__ void for_each_neighbor_p(const NbrView& nv, int s, const WaveList& wl, GatherFn&& gather,
                                                    Body&& body, Flip&& flip) {
  const int wave = __builtin_amdgcn_readfirstlane(s >> 6), lane = s & 63;
  const long long lb = wl.lb;
  const int cnt = nv.nbr_cnt[s];  // issued with the wave's words (unused on the mask path)
  if (lb >= 0) {
    const int past = kPadded ? nv.pad : s;
    const int lm = wl.lm;
    const int nfull = lm >> 2;
    if ((lm & 3) == kListDelta16) {
      const uint4* L = reinterpret_cast<const uint4*>(nv.list + lb) + lane;
      const int jb = wl.jb;
      list_loop_pairs<kPadded>(
          cnt, nfull, L,
          [&](uint2 e, int b, auto full) { return decode16<decltype(full)::value>(e, jb, b, cnt, past); }, gather,
          body);
    } else if ((lm & 3) == kListStaged) {
      // stage slots of the first three runs (then pads up to kA), then offsets
      const uint4* L = reinterpret_cast<const uint4*>(nv.list + lb) + lane;
      const int* st = nv.stage + kStageInts * wave;
      StageMap m;
      m.load(st);
      const int ka = st[7], ca = nv.cnt_a[s], cb = cnt - ca;
      list_loop_pairs<kPadded>(
          ca, 0, L, [&](uint2 e, int b, auto) { return decode16_staged<false>(e, m, b, ca, past); }, gather, body);
      const int jb = wl.jb;
      flip((ca & 1) != 0);  // the offsets part starts at index ca
      list_loop_pairs<kPadded>(
          cb, nfull, L + ((long long)(ka >> 3) << 6),
          [&](uint2 e, int b, auto full) { return decode16<decltype(full)::value>(e, jb, b, cb, past); }, gather,
          body);
      flip((ca & 1) != 0);
    } else {
      const int4* L = reinterpret_cast<const int4*>(nv.list + lb) + lane;
      list_loop<kPadded, int4>(
          cnt, nfull, [&](int b) { return L[(long long)b << 6]; },
          [&](int4 e, int b, auto full) { return decode32<decltype(full)::value>(e, b, cnt, past); }, gather, body);
    }
    return;
  }
  const float4 xni = nv.xn[s];
  int ci, cj, ck, xlo, xhi;
  float frac;
  cell_of(nv.g, xni.x, xni.y, xni.z, ci, cj, ck, frac);
  stencil_x(nv.g, ci, frac, xlo, xhi);
  const SceneWin win = nv.win[nv.ssid[s]];
  const long long mb = nv.mask_base[wave];
  int wofs = 0, k = 0;
  for (int q = 0; q < 9; ++q) {
    int ra, rl;
    run_range(nv.cell_start, win, ci, cj, ck, xlo, xhi, q, ra, rl);
    const int nw = (rl + 31) >> 5;
    for (int w = 0; w < nw; ++w) {
      const int jb = ra + (w << 5);
      uint32_t bits = mb >= 0 ? nv.mask[mb + ((long long)(wofs + w) << 6) + lane]
                              : test_word(nv.xn + jb, xni, min(32, rl - (w << 5)), nv.thr2);
      while (bits) {
        const int j = jb + pop_first(bits);
        if (k & 1)
          body(j, gather(j), OddEntry{});
        else
          body(j, gather(j), EvenEntry{});
        ++k;
      }
    }
    wofs += nw;
  }
}

// for_each_neighbor_p for a body with one running sum: body(j, gather(j)) in
// the canonical order.
template <bool kPadded = false, class GatherFn, class Body>
__device__ __forceinline__ void for_each_neighbor(const NbrView& nv, int s, GatherFn&& gather, Body&& body) {
  for_each_neighbor_p<kPadded>(
      nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)), gather,
      [&](int j, const auto& a, auto) { body(j, a); }, [](bool) {});
}

// for_each_neighbor with the neighbor's position (kW: with its w component).
template <bool kW, bool kPadded = false, class Body>
__device__ __forceinline__ void for_each_neighbor_x(const NbrView& nv, int s, const float4* __restrict__ xs,
                                                    Body&& body) {
  for_each_neighbor<kPadded>(nv, s, [&](int j) { return load_pos<kW>(xs, j); }, body);
}

// A lane of a lane pair takes every other entry of a 16-bit list: halfword
// hh of each 32-bit word of a 16-byte chunk, i.e. entries 8c + 2u + hh (u =
// 0..3) of chunk c, decoded into byte addresses (entry * 16 + base; sh =
// 16 hh). Entry k = 4c + u of the lane at or past cnt (the lane's count)
// becomes past.
template <bool kFull>
__device__ __forceinline__ Group4 decode_half_bytes(uint4 e, uint32_t sh, uint32_t base, int c, int cnt, int past) {
  const int k = c << 2;
  Group4 g;
  g.j[0] = entry_bytes<false>(e.x >> sh, base);
  g.j[1] = entry_bytes<false>(e.y >> sh, base);
  g.j[2] = entry_bytes<false>(e.z >> sh, base);
  g.j[3] = entry_bytes<false>(e.w >> sh, base);
  if (!kFull) {
#pragma unroll
    for (int u = 0; u < 4; ++u) g.j[u] = k + u < cnt ? g.j[u] : past;
  }
  return g;
}

// LDS records a wave needs for for_each_neighbor_staged.
constexpr int kStageRecords = kStageCap + 1;

// LDS byte address of a pointer into a __shared__ array, and a load from one.
__device__ __forceinline__ uint32_t lds_address(const float4* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) float4*)p;
}
typedef float lds_record4 __attribute__((ext_vector_type(4)));  // 16-byte aligned
typedef float lds_record3 __attribute__((ext_vector_type(3)));  // 16-byte aligned, 12 bytes read
template <bool kW>
__device__ __forceinline__ float4 lds_load_pos(uint32_t a) {
  if (kW) {
    const lds_record4 v = *(const __attribute__((address_space(3))) lds_record4*)(uintptr_t)a;
    return make_float4(v.x, v.y, v.z, v.w);
  }
  const lds_record3 v = *(const __attribute__((address_space(3))) lds_record3*)(uintptr_t)a;
  return make_float4(v.x, v.y, v.z, 0.0f);
}

// for_each_neighbor_p with the neighbor's position (kW: with its w
// component, padded lists) for the PBD passes: a staged wave copies its stage
// from xs into its LDS region buf (kStageRecords records; coalesced loads) and
// reads the neighbors of its first three runs from there (the padding particle
// sits at slot kStagePad), so only the others are gathered through the
// vector-memory path. Other waves take for_each_neighbor_p. The entries decode
// straight into byte addresses (LDS addresses of stage records, byte offsets
// into xs), which the body receives as j (the PBD bodies ignore it). Every
// path calls the body in the same order with the same parities.
template <bool kW, class Body, class Flip>
__device__ __forceinline__ void for_each_neighbor_staged(const NbrView& nv, int s, const float4* __restrict__ xs,
                                                         float4* buf, Body&& body, Flip&& flip) {
  const int wave = __builtin_amdgcn_readfirstlane(s >> 6), lane = s & 63;
  const WaveList wl = wave_list(nv, wave);
  const long long lb = wl.lb;
  const int lm = wl.lm;
  if (lb < 0 || (lm & 3) != kListStaged) {
    for_each_neighbor_p<true>(nv, s, wl, [&](int j) { return load_pos<kW>(xs, j); }, body, flip);
    return;
  }
  const int* st = nv.stage + kStageInts * wave;
  const int ca = nv.cnt_a[s], cb = nv.nbr_cnt[s] - ca;  // loads issued before the stage copy
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int a = st[q], p0 = st[3 + q], len = st[4 + q] - p0;
    for (int t = lane; t < len; t += 64) buf[p0 + t] = xs[a + t];
  }
  if (lane == 0) buf[kStagePad] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
  // LDS operations of one wave complete in order: keep the reads below after
  // the writes above
  asm volatile("" ::: "memory");
  __builtin_amdgcn_wave_barrier();
  const int ka = st[7];
  const uint4* L = reinterpret_cast<const uint4*>(nv.list + lb) + lane;
  // stage part: ka entries on every lane (pads read the padding particle)
  const uint32_t lds = __builtin_amdgcn_readfirstlane(lds_address(buf));
  list_loop_pairs<true>(
      ka, ka >> 2, L, [&](uint2 e, int b, auto) { return decode16_bytes<true>(e, lds, b, ka, 0); },
      [&](int a) { return lds_load_pos<kW>((uint32_t)a); }, body);
  const uint32_t jb16 = (uint32_t)wl.jb << 4;
  const int past = (int)((uint32_t)nv.pad << 4);
  flip((ca & 1) != 0);  // the offsets part starts at index ca
  list_loop_pairs<true>(
      cb, lm >> 2, L + ((long long)(ka >> 3) << 6),
      [&](uint2 e, int b, auto full) { return decode16_bytes<decltype(full)::value>(e, jb16, b, cb, past); },
      [&](int o) { return load_pos_bytes<kW>(xs, (uint32_t)o); }, body);
  flip((ca & 1) != 0);
}

// for_each_neighbor_staged for a lane pair: lane h (0 or 1) of the pair that
// holds particle s calls body(x_j) for the entries of s's list whose index in
// the canonical order has parity h, in that order (the one-lane passes' even
// and odd partial sums, one per lane). Entries past a lane's count and the
// stage pads add exact zeros (the padding particle). buf: this hardware wave's
// own stage region (both waves of a list copy its stage). Returns false, with
// nothing run, when s's wave has no 16-bit or staged list; the caller then
// runs the list on one lane of the pair.
template <bool kW, class Body>
__device__ __forceinline__ bool for_each_neighbor_half(const NbrView& nv, int s, int h, const float4* __restrict__ xs,
                                                       float4* buf, Body&& body) {
  const int wave = __builtin_amdgcn_readfirstlane(s >> 6), lane = s & 63;
  const WaveList wl = wave_list(nv, wave);
  const int mode = wl.lm & 3;
  if (wl.lb < 0 || (mode != kListDelta16 && mode != kListStaged)) return false;
  const uint4* L = reinterpret_cast<const uint4*>(nv.list + wl.lb) + lane;
  const int cnt = nv.nbr_cnt[s];
  const uint32_t jb16 = (uint32_t)wl.jb << 4;
  const int past = (int)((uint32_t)nv.pad << 4);
  auto each = [&](int, const float4& xj, auto) { body(xj); };
  auto gather = [&](int o) { return load_pos_bytes<kW>(xs, (uint32_t)o); };
  // groups of the lane that every lane of the wave fills: half the list's
  // (a lane's count is at least half its particle's, rounded down)
  const int nfull = (wl.lm >> 2) >> 1;
  if (mode == kListDelta16) {
    const int ch = (cnt - h + 1) >> 1;
    const uint32_t sh = (uint32_t)h << 4;
    list_loop<true, uint4>(
        ch, nfull, [&](int c) { return L[(long long)c << 6]; },
        [&](uint4 e, int c, auto full) {
          return decode_half_bytes<decltype(full)::value>(e, sh, jb16, c, ch, past);
        },
        gather, each);
    return true;
  }
  const int* st = nv.stage + kStageInts * wave;
  const int ca = nv.cnt_a[s], cb = cnt - ca;  // loads issued before the stage copy
  const int hl = (int)(threadIdx.x & 63);    // lane of this hardware wave
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int a = st[q], p0 = st[3 + q], len = st[4 + q] - p0;
    for (int t = hl; t < len; t += 64) buf[p0 + t] = xs[a + t];
  }
  if (hl == 0) buf[kStagePad] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
  // LDS operations of one wave complete in order: keep the reads below after
  // the writes above
  asm volatile("" ::: "memory");
  __builtin_amdgcn_wave_barrier();
  // stage part: ka entries (a multiple of 8, real ones first, then pads), so
  // every chunk is full and the lane's entries are ka / 2
  const int ka = st[7];
  const uint32_t lds = __builtin_amdgcn_readfirstlane(lds_address(buf));
  const uint32_t sh1 = (uint32_t)h << 4;
  list_loop<true, uint4>(
      ka >> 1, ka >> 3, [&](int c) { return L[(long long)c << 6]; },
      [&](uint4 e, int c, auto) { return decode_half_bytes<true>(e, sh1, lds, c, ka >> 1, 0); },
      [&](int a) { return lds_load_pos<kW>((uint32_t)a); }, each);
  // offsets part: its entry p has index ca + p, so the lane takes p of parity
  // h ^ (ca & 1)
  const int hh = h ^ (ca & 1);
  const int ch = (cb - hh + 1) >> 1;
  const uint32_t sh2 = (uint32_t)hh << 4;
  const uint4* L2 = L + ((long long)(ka >> 3) << 6);
  list_loop<true, uint4>(
      ch, nfull, [&](int c) { return L2[(long long)c << 6]; },
      [&](uint4 e, int c, auto full) {
        return decode_half_bytes<decltype(full)::value>(e, sh2, jb16, c, ch, past);
      },
      gather, each);
  return true;
}

// The lane-pair decode by index: halfword hh of each word of a 16-byte chunk
// (entries 8c + 2u + hh), as sorted indices jb + entry (16-bit offsets) or
// through the stage map (stage slots); entry k = 4c + u of the lane at or past
// cnt becomes past.
template <bool kFull>
__device__ __forceinline__ Group4 decode_half_idx(uint4 e, uint32_t sh, int jb, int c, int cnt, int past) {
  const int k = c << 2;
  Group4 g;
  g.j[0] = jb + (int)((e.x >> sh) & 0xffffu);
  g.j[1] = jb + (int)((e.y >> sh) & 0xffffu);
  g.j[2] = jb + (int)((e.z >> sh) & 0xffffu);
  g.j[3] = jb + (int)((e.w >> sh) & 0xffffu);
  if (!kFull) {
#pragma unroll
    for (int u = 0; u < 4; ++u) g.j[u] = k + u < cnt ? g.j[u] : past;
  }
  return g;
}
__device__ __forceinline__ Group4 decode_half_staged(uint4 e, uint32_t sh, const StageMap& m, int c, int cnt,
                                                     int past) {
  const int k = c << 2;
  Group4 g;
  g.j[0] = m.index_of((int)((e.x >> sh) & 0xffffu));
  g.j[1] = m.index_of((int)((e.y >> sh) & 0xffffu));
  g.j[2] = m.index_of((int)((e.z >> sh) & 0xffffu));
  g.j[3] = m.index_of((int)((e.w >> sh) & 0xffffu));
#pragma unroll
  for (int u = 0; u < 4; ++u) g.j[u] = k + u < cnt ? g.j[u] : past;
  return g;
}

// for_each_neighbor for a lane pair, by index: lane h (0 or 1) of the pair
// that holds particle s calls body(j, gather(j)) for the entries of s's list
// whose index in the canonical order has parity h, in that order (entries past
// the lane's count and the stage pads are not visited). Returns false, with
// nothing run, when s's wave has no 16-bit or staged list.
template <class GatherFn, class Body>
__device__ __forceinline__ bool for_each_neighbor_half_idx(const NbrView& nv, int s, int h, GatherFn&& gather,
                                                           Body&& body) {
  const int wave = __builtin_amdgcn_readfirstlane(s >> 6), lane = s & 63;
  const WaveList wl = wave_list(nv, wave);
  const int mode = wl.lm & 3;
  if (wl.lb < 0 || (mode != kListDelta16 && mode != kListStaged)) return false;
  const uint4* L = reinterpret_cast<const uint4*>(nv.list + wl.lb) + lane;
  const int cnt = nv.nbr_cnt[s];
  const int nfull = (wl.lm >> 2) >> 1;  // as in for_each_neighbor_half
  auto each = [&](int j, const auto& a, auto) { body(j, a); };
  if (mode == kListDelta16) {
    const int ch = (cnt - h + 1) >> 1;
    const uint32_t sh = (uint32_t)h << 4;
    list_loop<false, uint4>(
        ch, nfull, [&](int c) { return L[(long long)c << 6]; },
        [&](uint4 e, int c, auto full) { return decode_half_idx<decltype(full)::value>(e, sh, wl.jb, c, ch, s); },
        gather, each);
    return true;
  }
  // stage slots of the first ca entries (the pads after them are not visited),
  // then offsets from index ca on
  const int* st = nv.stage + kStageInts * wave;
  StageMap m;
  m.load(st);
  const int ka = st[7], ca = nv.cnt_a[s], cb = cnt - ca;
  const int c1 = (ca - h + 1) >> 1;
  const uint32_t sh1 = (uint32_t)h << 4;
  list_loop<false, uint4>(
      c1, 0, [&](int c) { return L[(long long)c << 6]; },
      [&](uint4 e, int c, auto) { return decode_half_staged(e, sh1, m, c, c1, s); }, gather, each);
  const int hh = h ^ (ca & 1);
  const int ch = (cb - hh + 1) >> 1;
  const uint32_t sh2 = (uint32_t)hh << 4;
  const uint4* L2 = L + ((long long)(ka >> 3) << 6);
  list_loop<false, uint4>(
      ch, nfull, [&](int c) { return L2[(long long)c << 6]; },
      [&](uint4 e, int c, auto full) { return decode_half_idx<decltype(full)::value>(e, sh2, wl.jb, c, ch, s); },
      gather, each);
  return true;
}

}  // namespace pbf
