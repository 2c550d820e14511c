// Neighbor lists and the PBD passes over them (NeighborQuery::compute,
// SummationDensity, DensityPBD::takeOneIteration).
//
// A particle's 27 stencil cells form 9 runs of x-consecutive sort cells; in
// cell order each run is one contiguous index range [a, a+len). The neighbor
// pass tests every candidate of the runs (the reference's set: every j in a
// valid stencil cell with sqrtf(|x_i - x_j|^2) < h, NeighborQuery.cu:245-276),
// keeps the result as bit masks (scratch) and decodes them into the wave's
// neighbor list (the encodings of pbf_internal.h, wave-interleaved in groups of
// four). The ten PBD passes of a step then read the list with coalesced loads
// and take each neighbor's position from LDS (staged lists) or a gather.
// Neighbor order is the run visit order, ascending index inside a run, on
// every path.
#include <climits>

#include <algorithm>

#include <hip/hip_ext.h>

#include "pbf_nbr.h"

namespace pbf {

// ---------------------------------------------------------------------------
// Neighbor pass: K_CalNeighborSize + K_GetNeighborElements
// (NeighborQuery.cu:245-276,337-367) in one traversal of the candidates.

// Mask words per lane a wave keeps in LDS during the neighbor pass, only those
// holding a neighbor (8 KB per wave: 20 waves per CU; 24 words, 13 waves per
// CU, ran 15% slower). A lane with more decodes its list from the candidates
// tested again. Waves with more than kNbMaxWords words per lane keep them all
// in the global mask reservation.
constexpr int kNbLdsWords = 16;
constexpr int kNbMaxWords = 28;

// Appends neighbor j to a lane's group of four and stores complete groups:
// 16-bit offsets from the wave base (mode 0: a 64-bit shift register, entry 0
// in the low half of e0) or 32-bit indices (mode 1).
template <int kMode>
struct ListWriter {
  uint16_t* list;
  long long lb;
  int lane, k = 0;
  uint32_t e0 = 0u, e1 = 0u, e2 = 0u, e3 = 0u;

  __device__ __forceinline__ void store(int group) {
    const long long slot = list_slot(lb, group << 2, lane, kMode);
    if (kMode == 0)
      *reinterpret_cast<uint2*>(list + slot) = make_uint2(e0, e1);
    else
      *reinterpret_cast<uint4*>(list + slot) = make_uint4(e0, e1, e2, e3);
  }
  __device__ __forceinline__ void push(uint32_t d) {
    if (kMode == 0) {
      e0 = __builtin_amdgcn_alignbit(e1, e0, 16);
      e1 = __builtin_amdgcn_alignbit(d, e1, 16);
    } else {
      e0 = e1;
      e1 = e2;
      e2 = e3;
      e3 = d;
    }
    ++k;
    if ((k & 3) == 0) store((k >> 2) - 1);
  }
  // Pads (kStagePad) up to entry `target` (a multiple of 8, 16-bit lists):
  // the open group is completed by pushes, the rest stored as whole groups.
  __device__ __forceinline__ void pad_to(int target) {
    while ((k & 3) != 0 && k < target) push(kStagePad);
    const uint32_t pp = kStagePad | (kStagePad << 16);
    for (int g = k >> 2; g < (target >> 2); ++g)
      *reinterpret_cast<uint2*>(list + list_slot(lb, g << 2, lane, 0)) = make_uint2(pp, pp);
    k = max(k, target);
  }
  // the last, partial group: entries moved to its start (the slots past the
  // count hold stale offsets of this wave, which no pass reads)
  __device__ __forceinline__ void finish() {
    const int m = k & 3;
    if (m == 0) return;
    if (kMode == 0) {
      const int sh = 16 * (4 - m);
      const unsigned long long v = (((unsigned long long)e1 << 32) | e0) >> sh;
      e0 = (uint32_t)v;
      e1 = (uint32_t)(v >> 32);
    } else {
      for (int t = m; t < 4; ++t) {
        e0 = e1;
        e1 = e2;
        e2 = e3;
      }
    }
    store(k >> 2);
  }
};

// Phase 2 over LDS words: each word holding a neighbor sits with the list
// value of its first candidate ({bits, base}; words without one are not
// stored); a flat loop over the lane's entries pops the next set bit and moves
// to the next word when one runs empty, so the wave runs as many iterations as
// its largest count. The base of a word: 32-bit j (kMode
// 1), or two 16-bit values: j - list_jbase (low half) and the stage slot of j
// (high half, the first three runs of a staged wave). The first n_slot entries
// are written as stage slots, followed by pads up to pad_to entries; the rest
// as offsets.
template <int kMode>
__device__ __forceinline__ void write_list_lds(uint16_t* list, long long lb, int lane, const uint2* lw, int cnt,
                                               int n_slot, int pad_to) {
  ListWriter<kMode> wr{list, lb, lane};
  uint32_t bits = 0u, ey = 0u;
  int wi = 0;
  // the next word is read one word ahead, so that a refill does not wait on LDS
  uint2 nxt = lw[0];
  auto run = [&](int k0, int k1, bool hi) {
    for (int k = k0; k < k1; ++k) {
      if (bits == 0u) {  // every stored word holds a set bit
        bits = nxt.x;
        ey = nxt.y;
        ++wi;
        nxt = lw[min(wi, kNbLdsWords - 1) << 6];
      }
      const uint32_t jb = kMode == 1 ? ey : (hi ? ey >> 16 : ey & 0xffffu);
      wr.push(jb + (uint32_t)pop_first(bits));
    }
  };
  run(0, n_slot, true);
  if (kMode == 0) wr.pad_to(pad_to);
  run(n_slot, cnt, false);
  wr.finish();
}

// Phase 2 run by run over words from word_of(q, w, wofs + w) (the global mask
// reservation, or the candidates tested again): the first n_slot entries (all
// in runs q < 3) are written as stage slots j - sdel[q], followed by pads up to
// pad_to entries; the rest as j - jbase.
template <int kMode, class WordFn>
__device__ __forceinline__ void write_list_runs(uint16_t* list, long long lb, int lane, const int (&ra)[9],
                                                const int (&rl)[9], WordFn&& word_of, int jbase,
                                                const int (&sdel)[3], int n_slot, int pad_to) {
  ListWriter<kMode> wr{list, lb, lane};
  int wofs = 0;
  bool padded = pad_to == 0;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int nw = (rl[q] + 31) >> 5;
    for (int w = 0; w < nw; ++w) {
      uint32_t bits = word_of(q, w, wofs + w);
      const int j0 = ra[q] + (w << 5);
      while (bits) {
        const int j = j0 + pop_first(bits);
        if (wr.k < n_slot) {
          wr.push((uint32_t)(j - sdel[q < 3 ? q : 0]));
        } else {
          if (!padded) {
            if (kMode == 0) wr.pad_to(pad_to);
            padded = true;
          }
          wr.push((uint32_t)(j - jbase));
        }
      }
    }
    wofs += nw;
  }
  if (!padded && kMode == 0) wr.pad_to(pad_to);
  wr.finish();
}

// One wave per block: phase 1 tests the candidates of the 9 runs (a lane's
// words in LDS, or in the global mask reservation when they do not fit),
// phase 2 writes the list into the wave's fixed region.
//
// Stage: the union over the lanes of each of the first three runs (the rows of
// the particle's own z-plane) is one index interval [A_q, E_q); a full wave
// whose three intervals hold at most kStageCap particles together (and contain
// every lane's runs, checked) writes those neighbors as stage slots, so that
// the PBD passes read them from LDS.
template <int kG>
__global__ __launch_bounds__(64) void k_neighbors(NbrView nv, NbrOut out, unsigned long long* __restrict__ counters) {
  __shared__ uint2 lword[kNbLdsWords * 64];
  const int wave = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const int s = (wave << 6) + lane;
  const bool active = s < nv.n;
  float4 xi = make_float4(0.f, 0.f, 0.f, 0.f);
  int ra[9], rl[9];
  int words = 0, cand = 0, jlo = INT_MAX, jhi = -1;
  bool oog_lane = false;
  if (active) {
    // the scene window's loads first: they do not wait on the position
    const SceneWin win = nv.win[nv.ssid[s]];
    xi = nv.xn[s];
    int ci, cj, ck, xlo, xhi;
    float frac;
    cell_of(nv.g, xi.x, xi.y, xi.z, ci, cj, ck, frac);
    stencil_x(nv.g, ci, frac, xlo, xhi);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      run_range(nv.cell_start, win, ci, cj, ck, xlo, xhi, q, ra[q], rl[q]);
      words += (rl[q] + 31) >> 5;
      cand += rl[q];
      if (rl[q] > 0) {
        jlo = min(jlo, ra[q]);
        jhi = max(jhi, ra[q] + rl[q] - 1);
      }
    }
    // An out-of-grid particle is not hashed (GridHash.cu:85,194) yet queries,
    // so each in-grid j it lists receives -dp_kj from it (DensityPBD.cu:162)
    // without listing it back (registered below for pass B). The window holds
    // every in-grid cell of the scene.
    oog_lane = !in_window(win, ci, cj, ck);
  } else {
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      ra[q] = 0;
      rl[q] = 0;
    }
  }
  const int wmax = wave_max(words);
  const int wjlo = wave_min(jlo), wjhi = wave_max(jhi);
  const int mode = (wjhi >= wjlo && wjhi - wjlo > 65535) ? kListIndex32 : kListDelta16;
  const int jbase = mode == kListDelta16 ? wjlo : 0;  // list offsets are j - jbase
  // the stage: interval q starts at index sa[q] and at slot sp[q] (its length
  // is sp[q + 1] - sp[q]); candidate j of run q has slot j - sdel[q]
  int sa[3], sp[4], sdel[3];
  bool fits = true, inside = true;
  sp[0] = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int a = wave_min_full(rl[q] > 0 ? ra[q] : INT_MAX);
    const int e = wave_max_full(rl[q] > 0 ? ra[q] + rl[q] : INT_MIN);
    const bool used = e > a;  // every lane's run q empty: a = INT_MAX, e = INT_MIN
    inside = inside && (rl[q] == 0 || (ra[q] >= a && ra[q] + rl[q] <= e));
    fits = fits && (!used || (a >= 0 && e - a <= kStageCap));
    sa[q] = used ? a : 0;
    sp[q + 1] = sp[q] + ((used && fits) ? e - a : 0);
    sdel[q] = sa[q] - sp[q];
  }
  const bool stage_ok = out.stage_lists && mode == kListDelta16 && fits && sp[3] <= kStageCap &&
                        __all(inside ? 1 : 0) && __all(active ? 1 : 0);
  // word storage: this wave's LDS block, else the global mask reservation
  const bool in_lds = wmax <= kNbMaxWords;
  long long mb = -1;
  if (!in_lds && lane == 0) {
    const unsigned long long need = 64ull * (unsigned long long)wmax;
    const unsigned long long at = atomicAdd(&counters[C_MASK_WORDS], need);
    if ((long long)(at + need) <= out.mask_cap)
      mb = (long long)at;
    else
      atomicAdd(&counters[C_FALLBACK], 1ull);
  }
  mb = __shfl(mb, 0);
  uint2* lcol = lword + lane;
  uint32_t* gcol = out.mask + (mb >= 0 ? mb + lane : 0);
  const bool keep = in_lds || mb >= 0;

  // phase 1: test every candidate, keep the words (in LDS only those holding
  // a neighbor, nlw of them), count (cnt_a: the first three runs)
  int cnt = 0, cnt_a = 0, wofs = 0, nlw = 0;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int nw = (rl[q] + 31) >> 5;
    for (int w = 0; w < nw; ++w) {
      const int jb = ra[q] + (w << 5);
      const uint32_t bits = test_word_planes<kG>(nv.cx, nv.cy, nv.cz, jb, xi, min(32, rl[q] - (w << 5)), nv.thr2);
      if (in_lds) {
        const uint32_t base = mode == kListIndex32
                                  ? (uint32_t)jb
                                  : (((uint32_t)(jb - jbase) & 0xffffu) |
                                     (q < 3 ? ((uint32_t)(jb - sdel[q < 3 ? q : 0]) & 0xffffu) << 16 : 0u));
        if (bits) {
          if (nlw < kNbLdsWords) lcol[nlw << 6] = make_uint2(bits, base);
          ++nlw;
        }
      } else if (mb >= 0) {
        gcol[(long long)(wofs + w) << 6] = bits;
      }
      cnt += __popc(bits);
    }
    wofs += nw;
    if (q == 2) {
      cnt_a = cnt;
    }
  }
  // An out-of-grid lane (the scenes' out-of-grid buckets only) registers
  // itself with each particle it lists, for pass B.
  if (oog_lane) {
    int wo = 0;
    for (int q = 0; q < 9; ++q) {
      const int nw = (rl[q] + 31) >> 5;
      for (int w = 0; w < nw; ++w) {
        const int jb = ra[q] + (w << 5);
        uint32_t bits = (!in_lds && mb >= 0)
                            ? gcol[(long long)(wo + w) << 6]
                            : test_word_planes<kG>(nv.cx, nv.cy, nv.cz, jb, xi, min(32, rl[q] - (w << 5)), nv.thr2);
        while (bits) {
          const int j = jb + pop_first(bits);
          const int slot = atomicAdd(&out.oog_rcnt[j], 1);
          if (slot < kOogRev) out.oog_rev[(size_t)j * kOogRev + slot] = s;
        }
      }
      wo += nw;
    }
  }
  // entries per lane the region must hold: whole pairs of groups for 16-bit
  // lists, whole groups for 32-bit ones. A staged list gives every lane a stage
  // part of kA entries (the wave's mean stage count rounded up to 8): a lane
  // with fewer pads it, a lane with more lists the excess as offsets after it
  // (the wave then runs about as many entries as its longest list).
  const int cmax = wave_max(cnt);
  const int need_plain = mode == kListDelta16 ? (cmax + 7) & ~7 : ((cmax + 3) & ~3) << 1;  // 16-bit slots per lane
  int kA = 0, need_staged = 0, nfull_b = 0, n_slot = 0;
  if (stage_ok) {
    kA = ((wave_sum(cnt_a) >> 6) + 7) & ~7;
    n_slot = min(cnt_a, kA);
    need_staged = kA + ((wave_max_full(cnt - n_slot) + 7) & ~7);
    nfull_b = wave_min_full(cnt - n_slot) >> 2;
  }
  // The wave's list lives in its own fixed region (no allocation atomic: a
  // same-address atomic per wave serializes in L2); a wave that needs more
  // than the region takes the mask or re-test path (or the unstaged list) and
  // reports its need.
  const bool staged = stage_ok && keep && need_staged <= out.list_per;
  const int need = stage_ok ? need_staged : need_plain;
  const int use = staged ? need_staged : need_plain;
  const long long lb = (keep && use <= out.list_per) ? (long long)wave * 64 * out.list_per : -1;
  const int nfull = staged ? nfull_b : wave_min(active ? cnt : INT_MAX) >> 2;  // groups every lane fills
  const int tot = wave_sum(cnt), tcand = wave_sum(cand);
  if (lane == 0) {
    out.mask_base[wave] = mb;
    out.list_base[wave] = lb;
    out.list_jbase[wave] = wjlo;
    out.list_mode[wave] = (staged ? kListStaged : mode) | (nfull << 2);
    out.wave_stat[wave] = make_int4(tot, tcand, need, staged ? 1 : 0);
    if (staged) {
      int* st = out.stage + kStageInts * wave;
      st[0] = sa[0];
      st[1] = sa[1];
      st[2] = sa[2];
      st[3] = sp[0];
      st[4] = sp[1];
      st[5] = sp[2];
      st[6] = sp[3];
      st[7] = kA;
    }
  }
  if (active) out.nbr_cnt[s] = cnt;
  if (staged) out.cnt_a[s] = (uint16_t)n_slot;

  // phase 2: decode this lane's words into the list, four entries per store
  if (active && lb >= 0) {
    const int n_sl = staged ? n_slot : 0, pad_to = staged ? kA : 0;
    if (in_lds && nlw <= kNbLdsWords) {
      if (mode == kListIndex32)
        write_list_lds<1>(out.list, lb, lane, lcol, cnt, 0, 0);
      else
        write_list_lds<0>(out.list, lb, lane, lcol, cnt, n_sl, pad_to);
    } else if (in_lds) {  // this lane's words did not fit: test its candidates again
      auto retest = [&](int q, int w, int) {
        return test_word_planes<kG>(nv.cx, nv.cy, nv.cz, ra[q] + (w << 5), xi, min(32, rl[q] - (w << 5)), nv.thr2);
      };
      if (mode == kListIndex32)
        write_list_runs<1>(out.list, lb, lane, ra, rl, retest, jbase, sdel, 0, 0);
      else
        write_list_runs<0>(out.list, lb, lane, ra, rl, retest, jbase, sdel, n_sl, pad_to);
    } else {
      auto global = [&](int, int, int wo) { return gcol[(long long)wo << 6]; };
      if (mode == kListIndex32)
        write_list_runs<1>(out.list, lb, lane, ra, rl, global, jbase, sdel, 0, 0);
      else
        write_list_runs<0>(out.list, lb, lane, ra, rl, global, jbase, sdel, n_sl, pad_to);
    }
  }
}

// A search of at most kWideGroupMaxParticles particles runs one wave or fewer
// per SIMD, and each wave waits on its own candidate loads: it tests sixteen
// candidates per load round trip. Larger searches keep groups of eight (issue
// slots bound them, and a run's partly empty last group costs more in groups
// of sixteen). Both give the same bits.
constexpr int kWideGroupMaxParticles = 65536;

template <int kG>
void launch_neighbors_g(const NbrView& nv, const NbrOut& out, unsigned long long* counters, hipStream_t st,
                        const KernelEvents* ev) {
  if (ev)
    hipExtLaunchKernelGGL(k_neighbors<kG>, dim3((nv.n + 63) / 64), dim3(64), 0, st, ev->start, ev->stop, 0, nv, out,
                          counters);
  else
    k_neighbors<kG><<<(nv.n + 63) / 64, 64, 0, st>>>(nv, out, counters);
}

void launch_neighbors(const NbrView& nv, const NbrOut& out, unsigned long long* counters, hipStream_t st,
                      const KernelEvents* ev) {
  if (nv.n <= kWideGroupMaxParticles)
    launch_neighbors_g<16>(nv, out, counters, st, ev);
  else
    launch_neighbors_g<8>(nv, out, counters, st, ev);
}

// One block: sums and maxima over the per-wave statistics.
__global__ __launch_bounds__(1024) void k_wave_stats(const int4* __restrict__ ws, int nwaves,
                                                     unsigned long long* __restrict__ counters) {
  __shared__ unsigned long long sp[16], sc[16];
  __shared__ int sn[16], ss[16];
  unsigned long long pairs = 0, cand = 0;
  int need = 0, staged = 0;
  for (int w = threadIdx.x; w < nwaves; w += blockDim.x) {
    const int4 v = ws[w];
    pairs += (unsigned long long)v.x;
    cand += (unsigned long long)v.y;
    need = max(need, v.z);
    staged += v.w;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    pairs += __shfl_xor(pairs, d);
    cand += __shfl_xor(cand, d);
    need = max(need, __shfl_xor(need, d));
    staged += __shfl_xor(staged, d);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sp[wv] = pairs;
    sc[wv] = cand;
    sn[wv] = need;
    ss[wv] = staged;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      pairs += sp[k];
      cand += sc[k];
      need = max(need, sn[k]);
      staged += ss[k];
    }
    counters[C_PAIRS] = pairs;
    counters[C_CANDIDATES] = cand;
    counters[C_LIST_NEED] = (unsigned long long)need;
    counters[C_STAGED] = (unsigned long long)staged;
    counters[C_LIST_SLOTS] = 64ull * (unsigned long long)nwaves * (unsigned long long)need;
  }
}

void launch_wave_stats(const int4* wave_stat, int nwaves, unsigned long long* counters, hipStream_t st) {
  if (nwaves > 0) k_wave_stats<<<1, 1024, 0, st>>>(wave_stat, nwaves, counters);
}

// ---------------------------------------------------------------------------
// PBD passes (DensityPBD::takeOneIteration, DensityPBD.cu:306-365).

// Pass A: K_ComputeDensity (SummationDensity.cu:13-36) and K_ComputeLambdas
// (DensityPBD.cu:46-90) fused: one traversal gives rho_i, sum g and sum g.g;
// lambda_i goes to xs[s].w for pass B. aux (taped steps only) receives
// (sum_j t^2/r d_ij, 1/(S + eps)) for the adjoint.

// Running sums of pass A for one particle (or one entry parity of it).
struct DensitySums {
  float rho = 0.f, sgx = 0.f, sgy = 0.f, sgz = 0.f, sgg = 0.f;
};

// Adds neighbor xj's terms to a.
__device__ __forceinline__ void density_term(DensitySums& a, float4 xi, float4 xj, const PhysConst& pc) {
#pragma clang fp contract(off)
  const PairGeom p = pair_geom(xi, xj, pc.inv_h);
  const float t2 = p.t * p.t;
  a.rho = fmaf(t2, p.t, a.rho);
  // |g| / |cG| = t^2 (|x_i - x_j| / r = 1); no gradient closer than pair_eps
  const float tg = p.r > pc.pair_eps ? t2 : 0.0f;
  const float gq = tg * p.rinv;
  a.sgx = fmaf(gq, p.dx, a.sgx);
  a.sgy = fmaf(gq, p.dy, a.sgy);
  a.sgz = fmaf(gq, p.dz, a.sgz);
  a.sgg = fmaf(tg, tg, a.sgg);
}

// The sums of the even and the odd entries added, the one order every path uses.
__device__ __forceinline__ DensitySums add_parities(const DensitySums& e, const DensitySums& o) {
  DensitySums t;
  t.rho = e.rho + o.rho;
  t.sgx = e.sgx + o.sgx;
  t.sgy = e.sgy + o.sgy;
  t.sgz = e.sgz + o.sgz;
  t.sgg = e.sgg + o.sgg;
  return t;
}

// Exchanges a and b when c holds (per lane: selects, no branch).
__device__ __forceinline__ void swap_if(bool c, float& a, float& b) {
  const float x = c ? b : a, y = c ? a : b;
  a = x;
  b = y;
}
__device__ __forceinline__ void swap_if(bool c, DensitySums& a, DensitySums& b) {
  swap_if(c, a.rho, b.rho);
  swap_if(c, a.sgx, b.sgx);
  swap_if(c, a.sgy, b.sgy);
  swap_if(c, a.sgz, b.sgz);
  swap_if(c, a.sgg, b.sgg);
}

// The odd lane's sums of a lane pair, read by both lanes.
__device__ __forceinline__ DensitySums partner_sums(const DensitySums& a) {
  DensitySums o;
  o.rho = __shfl_xor(a.rho, 1);
  o.sgx = __shfl_xor(a.sgx, 1);
  o.sgy = __shfl_xor(a.sgy, 1);
  o.sgz = __shfl_xor(a.sgz, 1);
  o.sgg = __shfl_xor(a.sgg, 1);
  return o;
}

// lambda_i from the particle's total sums, into xs[s].w; rho_i and the
// adjoint's sums when asked for.
__device__ __forceinline__ void density_lambda_finish(int s, float4* __restrict__ xs, const PhysConst& pc,
                                                      float* __restrict__ rho_out, const int* __restrict__ perm,
                                                      float4* __restrict__ aux, const DensitySums& t) {
  // nothing contracted, so every kernel that inlines this rounds it alike
#pragma clang fp contract(off)
  const float rho = t.rho * pc.cW;
  const float S = (pc.cG * pc.cG) * (t.sgg + (t.sgx * t.sgx + t.sgy * t.sgy + t.sgz * t.sgz));
  float lam = -(rho - pc.lam_rest) / (S + pc.lam_eps);
  lam = lam > 0.0f ? 0.0f : lam;
  reinterpret_cast<float*>(xs + s)[3] = lam;
  if (rho_out) rho_out[perm[s]] = rho;
  if (aux) aux[s] = make_float4(t.sgx, t.sgy, t.sgz, 1.0f / (S + pc.lam_eps));
}

// The whole list of s on one lane: one partial sum per entry parity.
template <class Traverse>
__device__ __forceinline__ DensitySums density_one_lane(float4 xi, const PhysConst& pc, Traverse&& traverse) {
  DensitySums a[2];
  traverse([&](int, float4 xj, auto par) { density_term(a[decltype(par)::value], xi, xj, pc); },
           [&](bool odd) { swap_if(odd, a[0], a[1]); });
  return add_parities(a[0], a[1]);
}

// One particle s of pass A (stage: the wave's LDS region).
__device__ __forceinline__ void density_lambda_at(const NbrView& nv, int s, float4* __restrict__ xs,
                                                  const PhysConst& pc, float* __restrict__ rho_out,
                                                  const int* __restrict__ perm, float4* __restrict__ aux,
                                                  float4* stage) {
  if (s >= nv.n) return;
  const float4 xi = xs[s];
  const DensitySums t = density_one_lane(xi, pc, [&](auto&& body, auto&& flip) {
    for_each_neighbor_staged<false>(nv, s, xs, stage, body, flip);
  });
  density_lambda_finish(s, xs, pc, rho_out, perm, aux, t);
}

// Pass A with a lane pair per particle (lane h = 0 the even entries, 1 the
// odd ones; stage: the hardware wave's LDS region). Same results as
// density_lambda_at, bit for bit.
__device__ __forceinline__ void density_lambda_at2(const NbrView& nv, int s, int h, float4* __restrict__ xs,
                                                   const PhysConst& pc, float* __restrict__ rho_out,
                                                   const int* __restrict__ perm, float4* __restrict__ aux,
                                                   float4* stage) {
  if (s >= nv.n) return;  // both lanes of the pair
  const float4 xi = xs[s];
  DensitySums a;
  if (for_each_neighbor_half<false>(nv, s, h, xs, stage, [&](float4 xj) { density_term(a, xi, xj, pc); })) {
    const DensitySums o = partner_sums(a);
    if (h == 0) density_lambda_finish(s, xs, pc, rho_out, perm, aux, add_parities(a, o));
    return;
  }
  if (h != 0) return;
  // neither a 16-bit nor a staged list: the even lane takes the whole list
  const DensitySums t = density_one_lane(xi, pc, [&](auto&& body, auto&& flip) {
    for_each_neighbor_p<true>(nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)),
                              [&](int j) { return load_pos<false>(xs, j); }, body, flip);
  });
  density_lambda_finish(s, xs, pc, rho_out, perm, aux, t);
}

template <int B>
__global__ __launch_bounds__(B) void k_density_lambda(NbrView nv, float4* __restrict__ xs, PhysConst pc,
                                                      float* __restrict__ rho_out, const int* __restrict__ perm,
                                                      float4* __restrict__ aux) {
  __shared__ float4 stage_buf[B / 64][kStageRecords];
  const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  density_lambda_at(nv, s, xs, pc, rho_out, perm, aux, stage_buf[threadIdx.x >> 6]);
}

// Lane-pair form: thread g of the grid takes entry parity g & 1 of particle
// g >> 1, so a wave holds 32 particles of one 64-particle list wave.
template <int B>
__global__ __launch_bounds__(B) void k_density_lambda2(NbrView nv, float4* __restrict__ xs, PhysConst pc,
                                                       float* __restrict__ rho_out, const int* __restrict__ perm,
                                                       float4* __restrict__ aux) {
  __shared__ float4 stage_buf[B / 64][kStageRecords];
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  density_lambda_at2(nv, g >> 1, g & 1, xs, pc, rho_out, perm, aux, stage_buf[threadIdx.x >> 6]);
}

// Persistent form: as many blocks as stay resident; each wave walks the
// 64-particle tiles of its XCD's contiguous eighth of the particles with a
// stride of that XCD's resident waves (a static schedule: no atomics). The
// dispatcher sends block b to XCD b % 8; a grid of fewer than 8 blocks splits
// the tiles into as many groups as it has blocks, so every tile is walked.
__device__ __forceinline__ void persistent_tiles(int ntiles, int& first, int& stride, int& end) {
  const int nb = gridDim.x, b = blockIdx.x;
  const int groups = nb < 8 ? nb : 8, x = b % groups;
  const int blocks_x = nb / groups + (x < nb % groups ? 1 : 0);  // blocks of group x
  const int wpb = blockDim.x >> 6;
  const int t0 = (int)((long long)ntiles * x / groups), t1 = (int)((long long)ntiles * (x + 1) / groups);
  first = t0 + (b / groups) * wpb + (threadIdx.x >> 6);
  stride = blocks_x * wpb;
  end = t1;
}

template <int B>
__global__ __launch_bounds__(B) void k_density_lambda_p(NbrView nv, float4* __restrict__ xs, PhysConst pc,
                                                        float* __restrict__ rho_out, const int* __restrict__ perm,
                                                        float4* __restrict__ aux) {
  __shared__ float4 stage_buf[B / 64][kStageRecords];
  int t, stride, end;
  persistent_tiles((nv.n + 63) >> 6, t, stride, end);
  for (; t < end; t += stride) {
    asm volatile("" ::: "memory");  // the last tile's stage reads stay before this tile's stage writes
    density_lambda_at(nv, (t << 6) + (threadIdx.x & 63), xs, pc, rho_out, perm, aux, stage_buf[threadIdx.x >> 6]);
  }
}

// Blocks of a persistent launch: the blocks of `kernel` (B threads) that stay
// resident on all CUs of the current device, at most one per B particles.
template <class K>
int resident_blocks(K kernel, int B) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, B, 0) != hipSuccess)
    cus = per_cu = 1;
  return std::max(cus * per_cu, 1);
}

int persistent_grid(int resident, int B, int n) { return std::max(std::min(resident, (n + B - 1) / B), 1); }

// A launch, with its dispatch timed by ev when given.
template <class K, class... Args>
void launch_timed(K kernel, int grid, int block, hipStream_t st, const KernelEvents* ev, Args... args) {
  if (ev)
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, ev->start, ev->stop, 0, args...);
  else
    kernel<<<grid, block, 0, st>>>(args...);
}

template <int B>
void density_lambda_as(const NbrView& nv, float4* xs, const PhysConst& pc, float* rho_out, const int* perm,
                       hipStream_t st, float4* aux, bool persistent, int lanes, const KernelEvents* ev) {
  if (lanes == 2) {
    launch_timed(k_density_lambda2<B>, (int)((2ll * nv.n + B - 1) / B), B, st, ev, nv, xs, pc, rho_out, perm, aux);
  } else if (persistent) {
    static const int resident = resident_blocks(k_density_lambda_p<B>, B);  // one device type per process
    launch_timed(k_density_lambda_p<B>, persistent_grid(resident, B, nv.n), B, st, ev, nv, xs, pc, rho_out, perm,
                 aux);
  } else {
    launch_timed(k_density_lambda<B>, (nv.n + B - 1) / B, B, st, ev, nv, xs, pc, rho_out, perm, aux);
  }
}

void launch_pass_density_lambda(const NbrView& nv, float4* xs, const PhysConst& pc, float* rho_out, const int* perm,
                                hipStream_t st, float4* aux, int block, bool persistent, int lanes,
                                const KernelEvents* ev) {
  switch (block) {
    case 64: density_lambda_as<64>(nv, xs, pc, rho_out, perm, st, aux, persistent, lanes, ev); break;
    case 128: density_lambda_as<128>(nv, xs, pc, rho_out, perm, st, aux, persistent, lanes, ev); break;
    case 512: density_lambda_as<512>(nv, xs, pc, rho_out, perm, st, aux, persistent, lanes, ev); break;
    default: density_lambda_as<256>(nv, xs, pc, rho_out, perm, st, aux, persistent, lanes, ev); break;
  }
}

// Pass B: K_ComputeDisplacement + K_UpdatePosition (DensityPBD.cu:135-179,
// 227-237) in gather form. The reference adds +dp_ij to i and -dp_ij to j from
// both endpoints of every listed pair and dp_ji == -dp_ij, so a hashed particle
// receives 2*sum_j dp_ij; an out-of-grid particle k is listed by nobody and
// receives sum_j dp_kj once, while each hashed j it lists receives dp_jk once
// more. The last iteration fuses DP_UpdateVelocity (DensityPBD.cu:368-378) and
// the scatter back to original order, or keeps x and v in cell order (with
// each particle's original index in id_out) for the viscosity passes or the
// next step of the call.
// Running sums of pass B for one particle (or one entry parity of it).
struct DispSums {
  float x = 0.f, y = 0.f, z = 0.f;
};

// Adds (lambda_i + lambda_j) t^2 / r (x_i - x_j) of neighbor xj to a.
__device__ __forceinline__ void disp_term(DispSums& a, float4 xi, float4 xj, const PhysConst& pc) {
#pragma clang fp contract(off)
  const PairGeom p = pair_geom(xi, xj, pc.inv_h);
  const float c = p.r > pc.pair_eps ? (xi.w + xj.w) * (p.t * p.t) * p.rinv : 0.0f;
  a.x = fmaf(c, p.dx, a.x);
  a.y = fmaf(c, p.dy, a.y);
  a.z = fmaf(c, p.dz, a.z);
}

__device__ __forceinline__ DispSums add_parities(const DispSums& e, const DispSums& o) {
  DispSums t;
  t.x = e.x + o.x;
  t.y = e.y + o.y;
  t.z = e.z + o.z;
  return t;
}

__device__ __forceinline__ void swap_if(bool c, DispSums& a, DispSums& b) {
  swap_if(c, a.x, b.x);
  swap_if(c, a.y, b.y);
  swap_if(c, a.z, b.z);
}

__device__ __forceinline__ DispSums partner_sums(const DispSums& a) {
  DispSums o;
  o.x = __shfl_xor(a.x, 1);
  o.y = __shfl_xor(a.y, 1);
  o.z = __shfl_xor(a.z, 1);
  return o;
}

template <class Traverse>
__device__ __forceinline__ DispSums disp_one_lane(float4 xi, const PhysConst& pc, Traverse&& traverse) {
  DispSums a[2];
  traverse([&](int, float4 xj, auto par) { disp_term(a[decltype(par)::value], xi, xj, pc); },
           [&](bool odd) { swap_if(odd, a[0], a[1]); });
  return add_parities(a[0], a[1]);
}

// The rest of pass B for particle s from its list's total sums a: the pairs of
// out-of-grid particles, the new position and, in the last iteration, the
// velocity.
__device__ __forceinline__ void displacement_finish(const NbrView& nv, int s, const float4* __restrict__ xs,
                                                    float4* __restrict__ xd, const PhysConst& pc, int last,
                                                    const float4* __restrict__ x0, const float4* __restrict__ vs,
                                                    const int* __restrict__ perm, float* __restrict__ x_out,
                                                    float* __restrict__ v_out, float4* __restrict__ v_sorted,
                                                    int* __restrict__ id_out, float4 xi, const DispSums& a) {
  // nothing contracted, so every kernel that inlines this rounds it alike
#pragma clang fp contract(off)
  float fa = 2.0f;
  DispSums b;  // pairs of out-of-grid particles that list s
  if (nv.counters[C_OOG] != 0) {  // some scene has out-of-grid particles
    const SceneWin win = nv.win[nv.ssid[s]];
    const int ob = win.cell_base + win.ncell;  // the scene's out-of-grid bucket
    const int o0 = nv.cell_start[ob], o1 = nv.cell_start[ob + 1];
    if (s >= o0) {
      fa = 1.0f;  // out-of-grid particle: its own pairs once
    } else if (o1 > o0) {
      // out-of-grid particles k that list s add -dp_ks = dp_sk to s once, in
      // ascending k
      const int rc = nv.oog_rcnt[s];
      if (rc <= kOogRev) {
        const int* rv = nv.oog_rev + (size_t)s * kOogRev;
        int prev = -1;
        for (int u = 0; u < rc; ++u) {
          int k = INT_MAX;
          for (int w = 0; w < rc; ++w) {
            const int c = rv[w];
            k = (c > prev && c < k) ? c : k;
          }
          disp_term(b, xi, xs[k], pc);
          prev = k;
        }
      } else {  // listed by more than kOogRev: scan the scene's out-of-grid particles
        const float4 xni = nv.xn[s];
        int ci, cj, ck;
        cell_of(nv.g, xni.x, xni.y, xni.z, ci, cj, ck);
        for (int k = o0; k < o1; ++k) {
          const float4 xnk = nv.xn[k];
          int ki, kj, kk;
          cell_of(nv.g, xnk.x, xnk.y, xnk.z, ki, kj, kk);
          if (abs((ki >> 1) - (ci >> 1)) <= 1 && abs(kj - cj) <= 1 && abs(kk - ck) <= 1 &&
              within(xnk, xni, nv.thr2))
            disp_term(b, xi, xs[k], pc);
        }
      }
    }
  }
  const float scale = pc.corr * pc.cG;
  const float nx = xi.x + scale * (fa * a.x + b.x);
  const float ny = xi.y + scale * (fa * a.y + b.y);
  const float nz = xi.z + scale * (fa * a.z + b.z);
  if (!last) {
    xd[s] = make_float4(nx, ny, nz, 0.0f);
    return;
  }
  const float4 xo = x0[s];
  const float4 vv = vs[s];
  if (v_sorted) {  // the viscosity passes or the next step follow
    xd[s] = make_float4(nx, ny, nz, 0.0f);
    v_sorted[s] = make_float4(vv.x + (nx - xo.x) / pc.dt, vv.y + (ny - xo.y) / pc.dt, vv.z + (nz - xo.z) / pc.dt,
                              0.0f);
    if (id_out) id_out[s] = perm[s];
    return;
  }
  const size_t i = (size_t)perm[s];
  x_out[3 * i] = nx;
  x_out[3 * i + 1] = ny;
  x_out[3 * i + 2] = nz;
  v_out[3 * i] = vv.x + (nx - xo.x) / pc.dt;
  v_out[3 * i + 1] = vv.y + (ny - xo.y) / pc.dt;
  v_out[3 * i + 2] = vv.z + (nz - xo.z) / pc.dt;
}



__device__ __forceinline__ void displacement_at(const NbrView& nv, int s, const float4* __restrict__ xs,
                                                float4* __restrict__ xd, const PhysConst& pc, int last,
                                                const float4* __restrict__ x0, const float4* __restrict__ vs,
                                                const int* __restrict__ perm, float* __restrict__ x_out,
                                                float* __restrict__ v_out, float4* __restrict__ v_sorted,
                                                int* __restrict__ id_out, float4* stage) {
  if (s == 0 && (!last || v_sorted)) xd[nv.pad] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
  if (s >= nv.n) return;
  const float4 xi = xs[s];
  // sum over the listed j of (lambda_i + lambda_j) t^2 / r (x_i - x_j)
  const DispSums a = disp_one_lane(xi, pc, [&](auto&& body, auto&& flip) {
    for_each_neighbor_staged<true>(nv, s, xs, stage, body, flip);
  });
  displacement_finish(nv, s, xs, xd, pc, last, x0, vs, perm, x_out, v_out, v_sorted, id_out, xi, a);
}

// Pass B with a lane pair per particle, as density_lambda_at2.
__device__ __forceinline__ void displacement_at2(const NbrView& nv, int s, int h, const float4* __restrict__ xs,
                                                 float4* __restrict__ xd, const PhysConst& pc, int last,
                                                 const float4* __restrict__ x0, const float4* __restrict__ vs,
                                                 const int* __restrict__ perm, float* __restrict__ x_out,
                                                 float* __restrict__ v_out, float4* __restrict__ v_sorted,
                                                 int* __restrict__ id_out, float4* stage) {
  if (s == 0 && h == 0 && (!last || v_sorted)) xd[nv.pad] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
  if (s >= nv.n) return;  // both lanes of the pair
  const float4 xi = xs[s];
  DispSums a;
  if (for_each_neighbor_half<true>(nv, s, h, xs, stage, [&](float4 xj) { disp_term(a, xi, xj, pc); })) {
    const DispSums o = partner_sums(a);
    if (h == 0)
      displacement_finish(nv, s, xs, xd, pc, last, x0, vs, perm, x_out, v_out, v_sorted, id_out, xi,
                          add_parities(a, o));
    return;
  }
  if (h != 0) return;
  // neither a 16-bit nor a staged list: the even lane takes the whole list
  const DispSums t = disp_one_lane(xi, pc, [&](auto&& body, auto&& flip) {
    for_each_neighbor_p<true>(nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)),
                              [&](int j) { return load_pos<true>(xs, j); }, body, flip);
  });
  displacement_finish(nv, s, xs, xd, pc, last, x0, vs, perm, x_out, v_out, v_sorted, id_out, xi, t);
}

template <int B>
__global__ __launch_bounds__(B) void k_displacement(NbrView nv, const float4* __restrict__ xs,
                                                    float4* __restrict__ xd, PhysConst pc, int last,
                                                    const float4* __restrict__ x0, const float4* __restrict__ vs,
                                                    const int* __restrict__ perm, float* __restrict__ x_out,
                                                    float* __restrict__ v_out, float4* __restrict__ v_sorted,
                                                    int* __restrict__ id_out) {
  __shared__ float4 stage_buf[B / 64][kStageRecords];
  const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  displacement_at(nv, s, xs, xd, pc, last, x0, vs, perm, x_out, v_out, v_sorted, id_out,
                  stage_buf[threadIdx.x >> 6]);
}

template <int B>
__global__ __launch_bounds__(B) void k_displacement2(NbrView nv, const float4* __restrict__ xs,
                                                     float4* __restrict__ xd, PhysConst pc, int last,
                                                     const float4* __restrict__ x0, const float4* __restrict__ vs,
                                                     const int* __restrict__ perm, float* __restrict__ x_out,
                                                     float* __restrict__ v_out, float4* __restrict__ v_sorted,
                                                     int* __restrict__ id_out) {
  __shared__ float4 stage_buf[B / 64][kStageRecords];
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  displacement_at2(nv, g >> 1, g & 1, xs, xd, pc, last, x0, vs, perm, x_out, v_out, v_sorted, id_out,
                   stage_buf[threadIdx.x >> 6]);
}

template <int B>
__global__ __launch_bounds__(B) void k_displacement_p(NbrView nv, const float4* __restrict__ xs,
                                                      float4* __restrict__ xd, PhysConst pc, int last,
                                                      const float4* __restrict__ x0, const float4* __restrict__ vs,
                                                      const int* __restrict__ perm, float* __restrict__ x_out,
                                                      float* __restrict__ v_out, float4* __restrict__ v_sorted,
                                                      int* __restrict__ id_out) {
  __shared__ float4 stage_buf[B / 64][kStageRecords];
  int t, stride, end;
  persistent_tiles((nv.n + 63) >> 6, t, stride, end);
  for (; t < end; t += stride) {
    asm volatile("" ::: "memory");  // the last tile's stage reads stay before this tile's stage writes
    displacement_at(nv, (t << 6) + (threadIdx.x & 63), xs, xd, pc, last, x0, vs, perm, x_out, v_out, v_sorted,
                    id_out, stage_buf[threadIdx.x >> 6]);
  }
}

template <int B>
void displacement_as(const NbrView& nv, const float4* xs, float4* xd, const PhysConst& pc, bool last,
                     const float4* x0, const float4* vs, const int* perm, float* x_out, float* v_out, hipStream_t st,
                     float4* v_sorted, int* id_out, bool persistent, int lanes, const KernelEvents* ev) {
  const int l = last ? 1 : 0;
  if (lanes == 2) {
    launch_timed(k_displacement2<B>, (int)((2ll * nv.n + B - 1) / B), B, st, ev, nv, xs, xd, pc, l, x0, vs, perm,
                 x_out, v_out, v_sorted, id_out);
  } else if (persistent) {
    static const int resident = resident_blocks(k_displacement_p<B>, B);  // one device type per process
    launch_timed(k_displacement_p<B>, persistent_grid(resident, B, nv.n), B, st, ev, nv, xs, xd, pc, l, x0, vs,
                 perm, x_out, v_out, v_sorted, id_out);
  } else {
    launch_timed(k_displacement<B>, (nv.n + B - 1) / B, B, st, ev, nv, xs, xd, pc, l, x0, vs, perm, x_out, v_out,
                 v_sorted, id_out);
  }
}

void launch_pass_displacement(const NbrView& nv, const float4* xs, float4* xd, const PhysConst& pc, bool last,
                              const float4* x0, const float4* vs, const int* perm, float* x_out, float* v_out,
                              hipStream_t st, float4* v_sorted, int* id_out, int block, bool persistent,
                              int lanes, const KernelEvents* ev) {
#define PBF_DISP(b) \
  displacement_as<b>(nv, xs, xd, pc, last, x0, vs, perm, x_out, v_out, st, v_sorted, id_out, persistent, lanes, ev)
  switch (block) {
    case 64: PBF_DISP(64); break;
    case 128: PBF_DISP(128); break;
    case 512: PBF_DISP(512); break;
    default: PBF_DISP(256); break;
  }
#undef PBF_DISP
}

// ---------------------------------------------------------------------------
// NeighborQuery::queryNeighborFixed (NeighborQuery.cu:514-581): rows of at most
// `limit` neighbors, the nearest ones (the reference keeps them by replacing
// the farthest kept neighbor; ties at the limit follow its visit order, here
// the smaller id wins). Distances are |x_i - x_j| as the reference forms them
// (glm's dot, then sqrtf). Rows sorted by (distance, id), in original ids:
// counts[i], elems[limit*i + k].
constexpr int kFixedBlock = 64;

__global__ __launch_bounds__(kFixedBlock) void k_nbr_fixed(NbrView nv, const int* __restrict__ perm, int limit,
                                                           int* __restrict__ counts, int* __restrict__ elems) {
  __shared__ float sd[kFixedMax][kFixedBlock];
  __shared__ int sid[kFixedMax][kFixedBlock];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nv.n) return;
  const int t = threadIdx.x;
  const float4 xi = nv.xn[s];
  int cnt = 0;
  for_each_neighbor(
      nv, s, [&](int j) { return nv.xn[j]; },
      [&](int j, float4 xj) {
#pragma clang fp contract(off)
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float d = sqrtf((dx * dx + dy * dy) + dz * dz);
        const int id = perm[j];
        int k;
        if (cnt < limit) {
          k = cnt++;
        } else {
          k = limit - 1;
          if (!(d < sd[k][t] || (d == sd[k][t] && id < sid[k][t]))) return;
        }
        // insertion into the sorted row
        while (k > 0 && (d < sd[k - 1][t] || (d == sd[k - 1][t] && id < sid[k - 1][t]))) {
          sd[k][t] = sd[k - 1][t];
          sid[k][t] = sid[k - 1][t];
          --k;
        }
        sd[k][t] = d;
        sid[k][t] = id;
      });
  const size_t i = (size_t)perm[s];
  counts[i] = cnt;
  for (int k = 0; k < cnt; ++k) elems[(size_t)limit * i + k] = sid[k][t];
}

void launch_nbr_fixed(const NbrView& nv, const int* perm, int limit, int* counts, int* elems, hipStream_t st) {
  k_nbr_fixed<<<(nv.n + kFixedBlock - 1) / kFixedBlock, kFixedBlock, 0, st>>>(nv, perm, limit, counts, elems);
}

// ---------------------------------------------------------------------------
// CSR export (NeighborList, Framework/Topology/NeighborList.h:27-66) in
// original ids.

__global__ __launch_bounds__(kBlock) void k_csr_fill(NbrView nv, const int* __restrict__ perm,
                                                     const int* __restrict__ offsets, int* __restrict__ elems,
                                                     long long cap) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nv.n) return;
  const int i = perm[s];
  int o = offsets[i];
  if ((long long)offsets[i + 1] > cap) return;
  for_each_neighbor_x<false>(nv, s, nv.xn, [&](int j, float4) { elems[o++] = perm[j]; });
}

void launch_csr_fill(const NbrView& nv, const int* perm, const int* offsets, int* elems, long long cap,
                     hipStream_t st) {
  k_csr_fill<<<(nv.n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nv, perm, offsets, elems, cap);
}

}  // namespace pbf
