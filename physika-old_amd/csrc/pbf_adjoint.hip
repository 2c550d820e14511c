// Adjoint (reverse-mode derivative) of one fluid step with the neighbor sets
// held fixed. Not part of the reference, which has no derivatives; the
// forward it differentiates is PositionBasedFluidModel::step without
// viscosity (PositionBasedFluidModel.cpp:83-101): integrate, K Jacobi
// iterations of DensityPBD (DensityPBD.cu:306-365), velocity update
// (DensityPBD.cu:368-378).
//
// Forward in the terms of the gather passes (all particles in the grid, so
// every listed pair is listed both ways):
//   u_ij   = t_ij^2 / r_ij  (0 when r_ij <= eps),  g_ij = cG u_ij d_ij,
//   rho_i  = cW sum_j t_ij^3,  G_i = sum_j g_ij,  S_i = sum_j |g_ij|^2 + |G_i|^2,
//   lam_i  = min(0, -(rho_i - rho0) / (S_i + eps)),
//   p'_i   = p_i + C sum_j (lam_i + lam_j) g_ij,   C = 2 corr.
// Reverse of one iteration, a = adjoint of p':
//   lam_bar_i = C sum_j (a_i - a_j) . g_ij
//   rho_bar_i = -lam_bar_i / (S_i + eps),  S_bar_i = -lam_bar_i lam_i / (S_i + eps)
//               (both 0 where lam_i was clamped)
//   p_bar_i   = a_i + sum_j [ J(d_ij) w_ij + (rho_bar_i + rho_bar_j) psi(d_ij) ]
//   w_ij      = C (lam_i + lam_j)(a_i - a_j) + 2 S_bar_i (g_ij + G_i)
//               + 2 (S_bar_j g_ij - S_bar_j G_j)
//   J(d) w    = cG [u w - (2 t/h + t^2/r) (d.w) d / r^2]   (Jacobian of g, symmetric)
//   psi(d)    = -3 cW t^2 d / (h r)                         (gradient of cW t^3)
// (g and psi are odd in d and J even, which turns every pair's scatter onto
// j into the gathered terms above.) Integration and velocity update:
//   x* = x + v* dt, v* = v + g dt, v' = v* + (p_K - x*)/dt, x' = p_K.
// The adjoint of p_K is x_bar' + c with c = v_bar'/dt, and that of x* is
// p_bar_0 - c. c can exceed the rest by orders of magnitude, so the iterates
// are carried as e = p_bar - c (e_K = x_bar', e_k = e_{k+1} + J_k^T (e_{k+1}
// + c), x*_bar = e_0) and f = e + c enters only the pair terms.
// Arithmetic is not contracted, so every neighbor-storage path of the list
// traversal gives the same bits.
#include "pbf_nbr.h"

namespace pbf {

struct F4x2 {
  float4 a, b;
};
struct F4x3 {
  float4 a, b, c;
};

// e = x_bar', c = v_bar'/dt, f = e + c, v*_bar = v_bar'.
__global__ __launch_bounds__(kBlock) void k_adj_begin(int n, const int* __restrict__ perm,
                                                      const float* __restrict__ xbar, const float* __restrict__ vbar,
                                                      float dt, float4* __restrict__ e, float4* __restrict__ c,
                                                      float4* __restrict__ f, float4* __restrict__ vsb,
                                                      const unsigned long long* __restrict__ counters,
                                                      int* __restrict__ flag) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0 && counters[C_OOG] != 0) *flag = 1;
  if (s >= n) return;
  const size_t i = (size_t)perm[s];
  const float4 xb = load3(xbar, (int)i), vb = load3(vbar, (int)i);
  const float cx = vb.x / dt, cy = vb.y / dt, cz = vb.z / dt;
  e[s] = xb;
  c[s] = make_float4(cx, cy, cz, 0.0f);
  f[s] = make_float4(xb.x + cx, xb.y + cy, xb.z + cz, 0.0f);
  vsb[s] = vb;
}

// lam_bar's pair sum: u (a_i - a_j) . d for neighbor (P_j, f_j).
__device__ __forceinline__ void adj_lambda_term(float& acc, float4 ps, float4 as, const F4x2& q,
                                                const PhysConst& pc) {
#pragma clang fp contract(off)
  const PairGeom g = pair_geom(ps, q.a, pc.inv_h);
  const float u = g.r > pc.pair_eps ? g.t * g.t * g.rinv : 0.0f;
  acc += u * ((as.x - q.b.x) * g.dx + (as.y - q.b.y) * g.dy + (as.z - q.b.z) * g.dz);
}

__device__ __forceinline__ void adj_lambda_finish(int s, float4 ps, float4 as, float acc,
                                                  const float4* __restrict__ aux, float4* __restrict__ q1,
                                                  float4* __restrict__ q2, const PhysConst& pc) {
#pragma clang fp contract(off)
  const float lam_bar = (2.0f * pc.corr * pc.cG) * acc;
  const float4 ax = aux[s];
  const bool active = ps.w < 0.0f;
  const float rho_bar = active ? -lam_bar * ax.w : 0.0f;
  const float s_bar = active ? -lam_bar * ps.w * ax.w : 0.0f;
  q1[s] = make_float4(as.x, as.y, as.z, rho_bar);
  q2[s] = make_float4(s_bar * pc.cG * ax.x, s_bar * pc.cG * ax.y, s_bar * pc.cG * ax.z, s_bar);
}

// Sums of the adjoint passes run as one partial sum per entry parity in the
// neighbor order, added at the end (as the forward passes), so that a lane
// pair and one lane give the same bits.
__global__ __launch_bounds__(kBlock) void k_adj_lambda(NbrView nv, const float4* __restrict__ P,
                                                       const float4* __restrict__ aux, const float4* __restrict__ f,
                                                       float4* __restrict__ q1, float4* __restrict__ q2,
                                                       PhysConst pc) {
#pragma clang fp contract(off)
  const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (s >= nv.n) return;
  const float4 ps = P[s], as = f[s];
  float acc[2] = {0.0f, 0.0f};
  for_each_neighbor_p(
      nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)), [&](int j) { return F4x2{P[j], f[j]}; },
      [&](int, const F4x2& q, auto par) { adj_lambda_term(acc[decltype(par)::value], ps, as, q, pc); },
      [&](bool odd) {
        const float x = odd ? acc[1] : acc[0], y = odd ? acc[0] : acc[1];
        acc[0] = x;
        acc[1] = y;
      });
  adj_lambda_finish(s, ps, as, acc[0] + acc[1], aux, q1, q2, pc);
}

// Lane-pair form (small searches): thread g takes entry parity g & 1 of
// particle g >> 1; waves with 32-bit lists or none run the list on the even lane.
__global__ __launch_bounds__(kBlock) void k_adj_lambda2(NbrView nv, const float4* __restrict__ P,
                                                        const float4* __restrict__ aux, const float4* __restrict__ f,
                                                        float4* __restrict__ q1, float4* __restrict__ q2,
                                                        PhysConst pc) {
#pragma clang fp contract(off)
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int s = g >> 1, h = g & 1;
  if (s >= nv.n) return;  // both lanes of the pair
  const float4 ps = P[s], as = f[s];
  auto gather = [&](int j) { return F4x2{P[j], f[j]}; };
  float acc = 0.0f;
  if (for_each_neighbor_half_idx(nv, s, h, gather,
                                 [&](int, const F4x2& q) { adj_lambda_term(acc, ps, as, q, pc); })) {
    const float odd = __shfl_xor(acc, 1);
    if (h == 0) adj_lambda_finish(s, ps, as, acc + odd, aux, q1, q2, pc);
    return;
  }
  if (h != 0) return;
  float a2[2] = {0.0f, 0.0f};
  for_each_neighbor_p(
      nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)), gather,
      [&](int, const F4x2& q, auto par) { adj_lambda_term(a2[decltype(par)::value], ps, as, q, pc); },
      [&](bool odd) {
        const float x = odd ? a2[1] : a2[0], y = odd ? a2[0] : a2[1];
        a2[0] = x;
        a2[1] = y;
      });
  adj_lambda_finish(s, ps, as, a2[0] + a2[1], aux, q1, q2, pc);
}

struct AdjSums {
  float x = 0.0f, y = 0.0f, z = 0.0f;
};

// p_bar's pair sum for neighbor (P_j, q1_j, q2_j).
__device__ __forceinline__ void adj_pos_term(AdjSums& b, float4 ps, float4 q1s, float4 q2s, float Gx, float Gy,
                                             float Gz, const F4x3& q, const PhysConst& pc) {
#pragma clang fp contract(off)
  const float C = 2.0f * pc.corr;
  const float cpsi = -3.0f * pc.cW * pc.inv_h;
  const PairGeom g = pair_geom(ps, q.a, pc.inv_h);
  const float t2 = g.t * g.t;
  // density term, every listed pair (zero for the self pair: d = 0)
  const float cr = (q1s.w + q.b.w) * cpsi * t2 * g.rinv;
  b.x += cr * g.dx;
  b.y += cr * g.dy;
  b.z += cr * g.dz;
  if (g.r > pc.pair_eps) {
    const float u = t2 * g.rinv;
    const float cg = pc.cG * u;
    const float gx = cg * g.dx, gy = cg * g.dy, gz = cg * g.dz;
    const float cl = C * (ps.w + q.a.w);
    const float wx = cl * (q1s.x - q.b.x) + 2.0f * q2s.w * (gx + Gx) + 2.0f * (q.c.w * gx - q.c.x);
    const float wy = cl * (q1s.y - q.b.y) + 2.0f * q2s.w * (gy + Gy) + 2.0f * (q.c.w * gy - q.c.y);
    const float wz = cl * (q1s.z - q.b.z) + 2.0f * q2s.w * (gz + Gz) + 2.0f * (q.c.w * gz - q.c.z);
    const float dw = g.dx * wx + g.dy * wy + g.dz * wz;
    const float k2 = (2.0f * g.t * pc.inv_h + u) * g.rinv * g.rinv * dw;
    b.x += pc.cG * (u * wx - k2 * g.dx);
    b.y += pc.cG * (u * wy - k2 * g.dy);
    b.z += pc.cG * (u * wz - k2 * g.dz);
  }
}

__device__ __forceinline__ void adj_pos_finish(int s, const AdjSums& b, const float4* __restrict__ e_in,
                                               const float4* __restrict__ c, float4* __restrict__ e_out,
                                               float4* __restrict__ f_out) {
#pragma clang fp contract(off)
  const float4 es = e_in[s], cs = c[s];
  const float ex = es.x + b.x, ey = es.y + b.y, ez = es.z + b.z;
  e_out[s] = make_float4(ex, ey, ez, 0.0f);
  f_out[s] = make_float4(ex + cs.x, ey + cs.y, ez + cs.z, 0.0f);
}

__device__ __forceinline__ AdjSums add_parities(const AdjSums& e, const AdjSums& o) {
  AdjSums t;
  t.x = e.x + o.x;
  t.y = e.y + o.y;
  t.z = e.z + o.z;
  return t;
}

__device__ __forceinline__ void swap_if(bool c, AdjSums& a, AdjSums& b) {
  const AdjSums x = c ? b : a, y = c ? a : b;
  a = x;
  b = y;
}

__global__ __launch_bounds__(kBlock) void k_adj_pos(NbrView nv, const float4* __restrict__ P,
                                                    const float4* __restrict__ aux, const float4* __restrict__ q1,
                                                    const float4* __restrict__ q2, const float4* __restrict__ e_in,
                                                    const float4* __restrict__ c, float4* __restrict__ e_out,
                                                    float4* __restrict__ f_out, PhysConst pc) {
#pragma clang fp contract(off)
  const int s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (s >= nv.n) return;
  const float4 ps = P[s], q1s = q1[s], q2s = q2[s], ax = aux[s];
  const float Gx = pc.cG * ax.x, Gy = pc.cG * ax.y, Gz = pc.cG * ax.z;
  AdjSums b[2];
  for_each_neighbor_p(
      nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)),
      [&](int j) { return F4x3{P[j], q1[j], q2[j]}; },
      [&](int, const F4x3& q, auto par) { adj_pos_term(b[decltype(par)::value], ps, q1s, q2s, Gx, Gy, Gz, q, pc); },
      [&](bool odd) { swap_if(odd, b[0], b[1]); });
  adj_pos_finish(s, add_parities(b[0], b[1]), e_in, c, e_out, f_out);
}

__global__ __launch_bounds__(kBlock) void k_adj_pos2(NbrView nv, const float4* __restrict__ P,
                                                     const float4* __restrict__ aux, const float4* __restrict__ q1,
                                                     const float4* __restrict__ q2, const float4* __restrict__ e_in,
                                                     const float4* __restrict__ c, float4* __restrict__ e_out,
                                                     float4* __restrict__ f_out, PhysConst pc) {
#pragma clang fp contract(off)
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int s = g >> 1, h = g & 1;
  if (s >= nv.n) return;  // both lanes of the pair
  const float4 ps = P[s], q1s = q1[s], q2s = q2[s], ax = aux[s];
  const float Gx = pc.cG * ax.x, Gy = pc.cG * ax.y, Gz = pc.cG * ax.z;
  auto gather = [&](int j) { return F4x3{P[j], q1[j], q2[j]}; };
  AdjSums b;
  if (for_each_neighbor_half_idx(nv, s, h, gather, [&](int, const F4x3& q) {
        adj_pos_term(b, ps, q1s, q2s, Gx, Gy, Gz, q, pc);
      })) {
    AdjSums o;
    o.x = __shfl_xor(b.x, 1);
    o.y = __shfl_xor(b.y, 1);
    o.z = __shfl_xor(b.z, 1);
    if (h == 0) adj_pos_finish(s, add_parities(b, o), e_in, c, e_out, f_out);
    return;
  }
  if (h != 0) return;
  AdjSums b2[2];
  for_each_neighbor_p(
      nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)), gather,
      [&](int, const F4x3& q, auto par) { adj_pos_term(b2[decltype(par)::value], ps, q1s, q2s, Gx, Gy, Gz, q, pc); },
      [&](bool odd) { swap_if(odd, b2[0], b2[1]); });
  adj_pos_finish(s, add_parities(b2[0], b2[1]), e_in, c, e_out, f_out);
}

// x_bar = e_0 (the adjoint of x*), v_bar = v*_bar + x_bar dt, original order.
__global__ __launch_bounds__(kBlock) void k_adj_end(int n, const int* __restrict__ perm,
                                                    const float4* __restrict__ e, const float4* __restrict__ vsb,
                                                    float dt, float* __restrict__ xbar, float* __restrict__ vbar) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = (size_t)perm[s];
  const float4 ee = e[s], vb = vsb[s];
  xbar[3 * i] = ee.x;
  xbar[3 * i + 1] = ee.y;
  xbar[3 * i + 2] = ee.z;
  vbar[3 * i] = vb.x + ee.x * dt;
  vbar[3 * i + 1] = vb.y + ee.y * dt;
  vbar[3 * i + 2] = vb.z + ee.z * dt;
}

void launch_adjoint_step(const NbrView& nv, float4* const* P, float4* const* aux, int K, const PhysConst& pc,
                         const int* perm, float* xbar, float* vbar, const AdjBufs& ab, hipStream_t st, int lanes) {
  const int n = nv.n;
  const int nb = (n + kBlock - 1) / kBlock;
  const int nb2 = (int)((2ll * n + kBlock - 1) / kBlock);
  k_adj_begin<<<nb, kBlock, 0, st>>>(n, perm, xbar, vbar, pc.dt, ab.e, ab.c, ab.f, ab.vsb, nv.counters, ab.flag);
  float4* e = ab.e;
  float4* e2 = ab.e2;
  for (int k = K - 1; k >= 0; --k) {
    if (lanes == 2) {
      k_adj_lambda2<<<nb2, kBlock, 0, st>>>(nv, P[k], aux[k], ab.f, ab.q1, ab.q2, pc);
      k_adj_pos2<<<nb2, kBlock, 0, st>>>(nv, P[k], aux[k], ab.q1, ab.q2, e, ab.c, e2, ab.f, pc);
    } else {
      k_adj_lambda<<<nb, kBlock, 0, st>>>(nv, P[k], aux[k], ab.f, ab.q1, ab.q2, pc);
      k_adj_pos<<<nb, kBlock, 0, st>>>(nv, P[k], aux[k], ab.q1, ab.q2, e, ab.c, e2, ab.f, pc);
    }
    float4* t = e;
    e = e2;
    e2 = t;
  }
  k_adj_end<<<nb, kBlock, 0, st>>>(n, perm, e, ab.vsb, pc.dt, xbar, vbar);
}

}  // namespace pbf
