// Double-precision forward and adjoint of the fluid step, for reverse mode
// whose gradients must match FP64 finite differences over long horizons. Not
// part of the reference, which has no derivatives; the forward is
// PositionBasedFluidModel::step without viscosity
// (PositionBasedFluidModel.cpp:83-101): integrate (ParticleIntegrator.cu:
// 60-70,103-112), K Jacobi iterations of DensityPBD (DensityPBD.cu:306-365)
// and the velocity update (DensityPBD.cu:368-378), every quantity in double.
//
// The neighbor sets come from the single-precision search on the positions
// rounded to float: they can differ from a double-precision search only for
// pairs within a few float ulps of h, where the kernel and its first
// derivative vanish (t = 1 - r/h ~ 1e-7), so values and gradients move by
// O(t^2) ~ 1e-14 relative.
//
// Forward in gather form (all particles in the grid, every pair listed both
// ways), u_ij = t^2/r (0 when r <= eps), g_ij = cG u_ij d_ij:
//   rho_i = cW sum_j t^3,  G_i = sum_j g_ij,  S_i = sum_j |g_ij|^2 + |G_i|^2,
//   lam_i = min(0, -(rho_i - rho0)/(S_i + eps)),
//   p'_i  = p_i + 2 corr sum_j (lam_i + lam_j) g_ij.
// The adjoint is that of the single-precision kernels (pbf_adjoint.hip),
// evaluated in double: iterates carried as e = p_bar - c with c = v_bar'/dt.
#include "pbf_nbr.h"

namespace pbf {

namespace {

struct PairGeom64 {
  double dx, dy, dz, r, rinv, t;
};

__device__ __forceinline__ PairGeom64 pair_geom64(const double4& a, const double4& b, double inv_h) {
#pragma clang fp contract(off)
  PairGeom64 p;
  p.dx = a.x - b.x;
  p.dy = a.y - b.y;
  p.dz = a.z - b.z;
  const double r2 = (p.dx * p.dx + p.dy * p.dy) + p.dz * p.dz;
  p.r = sqrt(r2);
  p.rinv = r2 > 0.0 ? 1.0 / p.r : 0.0;
  p.t = fmax(1.0 - p.r * inv_h, 0.0);
  return p;
}

struct D4x2 {
  double4 a, b;
};
struct D4x3 {
  double4 a, b, c;
};

__device__ __forceinline__ double4 load3d(const double* x, int i) {
  const size_t b = 3 * (size_t)i;
  return make_double4(x[b], x[b + 1], x[b + 2], 0.0);
}

__device__ __forceinline__ void store3d(double* x, int i, double a, double b, double c) {
  const size_t o = 3 * (size_t)i;
  x[o] = a;
  x[o + 1] = b;
  x[o + 2] = c;
}

// The positions x^n rounded to float (the neighbor search's input).
__global__ __launch_bounds__(kBlock) void k64_round(const double* __restrict__ x, float* __restrict__ xf, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * n) return;
  xf[i] = (float)x[i];
}

// Integration in sorted order: v* = v + (0 + g) dt, x* = x + v* dt; P0 = (x*,
// 0), vs = v*. Flags a search with particles outside the grid (their lists
// are one-sided, which this path does not handle).
__global__ __launch_bounds__(kBlock) void k64_begin(int n, const int* __restrict__ perm, const double* __restrict__ x,
                                                    const double* __restrict__ v, PhysConst64 pc,
                                                    double4* __restrict__ P0, double4* __restrict__ vs,
                                                    const unsigned long long* __restrict__ counters,
                                                    int* __restrict__ flag) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0 && counters[C_OOG] != 0) *flag = 1;
  if (s >= n) return;
  const int i = perm[s];
  const double4 xi = load3d(x, i), vi = load3d(v, i);
  const double vx = vi.x + (0.0 + pc.g[0]) * pc.dt;
  const double vy = vi.y + (0.0 + pc.g[1]) * pc.dt;
  const double vz = vi.z + (0.0 + pc.g[2]) * pc.dt;
  P0[s] = make_double4(xi.x + vx * pc.dt, xi.y + vy * pc.dt, xi.z + vz * pc.dt, 0.0);
  vs[s] = make_double4(vx, vy, vz, 0.0);
}

// Sums of a pass over a particle's neighbors: N doubles, kept as one partial
// sum per entry parity in the neighbor order and added at the end, as the
// float passes do, so that one lane and a lane pair give the same bits.
template <int N>
struct SumsD {
  double v[N] = {};
};

// The whole list of s on one lane: term(sums, gather(j)) per neighbor.
template <int N, class Gather, class Term>
__device__ __forceinline__ SumsD<N> sums64_one_lane(const NbrView& nv, int s, Gather&& gather, Term&& term) {
  SumsD<N> a[2];
  for_each_neighbor_p(
      nv, s, wave_list(nv, __builtin_amdgcn_readfirstlane(s >> 6)), gather,
      [&](int, const auto& q, auto par) { term(a[decltype(par)::value], q); },
      [&](bool odd) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double x = odd ? a[1].v[i] : a[0].v[i], y = odd ? a[0].v[i] : a[1].v[i];
          a[0].v[i] = x;
          a[1].v[i] = y;
        }
      });
  SumsD<N> t;
#pragma unroll
  for (int i = 0; i < N; ++i) t.v[i] = a[0].v[i] + a[1].v[i];
  return t;
}

// Lane h of the pair holding s sums the entries of parity h; t receives the
// even + odd sums (on both lanes). False, with nothing run, where s's wave has
// neither a 16-bit nor a staged list.
template <int N, class Gather, class Term>
__device__ __forceinline__ bool sums64_lane_pair(const NbrView& nv, int s, int h, Gather&& gather, Term&& term,
                                                 SumsD<N>& t) {
  SumsD<N> a;
  if (!for_each_neighbor_half_idx(nv, s, h, gather, [&](int, const auto& q) { term(a, q); })) return false;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double o = __shfl_xor(a.v[i], 1);
    t.v[i] = h == 0 ? a.v[i] + o : o + a.v[i];
  }
  return true;
}

// The sums of a pass for sorted particle s: thread g of the grid takes particle
// g (kPair false) or entry parity g & 1 of particle g >> 1 (kPair: a lane pair
// per particle, for small searches). Returns false on the lanes that must not
// write s's results (past the end, or the odd lane of a pair).
template <bool kPair, int N, class Gather, class Term>
__device__ __forceinline__ bool sums64_of(const NbrView& nv, int s, int h, Gather&& gather, Term&& term,
                                          SumsD<N>& t) {
  if (kPair && sums64_lane_pair<N>(nv, s, h, gather, term, t)) return h == 0;
  if (h != 0) return false;
  t = sums64_one_lane<N>(nv, s, gather, term);
  return true;
}

// Density and lambda; lambda goes to P[s].w, (G, 1/(S + eps)) to aux (taped
// steps), the density of the last iteration to rho_out (original order).
template <bool kPair>
__global__ __launch_bounds__(kBlock) void k64_density_lambda(NbrView nv, double4* __restrict__ P, PhysConst64 pc,
                                                             double4* __restrict__ aux, double* __restrict__ rho_out,
                                                             const int* __restrict__ perm) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = kPair ? g >> 1 : g, h = kPair ? g & 1 : 0;
  if (s >= nv.n) return;
  const double4 ps = P[s];
  SumsD<5> t;  // rho, G, |g|^2
  if (!sums64_of<kPair, 5>(
          nv, s, h, [&](int j) { return P[j]; },
          [&](auto& a, const double4& pj) {
            const PairGeom64 q = pair_geom64(ps, pj, pc.inv_h);
            const double t2 = q.t * q.t;
            a.v[0] += t2 * q.t;
            const double tg = q.r > pc.pair_eps ? t2 : 0.0;
            const double u = tg * q.rinv;
            a.v[1] += u * q.dx;
            a.v[2] += u * q.dy;
            a.v[3] += u * q.dz;
            a.v[4] += tg * tg;
          },
          t))
    return;
  const double rho = t.v[0] * pc.cW;
  const double sgx = t.v[1], sgy = t.v[2], sgz = t.v[3], sgg = t.v[4];
  const double S = (pc.cG * pc.cG) * (sgg + (sgx * sgx + sgy * sgy + sgz * sgz));
  double lam = -(rho - pc.lam_rest) / (S + pc.lam_eps);
  lam = lam > 0.0 ? 0.0 : lam;
  reinterpret_cast<double*>(P + s)[3] = lam;
  if (aux) aux[s] = make_double4(sgx, sgy, sgz, 1.0 / (S + pc.lam_eps));
  if (rho_out) rho_out[perm[s]] = rho;
}

// Displacement p' = p + 2 corr cG sum_j (lam_i + lam_j) u d; the last
// iteration writes x' = p_K and v' = v* + (p_K - x*)/dt in original order.
template <bool kPair>
__global__ __launch_bounds__(kBlock) void k64_displacement(NbrView nv, const double4* __restrict__ P,
                                                           double4* __restrict__ Pn, PhysConst64 pc, int last,
                                                           const double4* __restrict__ P0,
                                                           const double4* __restrict__ vs,
                                                           const int* __restrict__ perm, double* __restrict__ x_out,
                                                           double* __restrict__ v_out) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = kPair ? g >> 1 : g, h = kPair ? g & 1 : 0;
  if (s >= nv.n) return;
  const double4 ps = P[s];
  SumsD<3> t;
  if (!sums64_of<kPair, 3>(
          nv, s, h, [&](int j) { return P[j]; },
          [&](auto& a, const double4& pj) {
            const PairGeom64 q = pair_geom64(ps, pj, pc.inv_h);
            const double c = q.r > pc.pair_eps ? (ps.w + pj.w) * (q.t * q.t) * q.rinv : 0.0;
            a.v[0] += c * q.dx;
            a.v[1] += c * q.dy;
            a.v[2] += c * q.dz;
          },
          t))
    return;
  const double ax = t.v[0], ay = t.v[1], az = t.v[2];
  const double scale = pc.corr * pc.cG;
  const double nx = ps.x + scale * (2.0 * ax), ny = ps.y + scale * (2.0 * ay), nz = ps.z + scale * (2.0 * az);
  if (!last) {
    Pn[s] = make_double4(nx, ny, nz, 0.0);
    return;
  }
  const int i = perm[s];
  const double4 xo = P0[s], vv = vs[s];
  store3d(x_out, i, nx, ny, nz);
  store3d(v_out, i, vv.x + (nx - xo.x) / pc.dt, vv.y + (ny - xo.y) / pc.dt, vv.z + (nz - xo.z) / pc.dt);
}

// No PBD iterations: x' = x*, v' = v*.
__global__ __launch_bounds__(kBlock) void k64_write(int n, const int* __restrict__ perm,
                                                    const double4* __restrict__ P0, const double4* __restrict__ vs,
                                                    double* __restrict__ x_out, double* __restrict__ v_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const double4 p = P0[s], v = vs[s];
  store3d(x_out, i, p.x, p.y, p.z);
  store3d(v_out, i, v.x, v.y, v.z);
}

// ---- adjoint --------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k64_adj_begin(int n, const int* __restrict__ perm,
                                                        const double* __restrict__ xbar,
                                                        const double* __restrict__ vbar, double dt,
                                                        double4* __restrict__ e, double4* __restrict__ c,
                                                        double4* __restrict__ f, double4* __restrict__ vsb) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const double4 xb = load3d(xbar, i), vb = load3d(vbar, i);
  const double cx = vb.x / dt, cy = vb.y / dt, cz = vb.z / dt;
  e[s] = xb;
  c[s] = make_double4(cx, cy, cz, 0.0);
  f[s] = make_double4(xb.x + cx, xb.y + cy, xb.z + cz, 0.0);
  vsb[s] = vb;
}

// lam_bar_i = 2 corr cG sum_j u (f_i - f_j) . d; rho_bar, S_bar where lambda
// was not clamped.
template <bool kPair>
__global__ __launch_bounds__(kBlock) void k64_adj_lambda(NbrView nv, const double4* __restrict__ P,
                                                         const double4* __restrict__ aux,
                                                         const double4* __restrict__ f, double4* __restrict__ q1,
                                                         double4* __restrict__ q2, PhysConst64 pc) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = kPair ? g >> 1 : g, h = kPair ? g & 1 : 0;
  if (s >= nv.n) return;
  const double4 ps = P[s], as = f[s];
  SumsD<1> t;
  if (!sums64_of<kPair, 1>(
          nv, s, h, [&](int j) { return D4x2{P[j], f[j]}; },
          [&](auto& a, const D4x2& q) {
            const PairGeom64 r = pair_geom64(ps, q.a, pc.inv_h);
            const double u = r.r > pc.pair_eps ? r.t * r.t * r.rinv : 0.0;
            a.v[0] += u * ((as.x - q.b.x) * r.dx + (as.y - q.b.y) * r.dy + (as.z - q.b.z) * r.dz);
          },
          t))
    return;
  const double lam_bar = (2.0 * pc.corr * pc.cG) * t.v[0];
  const double4 ax = aux[s];
  const bool active = ps.w < 0.0;
  const double rho_bar = active ? -lam_bar * ax.w : 0.0;
  const double s_bar = active ? -lam_bar * ps.w * ax.w : 0.0;
  q1[s] = make_double4(as.x, as.y, as.z, rho_bar);
  q2[s] = make_double4(s_bar * pc.cG * ax.x, s_bar * pc.cG * ax.y, s_bar * pc.cG * ax.z, s_bar);
}

// e_k = e_{k+1} + sum_j [J(d) w + (rho_bar_i + rho_bar_j) psi(d)], f = e + c.
template <bool kPair>
__global__ __launch_bounds__(kBlock) void k64_adj_pos(NbrView nv, const double4* __restrict__ P,
                                                      const double4* __restrict__ aux,
                                                      const double4* __restrict__ q1,
                                                      const double4* __restrict__ q2,
                                                      const double4* __restrict__ e_in,
                                                      const double4* __restrict__ c, double4* __restrict__ e_out,
                                                      double4* __restrict__ f_out, PhysConst64 pc) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = kPair ? g >> 1 : g, h = kPair ? g & 1 : 0;
  if (s >= nv.n) return;
  const double4 ps = P[s], q1s = q1[s], q2s = q2[s], ax = aux[s];
  const double Gx = pc.cG * ax.x, Gy = pc.cG * ax.y, Gz = pc.cG * ax.z;
  const double C = 2.0 * pc.corr;
  const double cpsi = -3.0 * pc.cW * pc.inv_h;
  SumsD<3> t;
  if (!sums64_of<kPair, 3>(
          nv, s, h, [&](int j) { return D4x3{P[j], q1[j], q2[j]}; },
          [&](auto& a, const D4x3& q) {
            const PairGeom64 r = pair_geom64(ps, q.a, pc.inv_h);
            const double t2 = r.t * r.t;
            const double cr = (q1s.w + q.b.w) * cpsi * t2 * r.rinv;
            a.v[0] += cr * r.dx;
            a.v[1] += cr * r.dy;
            a.v[2] += cr * r.dz;
            if (r.r > pc.pair_eps) {
              const double u = t2 * r.rinv;
              const double cg = pc.cG * u;
              const double gx = cg * r.dx, gy = cg * r.dy, gz = cg * r.dz;
              const double cl = C * (ps.w + q.a.w);
              const double wx = cl * (q1s.x - q.b.x) + 2.0 * q2s.w * (gx + Gx) + 2.0 * (q.c.w * gx - q.c.x);
              const double wy = cl * (q1s.y - q.b.y) + 2.0 * q2s.w * (gy + Gy) + 2.0 * (q.c.w * gy - q.c.y);
              const double wz = cl * (q1s.z - q.b.z) + 2.0 * q2s.w * (gz + Gz) + 2.0 * (q.c.w * gz - q.c.z);
              const double dw = r.dx * wx + r.dy * wy + r.dz * wz;
              const double k2 = (2.0 * r.t * pc.inv_h + u) * r.rinv * r.rinv * dw;
              a.v[0] += pc.cG * (u * wx - k2 * r.dx);
              a.v[1] += pc.cG * (u * wy - k2 * r.dy);
              a.v[2] += pc.cG * (u * wz - k2 * r.dz);
            }
          },
          t))
    return;
  const double4 es = e_in[s], cs = c[s];
  const double ex = es.x + t.v[0], ey = es.y + t.v[1], ez = es.z + t.v[2];
  e_out[s] = make_double4(ex, ey, ez, 0.0);
  f_out[s] = make_double4(ex + cs.x, ey + cs.y, ez + cs.z, 0.0);
}

__global__ __launch_bounds__(kBlock) void k64_adj_end(int n, const int* __restrict__ perm,
                                                      const double4* __restrict__ e, const double4* __restrict__ vsb,
                                                      double dt, double* __restrict__ xbar,
                                                      double* __restrict__ vbar) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const double4 ee = e[s], vb = vsb[s];
  store3d(xbar, i, ee.x, ee.y, ee.z);
  store3d(vbar, i, vb.x + ee.x * dt, vb.y + ee.y * dt, vb.z + ee.z * dt);
}

inline int blocks_of(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

void launch64_round(const double* x, float* xf, int n, hipStream_t st) {
  if (n > 0) k64_round<<<blocks_of(3 * n), kBlock, 0, st>>>(x, xf, n);
}

void launch64_step(const NbrView& nv, const int* perm, const double* x_in, const double* v_in, double* x_out,
                   double* v_out, double* rho, double4* const* P, double4* const* aux, int K, double4* vs,
                   const PhysConst64& pc, const unsigned long long* counters, int* flag, hipStream_t st, int lanes) {
  const int n = nv.n, nb = blocks_of(n), nbp = (int)((2ll * n + kBlock - 1) / kBlock);
  k64_begin<<<nb, kBlock, 0, st>>>(n, perm, x_in, v_in, pc, P[0], vs, counters, flag);
  if (K == 0) {
    k64_write<<<nb, kBlock, 0, st>>>(n, perm, P[0], vs, x_out, v_out);
    return;
  }
  for (int k = 0; k < K; ++k) {
    const bool last = k + 1 == K;
    if (lanes == 2) {
      k64_density_lambda<true><<<nbp, kBlock, 0, st>>>(nv, P[k], pc, aux ? aux[k] : nullptr, last ? rho : nullptr,
                                                       perm);
      k64_displacement<true><<<nbp, kBlock, 0, st>>>(nv, P[k], last ? nullptr : P[k + 1], pc, last ? 1 : 0, P[0], vs,
                                                     perm, x_out, v_out);
    } else {
      k64_density_lambda<false><<<nb, kBlock, 0, st>>>(nv, P[k], pc, aux ? aux[k] : nullptr, last ? rho : nullptr,
                                                       perm);
      k64_displacement<false><<<nb, kBlock, 0, st>>>(nv, P[k], last ? nullptr : P[k + 1], pc, last ? 1 : 0, P[0], vs,
                                                     perm, x_out, v_out);
    }
  }
}

void launch64_adjoint_step(const NbrView& nv, double4* const* P, double4* const* aux, int K, const PhysConst64& pc,
                           const int* perm, double* xbar, double* vbar, const AdjBufs64& ab, hipStream_t st,
                           int lanes) {
  const int n = nv.n, nb = blocks_of(n), nbp = (int)((2ll * n + kBlock - 1) / kBlock);
  k64_adj_begin<<<nb, kBlock, 0, st>>>(n, perm, xbar, vbar, pc.dt, ab.e, ab.c, ab.f, ab.vsb);
  double4* e = ab.e;
  double4* e2 = ab.e2;
  for (int k = K - 1; k >= 0; --k) {
    if (lanes == 2) {
      k64_adj_lambda<true><<<nbp, kBlock, 0, st>>>(nv, P[k], aux[k], ab.f, ab.q1, ab.q2, pc);
      k64_adj_pos<true><<<nbp, kBlock, 0, st>>>(nv, P[k], aux[k], ab.q1, ab.q2, e, ab.c, e2, ab.f, pc);
    } else {
      k64_adj_lambda<false><<<nb, kBlock, 0, st>>>(nv, P[k], aux[k], ab.f, ab.q1, ab.q2, pc);
      k64_adj_pos<false><<<nb, kBlock, 0, st>>>(nv, P[k], aux[k], ab.q1, ab.q2, e, ab.c, e2, ab.f, pc);
    }
    double4* t = e;
    e = e2;
    e2 = t;
  }
  k64_adj_end<<<nb, kBlock, 0, st>>>(n, perm, e, ab.vsb, pc.dt, xbar, vbar);
}

}  // namespace pbf
