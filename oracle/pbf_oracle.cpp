// CPU oracle for the PhysIKA position-based-fluid step (test infrastructure:
// the checker, never the product). Built with -ffp-contract=off; see
// pbf_oracle.h for what pins each function.
#include "pbf_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace pbf_oracle {

// The reference's Vector<Real,3> wraps glm::tvec3 (Core/Vector/vector_3d.h:31).
// dot() is glm's compute_dot for tvec3: tmp = a*b; (tmp.x + tmp.y) + tmp.z
// (Extern/glm-0.9.9.7/glm/detail/func_geometric.inl:64-71), and norm() is
// glm::length = sqrt(dot(v, v)) (func_geometric.inl:95-100,
// Core/Vector/vector_3d.inl:208-212).
template <class R> static inline V3<R> sub(const V3<R>& a, const V3<R>& b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R> static inline R dot(const V3<R>& a, const V3<R>& b) {
  R tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}
template <class R> static inline R norm(const V3<R>& a) { return std::sqrt(dot(a, a)); }

// 1.0/r evaluated in double and stored to Real, as `* (1.0 / r)` does in
// K_ComputeDisplacement (DensityPBD.cu:158). For float this equals the
// correctly rounded float reciprocal (double rounding is innocuous for a
// quotient of floats).
template <class R> static inline R inv_double(R r) { return (R)(1.0 / (double)r); }

// M_PI literal of Core/Utility/cuda_utilities.h:31, cast to Real as in
// Kernel.h:43 ((Real)M_PI).
static constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// GridHash::setSpace, Framework/Topology/GridHash.cu:47-75.
//   padding = 2; ds = h; lo = lo - padding*ds; nSeg = (hi - lo_orig)/ds;
//   n = ceil(nSeg) + 1 + 2*padding.
template <class R> Grid<R> grid_setup(R h, const R lo[3], const R hi[3]) {
  Grid<R> g;
  const int padding = 2;
  g.ds = h;
  R pad = (R)padding * g.ds;
  int n[3];
  for (int c = 0; c < 3; ++c) {
    g.lo[c] = lo[c] - pad;
    R nseg = (hi[c] - lo[c]) / g.ds;
    n[c] = (int)(std::ceil(nseg) + (R)1 + (R)(2 * padding));
  }
  g.nx = n[0];
  g.ny = n[1];
  g.nz = n[2];
  g.num = (long long)g.nx * g.ny * g.nz;
  return g;
}

// GridHash::getIndex3 (GridHash.h:55-62): floor((p - lo)/ds) per axis with a
// true division. Out-of-range values are clamped to [-2, n+1] (the set of
// valid cells in the +-1 stencil is the same), NaN maps to -2.
template <class R> static inline int cell_coord(R p, R lo, R ds, int n) {
  R f = std::floor((p - lo) / ds);
  if (!(f >= (R)-2)) return -2;  // also NaN
  if (f > (R)(n + 1)) return n + 1;
  return (int)f;
}
template <class R> void cell3(const Grid<R>& g, const V3<R>& p, int out[3]) {
  out[0] = cell_coord(p.x, g.lo[0], g.ds, g.nx);
  out[1] = cell_coord(p.y, g.lo[1], g.ds, g.ny);
  out[2] = cell_coord(p.z, g.lo[2], g.ds, g.nz);
}

// GridHash::getIndex(i,j,k), GridHash.h:37-44.
template <class R> long long cell_id(const Grid<R>& g, int i, int j, int k) {
  if (i < 0 || i >= g.nx) return -1;
  if (j < 0 || j >= g.ny) return -1;
  if (k < 0 || k >= g.nz) return -1;
  return (long long)i + (long long)j * g.nx + (long long)k * g.nx * g.ny;
}

// The 27 stencil offsets in the order of NeighborQuery.cu:14-41 (centre, the
// 6 face neighbours, the 12 edge neighbours, the 8 corners).
static const int kOffsets[27][3] = {
    {0, 0, 0},   {0, 0, 1},   {0, 1, 0},   {1, 0, 0},   {0, 0, -1},  {0, -1, 0},  {-1, 0, 0},
    {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},
    {-1, 0, -1}, {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 1, 1},   {1, 1, -1},
    {1, -1, 1},  {-1, 1, 1},  {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

// SpikyKernel::Weight, Dynamics/ParticleSystem/Kernel.h:36-45.
template <class R> R spiky_w(R r, R h) {
  const R q = r / h;
  if (q > (R)1) return (R)0;
  const R d = (R)(1.0 - (double)q);
  const R hh = h * h;
  return (R)15 / ((R)kPi * hh * h) * d * d * d * (R)1;
}

// SpikyKernel::Gradient, Dynamics/ParticleSystem/Kernel.h:47-57.
template <class R> R spiky_grad(R r, R h) {
  const R q = r / h;
  if (q > (R)1) return (R)0;
  const R d = (R)(1.0 - (double)q);
  const R hh = h * h;
  return (R)-45 / ((R)kPi * hh * h) * d * d * (R)1;
}

// SummationDensity::calculateScalingFactor, SummationDensity.cu:113-138.
template <class R> R scaling_factor(R d, R H) {
  R V = d * d * d;
  R total = (R)0;
  int half_res = (int)(H / d + (R)1);
  for (int i = -half_res; i <= half_res; i++)
    for (int j = -half_res; j <= half_res; j++)
      for (int k = -half_res; k <= half_res; k++) {
        R x = (R)i * d;
        R y = (R)j * d;
        R z = (R)k * d;
        R r = std::sqrt(x * x + y * y + z * z);
        total += V * spiky_w(r, H);
      }
  return (R)(1.0 / (double)total);
}

// GridHash::construct (GridHash.cu:298-323) followed by
// NeighborQuery::queryNeighborDynamic (NeighborQuery.cu:450-478) with
// K_CalNeighborSize (:245-276) and K_GetNeighborElements (:337-367).
// Particles outside the grid are not inserted (GridHash.cu:85,194) but still
// query. Ids inside a cell are kept in ascending order (the reference's order
// comes from atomics and is not defined; the neighbor SET is the same).
template <class R>
void neighbor_query(const Grid<R>& g, const V3<R>* x, int n, R h, CSR& out) {
  std::vector<long long> cid(n);
  std::vector<int> c3(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    cell3(g, x[i], &c3[3 * (size_t)i]);
    cid[i] = cell_id(g, c3[3 * i], c3[3 * i + 1], c3[3 * i + 2]);
  }
  // counting sort of in-grid ids: index[c] = exclusive scan of counts
  std::vector<int> index((size_t)g.num + 1, 0);
  for (int i = 0; i < n; ++i)
    if (cid[i] >= 0) index[(size_t)cid[i] + 1]++;
  for (long long c = 0; c < g.num; ++c) index[(size_t)c + 1] += index[(size_t)c];
  std::vector<int> ids(index[(size_t)g.num]);
  {
    std::vector<int> fill(index.begin(), index.end() - 1);
    for (int i = 0; i < n; ++i)
      if (cid[i] >= 0) ids[fill[(size_t)cid[i]]++] = i;
  }
  std::vector<int> cnt(n);
  auto visit = [&](int i, auto&& emit) {
    const int* gi = &c3[3 * (size_t)i];
    for (int c = 0; c < 27; ++c) {
      long long cell = cell_id(g, gi[0] + kOffsets[c][0], gi[1] + kOffsets[c][1], gi[2] + kOffsets[c][2]);
      if (cell < 0) continue;
      for (int k = index[(size_t)cell]; k < index[(size_t)cell + 1]; ++k) {
        int j = ids[k];
        R dij = norm(sub(x[i], x[j]));
        if (dij < h) emit(j);
      }
    }
  };
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    int c = 0;
    visit(i, [&](int) { ++c; });
    cnt[i] = c;
  }
  out.offsets.assign((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) out.offsets[(size_t)i + 1] = out.offsets[(size_t)i] + cnt[i];
  out.elems.assign((size_t)out.offsets[(size_t)n], 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    int w = out.offsets[(size_t)i];
    visit(i, [&](int j) { out.elems[(size_t)w++] = j; });
  }
}

// NeighborQuery::queryNeighborFixed with K_ComputeNeighborFixed
// (NeighborQuery.cu:514-581): candidates visited in the 27-offset order,
// the first `limit` neighbors fill the row, each later one replaces the
// farthest kept one when closer (first farthest on ties). counts[i] = kept,
// elems[limit*i + k] in the row's final order.
template <class R>
void neighbor_query_fixed(const Grid<R>& g, const V3<R>* x, int n, R h, int limit, int* counts, int* elems) {
  CSR nb;
  neighbor_query(g, x, n, h, nb);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    std::vector<R> dist((size_t)limit);
    int* ids = elems + (size_t)limit * i;
    int counter = 0;
    for (int e = nb.offsets[i]; e < nb.offsets[i + 1]; ++e) {
      const int j = nb.elems[e];
      const R d = (R)(float)norm(sub(x[i], x[j]));  // float d_ij (NeighborQuery.cu:547)
      if (counter < limit) {
        ids[counter] = j;
        dist[counter] = d;
        ++counter;
      } else {
        int max_id = 0;
        R max_d = dist[0];
        for (int k = 1; k < limit; ++k)
          if (max_d < dist[k]) {
            max_d = dist[k];
            max_id = k;
          }
        if (d < dist[max_id]) {
          dist[max_id] = d;
          ids[max_id] = j;
        }
      }
    }
    counts[i] = counter;
  }
}

// K_ComputeDensity, SummationDensity.cu:13-36 (mass = m_factor * particle
// mass, SummationDensity.cu:104-109).
template <class R>
void density(const V3<R>* x, int n, const CSR& nb, R h, R mass, R* rho) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    R rho_i = (R)0;
    for (int k = nb.offsets[i]; k < nb.offsets[i + 1]; ++k) {
      int j = nb.elems[k];
      R r = norm(sub(x[i], x[j]));
      rho_i += mass * spiky_w(r, h);
    }
    rho[i] = rho_i;
  }
}

// K_ComputeLambdas, DensityPBD.cu:46-90; with minv (inverse masses) the
// massInv variant, DensityPBD.cu:93-131: sum_j |g|^2 m_j^-1 + |sum g|^2 m_i^-1.
template <class R>
void lambdas(const Params<R>& p, const R* rho, const V3<R>* x, int n, const CSR& nb, R* lam, const R* minv) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    R lamda_i = (R)0;
    V3<R> grad_ci{0, 0, 0};
    for (int k = nb.offsets[i]; k < nb.offsets[i + 1]; ++k) {
      int j = nb.elems[k];
      V3<R> dx = sub(x[i], x[j]);
      R r = norm(dx);
      if (r > p.pair_eps) {
        R gr = spiky_grad(r, p.h);
        R ir = (R)1 / r;
        V3<R> gv{(dx.x * gr) * ir, (dx.y * gr) * ir, (dx.z * gr) * ir};
        grad_ci.x += gv.x;
        grad_ci.y += gv.y;
        grad_ci.z += gv.z;
        lamda_i += minv ? dot(gv, gv) * minv[j] : dot(gv, gv);
      }
    }
    lamda_i += minv ? dot(grad_ci, grad_ci) * minv[i] : dot(grad_ci, grad_ci);
    R rho_i = rho[i];
    lamda_i = -(rho_i - p.lam_rest) / (lamda_i + p.lam_eps);
    lam[i] = lamda_i > (R)0 ? (R)0 : lamda_i;
  }
}

// dp_ij of K_ComputeDisplacement, DensityPBD.cu:158:
//   10*(x_i - x_j)*(lambda_i + lambda_j)*Gradient(r)*(1.0/r), left to right.
template <class R>
static inline bool pair_dp(const Params<R>& p, const V3<R>& xi, const V3<R>& xj, R li, R lj, V3<R>& dp) {
  V3<R> dx = sub(xi, xj);
  R r = norm(dx);
  if (!(r > p.pair_eps)) return false;
  R s = li + lj;
  R gr = spiky_grad(r, p.h);
  R ir = inv_double(r);
  dp.x = (((dx.x * p.corr) * s) * gr) * ir;
  dp.y = (((dx.y * p.corr) * s) * gr) * ir;
  dp.z = (((dx.z * p.corr) * s) * gr) * ir;
  return true;
}

// Gather form of K_ComputeDisplacement for a symmetric neighbor list: the
// reference adds +dp_ij to i and -dp_ij to j from both endpoints
// (DensityPBD.cu:161-174) and dp_ji == -dp_ij bit for bit, so
// dPos_i = 2 * sum_j dp_ij.
template <class R>
void displacement_gather(const Params<R>& p, const R* lam, const V3<R>* x, int n, const CSR& nb,
                         V3<R>* dpos) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    V3<R> acc{0, 0, 0};
    for (int k = nb.offsets[i]; k < nb.offsets[i + 1]; ++k) {
      int j = nb.elems[k];
      V3<R> dp;
      if (pair_dp(p, x[i], x[j], lam[i], lam[j], dp)) {
        acc.x += dp.x;
        acc.y += dp.y;
        acc.z += dp.z;
      }
    }
    dpos[i] = {(R)2 * acc.x, (R)2 * acc.y, (R)2 * acc.z};
  }
}

// Literal sequential emulation of the reference's pairwise atomic scatter
// (DensityPBD.cu:152-176): for every i in order, for every listed j,
// dPos[i] += dp_ij and dPos[j] -= dp_ij. Exact for any list, including the
// asymmetric lists produced when particles leave the grid. With minv, the
// massInv variant (DensityPBD.cu:182-224): dp_ji = -dp_ij m_j^-1 to j and
// dp_ij m_i^-1 to i.
template <class R>
void displacement_atomic_order(const Params<R>& p, const R* lam, const V3<R>* x, int n, const CSR& nb,
                               V3<R>* dpos, const R* minv) {
  for (int i = 0; i < n; ++i) dpos[i] = {0, 0, 0};  // m_deltaPos.reset(), :315
  for (int i = 0; i < n; ++i) {
    for (int k = nb.offsets[i]; k < nb.offsets[i + 1]; ++k) {
      int j = nb.elems[k];
      V3<R> dp;
      if (pair_dp(p, x[i], x[j], lam[i], lam[j], dp)) {
        V3<R> dj{-dp.x, -dp.y, -dp.z};
        if (minv) {
          dj = {dj.x * minv[j], dj.y * minv[j], dj.z * minv[j]};
          dp = {dp.x * minv[i], dp.y * minv[i], dp.z * minv[i]};
        }
        dpos[i].x += dp.x;
        dpos[j].x += dj.x;
        dpos[i].y += dp.y;
        dpos[j].y += dj.y;
        dpos[i].z += dp.z;
        dpos[j].z += dj.z;
      }
    }
  }
}

template <class R> static bool list_symmetric(const CSR& nb, int n) {
  // i in N_j for every j in N_i (rows need not be sorted)
  std::vector<std::vector<int>> rows(n);
  for (int i = 0; i < n; ++i) {
    rows[i].assign(nb.elems.begin() + nb.offsets[i], nb.elems.begin() + nb.offsets[i + 1]);
    std::sort(rows[i].begin(), rows[i].end());
  }
  for (int i = 0; i < n; ++i)
    for (int j : rows[i])
      if (!std::binary_search(rows[j].begin(), rows[j].end(), i)) return false;
  return true;
}

// DensityPBD::constrain (DensityPBD.cu:271-302), takeOneIteration (:306-365)
// with K_UpdatePosition (:227-237), and DP_UpdateVelocity (:368-392).
// rho receives the density of the last iteration, computed before its Δx
// (outDensity is SummationDensity's output, DensityPBD.cu:259).
template <class R>
void pbd_constrain(const Params<R>& p, V3<R>* x, V3<R>* v, int n, const CSR& nb, R* rho, const R* minv) {
  std::vector<V3<R>> x_old(x, x + n);
  std::vector<R> lam(n);
  std::vector<V3<R>> dpos(n);
  std::vector<R> rho_tmp(n);
  R* rho_w = rho ? rho : rho_tmp.data();
  const R mass = scaling_factor(p.d, p.h) * (p.rho0 * (p.d * p.d * p.d));
  const bool gather = !minv && (p.dx_mode == 1 || (p.dx_mode == 0 && list_symmetric<R>(nb, n)));
  for (int it = 0; it < p.iters; ++it) {
    density(x, n, nb, p.h, mass, rho_w);
    lambdas(p, rho_w, x, n, nb, lam.data(), minv);
    if (gather)
      displacement_gather(p, lam.data(), x, n, nb, dpos.data());
    else
      displacement_atomic_order(p, lam.data(), x, n, nb, dpos.data(), minv);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      x[i].x += dpos[i].x;
      x[i].y += dpos[i].y;
      x[i].z += dpos[i].z;
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    v[i].x += (x[i].x - x_old[i].x) / p.dt;
    v[i].y += (x[i].y - x_old[i].y) / p.dt;
    v[i].z += (x[i].z - x_old[i].z) / p.dt;
  }
}

// ParticleIntegrator::integrate (ParticleIntegrator.cu:129-138):
// K_UpdateVelocity (:60-70) v += dt*(f + g), then K_UpdatePosition (:103-112)
// x += dt*v. A null force means the zeroed ForceDensity of begin() (:32).
template <class R>
void integrate(const Params<R>& p, V3<R>* x, V3<R>* v, const V3<R>* f, int n) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    V3<R> fi = f ? f[i] : V3<R>{0, 0, 0};
    v[i].x += (fi.x + p.g[0]) * p.dt;
    v[i].y += (fi.y + p.g[1]) * p.dt;
    v[i].z += (fi.z + p.g[2]) * p.dt;
    x[i].x += v[i].x * p.dt;
    x[i].y += v[i].y * p.dt;
    x[i].z += v[i].z * p.dt;
  }
}

// One frame of a scene holding StaticBoundary (the box, flags & 2) and the
// fluid: the boundary node advances first (SceneGraph::takeOneFrame traverses
// top-down, Framework/Framework/SceneGraph.cpp:101-114, Node.cpp:393-405),
// then PositionBasedFluidModel::step (PositionBasedFluidModel.cpp:83-101):
// begin (force := 0) -> NeighborQuery::compute on x^n -> integrate ->
// DensityPBD::constrain -> ImplicitViscosity::constrain (flags & 1).
template <class R>
void step(const Params<R>& p, V3<R>* x, V3<R>* v, R* rho, int n, CSR* nb_out, const SDF<R>* sdf) {
  if (p.flags & 2) {
    SDF<R> local_sdf;
    if (!sdf) {
      local_sdf = box_sdf(p.box_lo, p.box_hi, p.box_spacing, true);
      sdf = &local_sdf;
    }
    box_constrain(p, *sdf, x, v, n);
  }
  Grid<R> g = grid_setup(p.h, p.lo, p.hi);
  CSR local;
  CSR& nb = nb_out ? *nb_out : local;
  neighbor_query(g, x, n, p.h, nb);
  // The list is symmetric exactly when every particle lies inside the grid
  // (out-of-grid particles query but are not inserted, GridHash.cu:85,194).
  Params<R> q = p;
  if (q.dx_mode == 0) {
    q.dx_mode = 1;
    for (int i = 0; i < n; ++i) {
      int c[3];
      cell3(g, x[i], c);
      if (cell_id(g, c[0], c[1], c[2]) < 0) {
        q.dx_mode = 2;
        break;
      }
    }
  }
  integrate<R>(q, x, v, nullptr, n);
  pbd_constrain(q, x, v, n, nb, rho);
  if (p.flags & 1) viscosity(p, x, v, n, nb);
}

// ---------------------------------------------------------------------------
// ImplicitViscosity (Dynamics/ParticleSystem/ImplicitViscosity.cu).

// VB_VisWeight, ImplicitViscosity.cu:9-18: 45/(13 pi h^3) (1 - r/h), 0 beyond h.
template <class R> R vis_weight(R r, R h) {
  const R q = r / h;
  if (q > (R)1) return (R)0;
  const R d = (R)1 - q;
  const R RR = h * h;
  return (R)45 / ((R)13 * (R)kPi * RR * h) * d;
}

// ImplicitViscosity::constrain (:99-128) with K_ApplyViscosity (:20-62):
// Jacobi iterations v_i = v_old,i/(1+b) + (sum_j w_ij v_j / sum_j w_ij) b/(1+b),
// b = dt mu / h (0 when sum_j w_ij < EPSILON), pairs with r > EPSILON.
template <class R>
void viscosity(const Params<R>& p, const V3<R>* x, V3<R>* v, int n, const CSR& nb) {
  std::vector<V3<R>> v_old(v, v + n), v_buf(n);
  for (int t = 0; t < p.vis_iters; ++t) {
    std::copy(v, v + n, v_buf.begin());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      V3<R> dv{0, 0, 0};
      R tw = (R)0;
      for (int e = nb.offsets[i]; e < nb.offsets[i + 1]; ++e) {
        const int j = nb.elems[e];
        const R r = norm(sub(x[i], x[j]));
        if (r > p.pair_eps) {
          const R w = vis_weight(r, p.h);
          tw += w;
          dv.x += w * v_buf[j].x;
          dv.y += w * v_buf[j].y;
          dv.z += w * v_buf[j].z;
        }
      }
      R b = p.dt * p.mu / p.h;
      b = tw < p.pair_eps ? (R)0 : b;
      tw = tw < p.pair_eps ? (R)1 : tw;
      dv = {dv.x / tw, dv.y / tw, dv.z / tw};
      const R ob = (R)1 + b;
      v[i] = {v_old[i].x / ob + dv.x * b / ob, v_old[i].y / ob + dv.y * b / ob, v_old[i].z / ob + dv.z * b / ob};
    }
  }
}

// ---------------------------------------------------------------------------
// Box boundary: DistanceField3D (Framework/Topology/DistanceField3D.cu/.h) and
// K_ConstrainSDF (Dynamics/ParticleSystem/BoundaryConstraint.cu:30-67).
// CUDA's min/max on floats return the non-NaN operand, as std::fmin/fmax do.

template <class R> static inline R seg_distance(const V3<R>& pos, const V3<R>& lo, const V3<R>& hi) {
  // DistanceToSegment, DistanceField3D.cu:14-32
  V3<R> seg = sub(hi, lo);
  const V3<R> e1 = sub(pos, lo), e2 = sub(pos, hi);
  if (dot(e1, seg) < (R)0) return norm(e1);
  const V3<R> nseg{-seg.x, -seg.y, -seg.z};
  if (dot(e2, nseg) < (R)0) return norm(e2);
  const R length1 = dot(e1, e1);
  // Vector::normalize (Core/Vector/vector_3d.inl:220-224): glm::normalize =
  // x * (1/sqrt(dot(x, x))) when |x| > FLT_EPSILON, else 0
  const R len = norm(seg);
  if (len > std::numeric_limits<R>::epsilon()) {
    const R is = (R)1 / std::sqrt(dot(seg, seg));
    seg = {seg.x * is, seg.y * is, seg.z * is};
  } else {
    seg = {0, 0, 0};
  }
  const R length2 = dot(e1, seg);
  return std::sqrt(length1 - length2 * length2);
}

template <class R>
static inline R square_distance(const V3<R>& pos, const V3<R>& lo, const V3<R>& hi, int axis) {
  // DistanceToSqure, DistanceField3D.cu:34-88
  V3<R> n, c1, c2, c3, c4;
  R loc[2], hic[2], pp[2];
  if (axis == 0) {
    c1 = {lo.x, lo.y, lo.z}; c2 = {lo.x, hi.y, lo.z}; c3 = {lo.x, hi.y, hi.z}; c4 = {lo.x, lo.y, hi.z};
    n = {1, 0, 0};
    loc[0] = lo.y; loc[1] = lo.z; hic[0] = hi.y; hic[1] = hi.z; pp[0] = pos.y; pp[1] = pos.z;
  } else if (axis == 1) {
    c1 = {lo.x, lo.y, lo.z}; c2 = {lo.x, lo.y, hi.z}; c3 = {hi.x, lo.y, hi.z}; c4 = {hi.x, lo.y, lo.z};
    n = {0, 1, 0};
    loc[0] = lo.x; loc[1] = lo.z; hic[0] = hi.x; hic[1] = hi.z; pp[0] = pos.x; pp[1] = pos.z;
  } else {
    c1 = {lo.x, lo.y, lo.z}; c2 = {hi.x, lo.y, lo.z}; c3 = {hi.x, hi.y, lo.z}; c4 = {lo.x, hi.y, lo.z};
    n = {0, 0, 1};
    loc[0] = lo.x; loc[1] = lo.y; hic[0] = hi.x; hic[1] = hi.y; pp[0] = pos.x; pp[1] = pos.y;
  }
  const R d1 = seg_distance(pos, c1, c2), d2 = seg_distance(pos, c2, c3);
  const R d3 = seg_distance(pos, c3, c4), d4 = seg_distance(pos, c4, c1);
  const R d5 = std::fabs(dot(n, sub(pos, c1)));
  if (pp[0] < hic[0] && pp[0] > loc[0] && pp[1] < hic[1] && pp[1] > loc[1]) return d5;
  return std::fmin(std::fmin(d1, d2), std::fmin(d3, d4));
}

template <class R> static inline R box_distance(const V3<R>& pos, const R lo[3], const R hi[3]) {
  // DistanceToBox, DistanceField3D.cu:90-129 (the corner norms it computes
  // are unused there)
  const V3<R> c0{lo[0], lo[1], lo[2]}, c1{hi[0], lo[1], lo[2]}, c2{hi[0], hi[1], lo[2]}, c3{lo[0], hi[1], lo[2]};
  const V3<R> c4{lo[0], lo[1], hi[2]}, c5{hi[0], lo[1], hi[2]}, c6{hi[0], hi[1], hi[2]}, c7{lo[0], hi[1], hi[2]};
  if (pos.x < hi[0] && pos.x > lo[0] && pos.y < hi[1] && pos.y > lo[1] && pos.z < hi[2] && pos.z > lo[2]) {
    const R dx = std::fmin(std::fabs(pos.x - hi[0]), std::fabs(pos.x - lo[0]));
    const R dy = std::fmin(std::fabs(pos.y - hi[1]), std::fabs(pos.y - lo[1]));
    const R dz = std::fmin(std::fabs(pos.z - hi[2]), std::fabs(pos.z - lo[2]));
    return std::fmin(std::fmin(dx, dy), dz);
  }
  const R x1 = square_distance(pos, c0, c7, 0), x2 = square_distance(pos, c1, c6, 0);
  const R y1 = square_distance(pos, c0, c5, 1), y2 = square_distance(pos, c3, c6, 1);
  const R z1 = square_distance(pos, c0, c2, 2), z2 = square_distance(pos, c4, c6, 2);
  return -std::fmin(std::fmin(std::fmin(x1, x2), std::fmin(y1, y2)), std::fmin(z1, z2));
}

template <class R> SDF<R> box_sdf(const R lo[3], const R hi[3], R spacing, bool inverted) {
  SDF<R> f;
  // setCube: n = floor((hi - lo)/distance); setSpace(lo - 5d, hi + 5d, n + 10)
  int nb[3];
  R p0[3], p1[3];
  for (int a = 0; a < 3; ++a) {
    nb[a] = (int)std::floor((hi[a] - lo[a]) / spacing) + 10;
    const R pad = (R)5 * spacing;
    p0[a] = lo[a] - pad;
    p1[a] = hi[a] + pad;
  }
  // setSpace: h = (p1 - p0) * (1.0/(n+1)) with the reciprocal in double,
  // n + 1 samples per axis
  for (int a = 0; a < 3; ++a) {
    f.left[a] = p0[a];
    f.hs[a] = (p1[a] - p0[a]) * (R)(1.0 / (double)(R)(nb[a] + 1));
    f.inv_h[a] = (R)(1.0 / (double)f.hs[a]);
  }
  f.nx = nb[0] + 1;
  f.ny = nb[1] + 1;
  f.nz = nb[2] + 1;
  f.inverted = inverted;
  f.d.assign((size_t)f.nx * f.ny * f.nz, (R)0);
  const R sign = inverted ? (R)1 : (R)-1;
  // K_DistanceFieldToBox: p = start + (i,j,k)*h
#pragma omp parallel for schedule(static)
  for (int k = 0; k < f.nz; ++k)
    for (int j = 0; j < f.ny; ++j)
      for (int i = 0; i < f.nx; ++i) {
        const V3<R> q{f.left[0] + (R)i * f.hs[0], f.left[1] + (R)j * f.hs[1], f.left[2] + (R)k * f.hs[2]};
        f.d[(size_t)i + (size_t)f.nx * ((size_t)j + (size_t)f.ny * k)] = sign * box_distance(q, lo, hi);
      }
  return f;
}

template <class R> void sdf_distance(const SDF<R>& f, const V3<R>& p, R& d, V3<R>& normal) {
  const R fx = (p.x - f.left[0]) * f.inv_h[0];
  const R fy = (p.y - f.left[1]) * f.inv_h[1];
  const R fz = (p.z - f.left[2]) * f.inv_h[2];
  const R ffx = std::floor(fx), ffy = std::floor(fy), ffz = std::floor(fz);
  // out of the field (also NaN): +-100000 and no normal
  if (!(ffx >= (R)0 && ffx < (R)(f.nx - 1) && ffy >= (R)0 && ffy < (R)(f.ny - 1) && ffz >= (R)0 &&
        ffz < (R)(f.nz - 1))) {
    d = f.inverted ? (R)-100000 : (R)100000;
    normal = {0, 0, 0};
    return;
  }
  const int i = (int)ffx, j = (int)ffy, k = (int)ffz;
  const R alpha = fx - (R)i, beta = fy - (R)j, gamma = fz - (R)k;
  auto D = [&](int a, int b, int c) { return f.d[(size_t)a + (size_t)f.nx * ((size_t)b + (size_t)f.ny * c)]; };
  auto lerp = [](R a, R b, R t) { return ((R)1 - t) * a + t * b; };
  const R d000 = D(i, j, k), d100 = D(i + 1, j, k), d010 = D(i, j + 1, k), d110 = D(i + 1, j + 1, k);
  const R d001 = D(i, j, k + 1), d101 = D(i + 1, j, k + 1), d011 = D(i, j + 1, k + 1), d111 = D(i + 1, j + 1, k + 1);
  const R dx00 = lerp(d000, d100, alpha), dx10 = lerp(d010, d110, alpha), dxy0 = lerp(dx00, dx10, beta);
  const R dx01 = lerp(d001, d101, alpha), dx11 = lerp(d011, d111, alpha), dxy1 = lerp(dx01, dx11, beta);
  const R d0y0 = lerp(d000, d010, beta), d0y1 = lerp(d001, d011, beta), d0yz = lerp(d0y0, d0y1, gamma);
  const R d1y0 = lerp(d100, d110, beta), d1y1 = lerp(d101, d111, beta), d1yz = lerp(d1y0, d1y1, gamma);
  const R dx0z = lerp(dx00, dx01, gamma), dx1z = lerp(dx10, dx11, gamma);
  V3<R> nrm{d0yz - d1yz, dx0z - dx1z, dxy0 - dxy1};
  const R l = norm(nrm);
  if (l < (R)0.0001f) {
    nrm = {0, 0, 0};
  } else if (l > std::numeric_limits<R>::epsilon()) {
    const R is = (R)1 / std::sqrt(dot(nrm, nrm));
    nrm = {nrm.x * is, nrm.y * is, nrm.z * is};
  }
  normal = nrm;
  d = ((R)1 - gamma) * dxy0 + gamma * dxy1;
}

float box_jitter(int id) {
  uint64_t z = ((uint64_t)(uint32_t)id + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const float u = (float)((z >> 40) + 1) * (1.0f / 16777216.0f);  // (0, 1]
  return 0.0001f * u;
}

template <class R>
void box_constrain(const Params<R>& p, const SDF<R>& f, V3<R>* x, V3<R>* v, int n) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    V3<R> pos = x[i], vec = v[i];
    R dist;
    V3<R> nrm;
    sdf_distance(f, pos, dist, nrm);
    if (dist <= (R)0) {
      const R olddist = -dist;
      dist = (R)box_jitter(i);
      const R s = olddist + dist;
      pos = {pos.x - s * nrm.x, pos.y - s * nrm.y, pos.z - s * nrm.z};
      const R vn = dot(vec, nrm);
      V3<R> vnor{vn * nrm.x, vn * nrm.y, vn * nrm.z};
      const V3<R> vtan{vec.x - vnor.x, vec.y - vnor.y, vec.z - vnor.z};
      if (vn > (R)0) vnor = {-vnor.x, -vnor.y, -vnor.z};
      const R keep = (R)1 - p.nfric;
      vnor = {vnor.x * keep, vnor.y * keep, vnor.z * keep};
      vec = {vnor.x + vtan.x, vnor.y + vtan.y, vnor.z + vtan.z};
      const R damp = std::pow((R)2.71828182845904523536, -p.dt * p.tfric);
      vec = {vec.x * damp, vec.y * damp, vec.z * damp};
    }
    x[i] = pos;
    v[i] = vec;
  }
}

// ---------------------------------------------------------------------------
// Adjoint of one step (not in the reference). Forward quantities of every
// PBD iteration are recomputed in the order of pbd_constrain; the reverse
// sweep treats each listed ordered pair (i, j) as the reference's scatter
// does: dp_ij = corr (lambda_i + lambda_j) g(x_i - x_j) is added to i and
// subtracted from j, with g(d) = Gradient(|d|) d / |d|.

// g(d) = phi(r) d/r, phi = Gradient (Kernel.h:47-57); returns J(d) w with
// J = dg/dd = phi'(r) dh dh^T + phi/r (I - dh dh^T), dh = d/r.
template <class R>
static inline V3<R> jac_g(const V3<R>& d, R r, R h, const V3<R>& w) {
  const R q = r / h;
  if (q > (R)1) return {0, 0, 0};
  const R c = (R)-45 / ((R)kPi * h * h * h);
  const R phi = c * (1 - q) * (1 - q);
  const R dphi = c * (R)-2 * (1 - q) / h;
  const R ir = (R)1 / r;
  const V3<R> dh{d.x * ir, d.y * ir, d.z * ir};
  const R dw = dh.x * w.x + dh.y * w.y + dh.z * w.z;
  const R a = phi * ir, b = (dphi - phi * ir) * dw;
  return {a * w.x + b * dh.x, a * w.y + b * dh.y, a * w.z + b * dh.z};
}

// d/dd of W(|d|) (Kernel.h:36-45): W'(r) d/r, zero at d = 0.
template <class R> static inline V3<R> grad_w(const V3<R>& d, R r, R h) {
  const R q = r / h;
  if (q > (R)1 || !(r > (R)0)) return {0, 0, 0};
  const R dw = (R)-45 / ((R)kPi * h * h * h * h) * (1 - q) * (1 - q);
  return {dw * d.x / r, dw * d.y / r, dw * d.z / r};
}

template <class R>
void step_adjoint(const Params<R>& p, const V3<R>* x, const V3<R>* v, int n, V3<R>* xbar, V3<R>* vbar) {
  const Grid<R> g = grid_setup(p.h, p.lo, p.hi);
  CSR nb;
  neighbor_query(g, x, n, p.h, nb);
  const int K = p.iters;
  const R mass = scaling_factor(p.d, p.h) * (p.rho0 * (p.d * p.d * p.d));
  // forward: integrate, then the PBD iterations with their per-iteration state
  std::vector<V3<R>> xs(x, x + n), vs(v, v + n);
  integrate<R>(p, xs.data(), vs.data(), nullptr, n);
  std::vector<std::vector<V3<R>>> P(K + 1), G(K);
  std::vector<std::vector<R>> rho(K), lamu(K), S(K);
  P[0] = xs;
  Params<R> q = p;
  q.dx_mode = 2;  // the scatter form, valid for any list
  for (int k = 0; k < K; ++k) {
    const V3<R>* pk = P[k].data();
    rho[k].assign(n, 0);
    lamu[k].assign(n, 0);
    S[k].assign(n, 0);
    G[k].assign(n, {0, 0, 0});
    density(pk, n, nb, p.h, mass, rho[k].data());
    std::vector<R> lam(n);
    for (int i = 0; i < n; ++i) {
      R s = 0;
      V3<R> gi{0, 0, 0};
      for (int e = nb.offsets[i]; e < nb.offsets[i + 1]; ++e) {
        const V3<R> d = sub(pk[i], pk[nb.elems[e]]);
        const R r = norm(d);
        if (!(r > p.pair_eps)) continue;
        const R gr = spiky_grad(r, p.h);
        const V3<R> gv{d.x * gr / r, d.y * gr / r, d.z * gr / r};
        gi.x += gv.x;
        gi.y += gv.y;
        gi.z += gv.z;
        s += dot(gv, gv);
      }
      s += dot(gi, gi);
      S[k][i] = s;
      G[k][i] = gi;
      lamu[k][i] = -(rho[k][i] - p.lam_rest) / (s + p.lam_eps);
      lam[i] = lamu[k][i] > 0 ? 0 : lamu[k][i];
    }
    std::vector<V3<R>> dpos(n);
    displacement_atomic_order(q, lam.data(), pk, n, nb, dpos.data());
    P[k + 1] = P[k];
    for (int i = 0; i < n; ++i) {
      P[k + 1][i].x += dpos[i].x;
      P[k + 1][i].y += dpos[i].y;
      P[k + 1][i].z += dpos[i].z;
    }
  }
  // reverse: v' = v* + (P_K - x*)/dt, x' = P_K
  std::vector<V3<R>> a(n), xsb(n), vsb(n);
  for (int i = 0; i < n; ++i) {
    a[i] = {xbar[i].x + vbar[i].x / p.dt, xbar[i].y + vbar[i].y / p.dt, xbar[i].z + vbar[i].z / p.dt};
    xsb[i] = {-vbar[i].x / p.dt, -vbar[i].y / p.dt, -vbar[i].z / p.dt};
    vsb[i] = vbar[i];
  }
  for (int k = K - 1; k >= 0; --k) {
    const V3<R>* pk = P[k].data();
    std::vector<R> lam(n), lamb(n, 0), rhob(n, 0), Sb(n, 0);
    for (int i = 0; i < n; ++i) lam[i] = lamu[k][i] > 0 ? 0 : lamu[k][i];
    // lambda adjoint: each pair's term (a_i - a_j) . dp_ij
    for (int i = 0; i < n; ++i)
      for (int e = nb.offsets[i]; e < nb.offsets[i + 1]; ++e) {
        const int j = nb.elems[e];
        const V3<R> d = sub(pk[i], pk[j]);
        const R r = norm(d);
        if (!(r > p.pair_eps)) continue;
        const R gr = spiky_grad(r, p.h);
        const V3<R> da = sub(a[i], a[j]);
        const R t = p.corr * gr / r * dot(da, d);
        lamb[i] += t;
        lamb[j] += t;
      }
    for (int i = 0; i < n; ++i) {
      if (!(lamu[k][i] < 0)) continue;  // lambda clamped to 0
      const R den = S[k][i] + p.lam_eps;
      rhob[i] = -lamb[i] / den;
      Sb[i] = lamb[i] * (rho[k][i] - p.lam_rest) / (den * den);
    }
    std::vector<V3<R>> pb(a);
    for (int i = 0; i < n; ++i)
      for (int e = nb.offsets[i]; e < nb.offsets[i + 1]; ++e) {
        const int j = nb.elems[e];
        const V3<R> d = sub(pk[i], pk[j]);
        const R r = norm(d);
        const V3<R> gw = grad_w(d, r, p.h);
        V3<R> db{rhob[i] * mass * gw.x, rhob[i] * mass * gw.y, rhob[i] * mass * gw.z};
        if (r > p.pair_eps) {
          const R gr = spiky_grad(r, p.h);
          const V3<R> gv{d.x * gr / r, d.y * gr / r, d.z * gr / r};
          const R cl = p.corr * (lam[i] + lam[j]);
          const V3<R> w{cl * (a[i].x - a[j].x) + 2 * Sb[i] * (gv.x + G[k][i].x),
                        cl * (a[i].y - a[j].y) + 2 * Sb[i] * (gv.y + G[k][i].y),
                        cl * (a[i].z - a[j].z) + 2 * Sb[i] * (gv.z + G[k][i].z)};
          const V3<R> jw = jac_g(d, r, p.h, w);
          db.x += jw.x;
          db.y += jw.y;
          db.z += jw.z;
        }
        pb[i].x += db.x;
        pb[i].y += db.y;
        pb[i].z += db.z;
        pb[j].x -= db.x;
        pb[j].y -= db.y;
        pb[j].z -= db.z;
      }
    a.swap(pb);
  }
  // x* = x + v* dt, v* = v + g dt
  for (int i = 0; i < n; ++i) {
    const V3<R> xb{xsb[i].x + a[i].x, xsb[i].y + a[i].y, xsb[i].z + a[i].z};
    xbar[i] = xb;
    vbar[i] = {vsb[i].x + xb.x * p.dt, vsb[i].y + xb.y * p.dt, vsb[i].z + xb.z * p.dt};
  }
}

template <class R>
void rollout_adjoint(const Params<R>& p, const V3<R>* x0, const V3<R>* v0, int n, int nsteps, V3<R>* xbar,
                     V3<R>* vbar) {
  std::vector<std::vector<V3<R>>> X(nsteps), Vv(nsteps);
  std::vector<V3<R>> x(x0, x0 + n), v(v0, v0 + n);
  for (int s = 0; s < nsteps; ++s) {
    X[s] = x;
    Vv[s] = v;
    step(p, x.data(), v.data(), (R*)nullptr, n, (CSR*)nullptr);
  }
  for (int s = nsteps - 1; s >= 0; --s) step_adjoint(p, X[s].data(), Vv[s].data(), n, xbar, vbar);
}

#define PBF_ORACLE_INSTANTIATE(R)                                                                  \
  template Grid<R> grid_setup<R>(R, const R*, const R*);                                           \
  template void cell3<R>(const Grid<R>&, const V3<R>&, int*);                                      \
  template long long cell_id<R>(const Grid<R>&, int, int, int);                                    \
  template R spiky_w<R>(R, R);                                                                     \
  template R spiky_grad<R>(R, R);                                                                  \
  template R scaling_factor<R>(R, R);                                                              \
  template void neighbor_query<R>(const Grid<R>&, const V3<R>*, int, R, CSR&);                     \
  template void density<R>(const V3<R>*, int, const CSR&, R, R, R*);                               \
  template void lambdas<R>(const Params<R>&, const R*, const V3<R>*, int, const CSR&, R*, const R*); \
  template void neighbor_query_fixed<R>(const Grid<R>&, const V3<R>*, int, R, int, int*, int*);     \
  template void displacement_gather<R>(const Params<R>&, const R*, const V3<R>*, int, const CSR&,  \
                                       V3<R>*);                                                    \
  template void displacement_atomic_order<R>(const Params<R>&, const R*, const V3<R>*, int,        \
                                             const CSR&, V3<R>*, const R*);                        \
  template void pbd_constrain<R>(const Params<R>&, V3<R>*, V3<R>*, int, const CSR&, R*, const R*); \
  template void integrate<R>(const Params<R>&, V3<R>*, V3<R>*, const V3<R>*, int);                 \
  template void step<R>(const Params<R>&, V3<R>*, V3<R>*, R*, int, CSR*, const SDF<R>*);         \
  template R vis_weight<R>(R, R);                                                                  \
  template void viscosity<R>(const Params<R>&, const V3<R>*, V3<R>*, int, const CSR&);             \
  template SDF<R> box_sdf<R>(const R*, const R*, R, bool);                                         \
  template void sdf_distance<R>(const SDF<R>&, const V3<R>&, R&, V3<R>&);                          \
  template void box_constrain<R>(const Params<R>&, const SDF<R>&, V3<R>*, V3<R>*, int);            \
  template void step_adjoint<R>(const Params<R>&, const V3<R>*, const V3<R>*, int, V3<R>*, V3<R>*); \
  template void rollout_adjoint<R>(const Params<R>&, const V3<R>*, const V3<R>*, int, int, V3<R>*,  \
                                   V3<R>*);

PBF_ORACLE_INSTANTIATE(float)
PBF_ORACLE_INSTANTIATE(double)

}  // namespace pbf_oracle
