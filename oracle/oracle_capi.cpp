// C entry points of the CPU oracle, loaded with ctypes by tests/, smoke() and
// bench.py's cpu_baseline leg (test infrastructure: the checker, never the
// product). Positions/velocities are packed xyz triples, as the reference's
// Vector<Real,3> (Core/Vector/vector_3d.h:31).
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "pbf_oracle.h"

using namespace pbf_oracle;

namespace {

template <class R> CSR to_csr(const int* off, const int* el, int n) {
  CSR c;
  c.offsets.assign(off, off + n + 1);
  c.elems.assign(el, el + off[n]);
  return c;
}

template <class R> R mass_of(const Params<R>& p) {
  return scaling_factor(p.d, p.h) * (p.rho0 * (p.d * p.d * p.d));
}

template <class R>
long long nq(const Params<R>* p, const R* x, int n, int* offsets, int* elems, long long cap) {
  Grid<R> g = grid_setup(p->h, p->lo, p->hi);
  CSR c;
  neighbor_query(g, reinterpret_cast<const V3<R>*>(x), n, p->h, c);
  long long total = c.offsets[n];
  if (offsets) std::memcpy(offsets, c.offsets.data(), sizeof(int) * ((size_t)n + 1));
  if (elems && cap >= total) std::memcpy(elems, c.elems.data(), sizeof(int) * (size_t)total);
  return total;
}

template <class R>
long long run_steps(const Params<R>* p, R* x, R* v, R* rho, int n, int nsteps) {
  CSR c;
  SDF<R> sdf;
  if (p->flags & 2) sdf = box_sdf(p->box_lo, p->box_hi, p->box_spacing, true);
  for (int s = 0; s < nsteps; ++s)
    step(*p, reinterpret_cast<V3<R>*>(x), reinterpret_cast<V3<R>*>(v), rho, n, &c, (p->flags & 2) ? &sdf : nullptr);
  return nsteps > 0 ? (long long)c.offsets[n] : 0;
}

}  // namespace

extern "C" {

int orc_abi_version(void) { return 1; }

int orc_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

void orc_grid_f(const Params<float>* p, float* ds, float* lo3, int* dims3) {
  Grid<float> g = grid_setup(p->h, p->lo, p->hi);
  *ds = g.ds;
  for (int c = 0; c < 3; ++c) lo3[c] = g.lo[c];
  dims3[0] = g.nx;
  dims3[1] = g.ny;
  dims3[2] = g.nz;
}

void orc_cell3_f(const Params<float>* p, const float* x, int n, int* c3, long long* cid) {
  Grid<float> g = grid_setup(p->h, p->lo, p->hi);
  const V3<float>* xv = reinterpret_cast<const V3<float>*>(x);
  for (int i = 0; i < n; ++i) {
    cell3(g, xv[i], c3 + 3 * (size_t)i);
    if (cid) cid[i] = cell_id(g, c3[3 * i], c3[3 * i + 1], c3[3 * i + 2]);
  }
}

float orc_spiky_w_f(float r, float h) { return spiky_w(r, h); }
float orc_spiky_grad_f(float r, float h) { return spiky_grad(r, h); }
double orc_spiky_w_d(double r, double h) { return spiky_w(r, h); }
double orc_spiky_grad_d(double r, double h) { return spiky_grad(r, h); }
float orc_scaling_factor_f(float d, float H) { return scaling_factor(d, H); }
double orc_scaling_factor_d(double d, double H) { return scaling_factor(d, H); }
float orc_mass_f(const Params<float>* p) { return mass_of(*p); }

long long orc_neighbor_query_f(const Params<float>* p, const float* x, int n, int* offsets, int* elems,
                               long long cap) {
  return nq(p, x, n, offsets, elems, cap);
}
long long orc_neighbor_query_d(const Params<double>* p, const double* x, int n, int* offsets,
                               int* elems, long long cap) {
  return nq(p, x, n, offsets, elems, cap);
}

void orc_density_f(const Params<float>* p, const float* x, int n, const int* off, const int* el,
                   float* rho) {
  CSR c = to_csr<float>(off, el, n);
  density(reinterpret_cast<const V3<float>*>(x), n, c, p->h, mass_of(*p), rho);
}

void orc_lambdas_f(const Params<float>* p, const float* rho, const float* x, int n, const int* off,
                   const int* el, float* lam) {
  CSR c = to_csr<float>(off, el, n);
  lambdas(*p, rho, reinterpret_cast<const V3<float>*>(x), n, c, lam);
}

// mode 1 = gather (2*sum), 2 = literal atomic-order emulation
void orc_displacement_f(const Params<float>* p, const float* lam, const float* x, int n, const int* off,
                        const int* el, float* dpos, int mode) {
  CSR c = to_csr<float>(off, el, n);
  if (mode == 2)
    displacement_atomic_order(*p, lam, reinterpret_cast<const V3<float>*>(x), n, c,
                              reinterpret_cast<V3<float>*>(dpos));
  else
    displacement_gather(*p, lam, reinterpret_cast<const V3<float>*>(x), n, c,
                        reinterpret_cast<V3<float>*>(dpos));
}

void orc_pbd_constrain_f(const Params<float>* p, float* x, float* v, int n, const int* off,
                         const int* el, float* rho) {
  CSR c = to_csr<float>(off, el, n);
  pbd_constrain(*p, reinterpret_cast<V3<float>*>(x), reinterpret_cast<V3<float>*>(v), n, c, rho);
}
void orc_pbd_constrain_d(const Params<double>* p, double* x, double* v, int n, const int* off,
                         const int* el, double* rho) {
  CSR c = to_csr<double>(off, el, n);
  pbd_constrain(*p, reinterpret_cast<V3<double>*>(x), reinterpret_cast<V3<double>*>(v), n, c, rho);
}

void orc_integrate_f(const Params<float>* p, float* x, float* v, const float* f, int n) {
  integrate(*p, reinterpret_cast<V3<float>*>(x), reinterpret_cast<V3<float>*>(v),
            reinterpret_cast<const V3<float>*>(f), n);
}
void orc_integrate_d(const Params<double>* p, double* x, double* v, const double* f, int n) {
  integrate(*p, reinterpret_cast<V3<double>*>(x), reinterpret_cast<V3<double>*>(v),
            reinterpret_cast<const V3<double>*>(f), n);
}

// nsteps of PositionBasedFluidModel::step; rho receives the last step's
// density; returns the last step's neighbor-pair count (sum of list sizes).
long long orc_step_f(const Params<float>* p, float* x, float* v, float* rho, int n, int nsteps) {
  return run_steps(p, x, v, rho, n, nsteps);
}
long long orc_step_d(const Params<double>* p, double* x, double* v, double* rho, int n, int nsteps) {
  return run_steps(p, x, v, rho, n, nsteps);
}

// Adjoint of nsteps steps from (x0, v0): xbar, vbar hold the adjoint of the
// final state on entry and of the initial state on return.
void orc_rollout_adjoint_d(const Params<double>* p, const double* x0, const double* v0, int n, int nsteps,
                           double* xbar, double* vbar) {
  rollout_adjoint(*p, reinterpret_cast<const V3<double>*>(x0), reinterpret_cast<const V3<double>*>(v0), n, nsteps,
                  reinterpret_cast<V3<double>*>(xbar), reinterpret_cast<V3<double>*>(vbar));
}
void orc_rollout_adjoint_f(const Params<float>* p, const float* x0, const float* v0, int n, int nsteps,
                           float* xbar, float* vbar) {
  rollout_adjoint(*p, reinterpret_cast<const V3<float>*>(x0), reinterpret_cast<const V3<float>*>(v0), n, nsteps,
                  reinterpret_cast<V3<float>*>(xbar), reinterpret_cast<V3<float>*>(vbar));
}

// ImplicitViscosity::constrain on a caller list.
void orc_viscosity_f(const Params<float>* p, const float* x, float* v, int n, const int* off, const int* el) {
  CSR c = to_csr<float>(off, el, n);
  viscosity(*p, reinterpret_cast<const V3<float>*>(x), reinterpret_cast<V3<float>*>(v), n, c);
}
float orc_vis_weight_f(float r, float h) { return vis_weight(r, h); }

// Box SDF of setCube(lo, hi, spacing, inverted): dims (3 ints), left, h (3
// floats each); samples copied when `samples` is non-null and holds cap.
long long orc_box_sdf_f(const float* lo, const float* hi, float spacing, int inverted, int* dims, float* left,
                        float* h, float* samples, long long cap) {
  SDF<float> f = box_sdf(lo, hi, spacing, inverted != 0);
  dims[0] = f.nx;
  dims[1] = f.ny;
  dims[2] = f.nz;
  for (int a = 0; a < 3; ++a) {
    left[a] = f.left[a];
    h[a] = f.hs[a];
  }
  const long long total = (long long)f.d.size();
  if (samples && cap >= total) std::memcpy(samples, f.d.data(), sizeof(float) * (size_t)total);
  return total;
}

// Distance and normal at points of the box field (inverted: fluid inside).
void orc_sdf_distance_f(const float* lo, const float* hi, float spacing, const float* x, int n, float* d,
                        float* normal) {
  SDF<float> f = box_sdf(lo, hi, spacing, true);
  for (int i = 0; i < n; ++i) {
    V3<float> nrm;
    sdf_distance(f, reinterpret_cast<const V3<float>*>(x)[i], d[i], nrm);
    normal[3 * i] = nrm.x;
    normal[3 * i + 1] = nrm.y;
    normal[3 * i + 2] = nrm.z;
  }
}

void orc_box_constrain_f(const Params<float>* p, float* x, float* v, int n) {
  SDF<float> f = box_sdf(p->box_lo, p->box_hi, p->box_spacing, true);
  box_constrain(*p, f, reinterpret_cast<V3<float>*>(x), reinterpret_cast<V3<float>*>(v), n);
}

float orc_box_jitter(int id) { return box_jitter(id); }

void orc_pbd_constrain_minv_f(const Params<float>* p, float* x, float* v, int n, const int* off, const int* el,
                              const float* minv, float* rho) {
  CSR c = to_csr<float>(off, el, n);
  pbd_constrain(*p, reinterpret_cast<V3<float>*>(x), reinterpret_cast<V3<float>*>(v), n, c, rho, minv);
}
void orc_pbd_constrain_minv_d(const Params<double>* p, double* x, double* v, int n, const int* off, const int* el,
                              const double* minv, double* rho) {
  CSR c = to_csr<double>(off, el, n);
  pbd_constrain(*p, reinterpret_cast<V3<double>*>(x), reinterpret_cast<V3<double>*>(v), n, c, rho, minv);
}

void orc_neighbor_query_fixed_f(const Params<float>* p, const float* x, int n, int limit, int* counts, int* elems) {
  Grid<float> g = grid_setup(p->h, p->lo, p->hi);
  neighbor_query_fixed(g, reinterpret_cast<const V3<float>*>(x), n, p->h, limit, counts, elems);
}

}  // extern "C"
