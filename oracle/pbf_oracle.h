// CPU oracle for the PhysIKA position-based-fluid step.
//
// TEST INFRASTRUCTURE ONLY. This is the checker: a plain C++ restatement of the
// reference algorithm that the GPU product is compared against. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product library never links, calls or falls back to it.
//
// Every function follows a reference file:line (paths relative to the
// reference's Source/ tree unless stated) and keeps its fp evaluation order.
// The file is compiled with -ffp-contract=off so that no multiply-add is fused.
// Templated on the scalar type: float mirrors the reference (PRECISION_FLOAT,
// Core/Platform.h:48); double serves finite-difference and adjoint checks.
//
// Pinning (what each function is checked against, tests/test_oracle_*.py):
//   grid/cell index   closed form on hand-picked points (cell faces, padding,
//                     out-of-grid)
//   neighbor query    brute-force O(N^2) sets, symmetry, self-inclusion,
//                     cell-size independence, h +- 1 ulp pairs
//   spiky W / dW      FP64 closed form; dW = d/dr W by central differences
//   scaling factor    independent 343-term sum; perfect lattice interior
//                     density == rest density
//   density           lattice closed form; isolated particle closed form
//   lambda            rho <= 1000 -> 0; two-particle closed form
//   displacement      momentum conservation; literal atomic-order emulation of
//                     the reference's pairwise scatter (factor 2); 2-particle
//                     closed form
//   integrator        free fall closed form (bit-exact)
//   massInv variant   unit inverse masses reproduce the plain path; mass-
//                     weighted momentum conservation; two-particle closed form
//   fixed neighbors   the replace-farthest rows hold the `limit` nearest of the
//                     dynamic set (sorted selection on tie-free inputs)
//   step              committed golden fixtures (regression pins only)
//   step_adjoint,     FP64 central finite differences of step()/rollouts on
//   rollout_adjoint   small lattices (compressed, so that lambda is active),
//                     and the closed form of the free-fall adjoint
#pragma once
#include <cstdint>
#include <vector>

namespace pbf_oracle {

template <class R> struct V3 { R x, y, z; };

// Grid geometry of GridHash::setSpace (Framework/Topology/GridHash.cu:47-75).
template <class R> struct Grid {
  R ds;
  R lo[3];
  int nx, ny, nz;
  long long num;  // nx*ny*nz
};

template <class R> struct Params {
  R h;             // smoothing length = search radius = grid cell
  R d;             // sampling distance (sets the mass and the lattice factor)
  R rho0;          // rest density (mass only)
  R dt;
  R g[3];
  int iters;       // DensityPBD iteration number
  R lam_rest;      // 1000.0f literal, DensityPBD.cu:87
  R lam_eps;       // 0.1f literal, DensityPBD.cu:87
  R corr;          // 10.0f literal, DensityPBD.cu:158
  R pair_eps;      // EPSILON, Core/Utility/cuda_utilities.h:30
  R lo[3], hi[3];  // scene bounds, SceneGraph.h:77-78
  // Δx accumulation: 0 = gather when the list is symmetric, else the
  // atomic-order emulation; 1 = gather; 2 = atomic-order emulation.
  int dx_mode;
  // ImplicitViscosity (ImplicitViscosity.cu:71-81; the model sets
  // viscosity 1, PositionBasedFluidModel.cpp:70) and the box boundary of
  // StaticBoundary::loadCube (StaticBoundary.cpp:72-77; normal friction 0.95,
  // tangential 0, BoundaryConstraint.h:32-33). flags: 1 = viscosity after the
  // PBD solve, 2 = box boundary before every step.
  R mu;
  int vis_iters;
  R box_lo[3], box_hi[3];
  R box_spacing;
  R nfric, tfric;
  int flags;
};

// Sampled signed distance field of DistanceField3D (Framework/Topology/
// DistanceField3D.h:17-110): samples at left + (i,j,k)*h, i < n.
template <class R> struct SDF {
  R left[3], hs[3], inv_h[3];
  int nx, ny, nz;
  bool inverted;
  std::vector<R> d;  // x fastest
};

struct CSR {
  std::vector<int> offsets;  // n+1 exclusive offsets (NeighborList m_index + total)
  std::vector<int> elems;    // neighbor ids
};

template <class R> Grid<R> grid_setup(R h, const R lo[3], const R hi[3]);
// Cell coordinates of GridHash::getIndex3 (GridHash.h:55-62). Values far out of
// range are clamped to [-2, n+1], which keeps the set of valid 27-stencil cells
// unchanged; a NaN coordinate yields -2 (no valid stencil cell).
template <class R> void cell3(const Grid<R>& g, const V3<R>& p, int out[3]);
// GridHash::getIndex(i,j,k) (GridHash.h:37-44): linear id or -1.
template <class R> long long cell_id(const Grid<R>& g, int i, int j, int k);

template <class R> R spiky_w(R r, R h);     // Dynamics/ParticleSystem/Kernel.h:36-45
template <class R> R spiky_grad(R r, R h);  // Dynamics/ParticleSystem/Kernel.h:47-57
template <class R> R scaling_factor(R d, R H);  // SummationDensity.cu:113-138

template <class R>
void neighbor_query(const Grid<R>& g, const V3<R>* x, int n, R h, CSR& out);
template <class R>
void density(const V3<R>* x, int n, const CSR& nb, R h, R mass, R* rho);
// minv: inverse masses (the massInv variant) or null.
template <class R>
void lambdas(const Params<R>& p, const R* rho, const V3<R>* x, int n, const CSR& nb, R* lam,
             const R* minv = nullptr);
template <class R>
void displacement_gather(const Params<R>& p, const R* lam, const V3<R>* x, int n,
                         const CSR& nb, V3<R>* dpos);
template <class R>
void displacement_atomic_order(const Params<R>& p, const R* lam, const V3<R>* x, int n,
                               const CSR& nb, V3<R>* dpos, const R* minv = nullptr);
template <class R>
void pbd_constrain(const Params<R>& p, V3<R>* x, V3<R>* v, int n, const CSR& nb, R* rho,
                   const R* minv = nullptr);
// NeighborQuery with a neighbor limit (NeighborList::isLimited, the
// fixed-capacity layout elems[limit*i + k]).
template <class R>
void neighbor_query_fixed(const Grid<R>& g, const V3<R>* x, int n, R h, int limit, int* counts, int* elems);
template <class R>
void integrate(const Params<R>& p, V3<R>* x, V3<R>* v, const V3<R>* f, int n);
// sdf: the box field when flags & 2 (built here when null).
template <class R>
void step(const Params<R>& p, V3<R>* x, V3<R>* v, R* rho, int n, CSR* nb_out, const SDF<R>* sdf = nullptr);

// VB_VisWeight (ImplicitViscosity.cu:9-18) and ImplicitViscosity::constrain
// (ImplicitViscosity.cu:99-128) over the list nb on positions x.
template <class R> R vis_weight(R r, R h);
template <class R>
void viscosity(const Params<R>& p, const V3<R>* x, V3<R>* v, int n, const CSR& nb);
// BoundaryConstraint::setCube (BoundaryConstraint.cu:109-117) with
// DistanceField3D::setSpace / loadBox (DistanceField3D.cu:194-201,283-291).
template <class R> SDF<R> box_sdf(const R lo[3], const R hi[3], R spacing, bool inverted);
// DistanceField3D::getDistance (DistanceField3D.h:112-170).
template <class R> void sdf_distance(const SDF<R>& f, const V3<R>& p, R& d, V3<R>& normal);
// The jitter of K_ConstrainSDF: 1e-4 times a uniform in (0, 1] drawn per
// particle id. The reference draws it with curand (RandNumber(pId),
// Core/Utility/CudaRand.h); the library and this oracle use a counter-based draw
// (SplitMix64 of the id, top 24 bits), so jittered positions differ from the
// reference's by at most 1e-4 m along the wall normal.
float box_jitter(int id);
// K_ConstrainSDF (BoundaryConstraint.cu:30-67); ids are the particles'
// indices in x (the reference's pId).
template <class R>
void box_constrain(const Params<R>& p, const SDF<R>& f, V3<R>* x, V3<R>* v, int n);

// Reverse-mode derivative of step() with the neighbor sets held fixed (they
// are piecewise constant in x; W and the kernel gradient vanish with their
// first derivative at r = h, so the step is C1 across a set change). x, v:
// the state before the step. xbar, vbar: in, the adjoint of the state after
// the step; out, the adjoint of x, v. Pairwise scatter over the listed
// ordered pairs, valid for symmetric and asymmetric lists alike. Not part of
// the reference (which has no derivatives).
template <class R>
void step_adjoint(const Params<R>& p, const V3<R>* x, const V3<R>* v, int n, V3<R>* xbar, V3<R>* vbar);
// nsteps of step() from (x0, v0) kept in memory, then step_adjoint in
// reverse. xbar, vbar as for step_adjoint (final state in, initial out).
template <class R>
void rollout_adjoint(const Params<R>& p, const V3<R>* x0, const V3<R>* v0, int n, int nsteps, V3<R>* xbar,
                     V3<R>* vbar);

}  // namespace pbf_oracle
