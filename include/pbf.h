/* pbf.h - C ABI of the MI355X position-based-fluid step (libpbf.so).
 *
 * Replaces, for PhysIKA's PBD particle fluid, the module methods that
 * PositionBasedFluidModel::step calls
 * (Dynamics/ParticleSystem/PositionBasedFluidModel.cpp:83-101):
 *   NeighborQuery::compute            (Framework/Topology/NeighborQuery.cu:152-199)
 *   ParticleIntegrator::begin/integrate (Dynamics/ParticleSystem/ParticleIntegrator.cu:20-34,129-138)
 *   DensityPBD::constrain             (Dynamics/ParticleSystem/DensityPBD.cu:271-302)
 *   SummationDensity::compute         (Dynamics/ParticleSystem/SummationDensity.cu:59-110)
 * with field semantics kept: positions and velocities are packed xyz float
 * triples (the reference's Vector<float,3>, Core/Vector/vector_3d.h:31), the
 * neighborhood is the reference's NeighborList layout
 * (Framework/Topology/NeighborList.h:27-66): `index`, n exclusive row starts
 * (NeighborList::m_index, n entries), and `elements`, nelems ints
 * (m_elements); row i ends at index[i+1], the last row at nelems
 * (NeighborList::getNeighborSize). The density is one float per particle.
 *
 * Conventions
 *  - Every function returns an int status: PBF_OK (0) or a negative
 *    pbf_status; pbf_last_error(ctx) gives the text. No C++ exception crosses
 *    the ABI.
 *  - Pointers named d_* are device pointers (hipMalloc'd or any device memory
 *    of the current HIP device); `stream` is a hipStream_t (NULL = the null
 *    stream). Calls are stream-ordered and asynchronous unless documented as
 *    blocking.
 *  - The caller owns x, v, rho, forces and CSR buffers. The context owns all
 *    scratch, sized by pbf_reserve() or on first use; a steady-state call with
 *    a particle count within the reservation allocates nothing.
 *  - A context is used by one host thread at a time; contexts are independent.
 */
#ifndef PBF_H_
#define PBF_H_

#ifdef __cplusplus
extern "C" {
#endif

#define PBF_ABI_VERSION 9

typedef enum pbf_status {
  PBF_OK = 0,
  PBF_ERR_INVALID = -1,     /* bad argument: null pointer, negative size, bad params */
  PBF_ERR_HIP = -2,         /* a HIP runtime call failed */
  PBF_ERR_CAPACITY = -3,    /* caller buffer too small (the needed size is reported) */
  PBF_ERR_ALLOC = -4,       /* device allocation failed */
  PBF_ERR_UNSUPPORTED = -5  /* size or configuration outside what the library handles */
} pbf_status;

/* Parameters of the step. pbf_params_default() gives the benchmark workload's
 * values (see there); the reference's own defaults differ in h and K. */
typedef struct pbf_params {
  float smoothing_length;  /* h: NeighborQuery.inRadius (search radius and grid
                              cell, NeighborQuery.cu:142) and DensityPBD /
                              SummationDensity varSmoothingLength */
  float sampling_distance; /* d: DensityPBD.varSamplingDistance; particle mass
                              rho0*d^3 and the lattice normalisation
                              (SummationDensity.cu:113-138) */
  float rest_density;      /* DensityPBD.varRestDensity (mass only) */
  float dt;                /* Node::m_dt (Framework/Framework/Node.cpp:15) */
  float gravity[3];        /* SceneGraph::getGravity (SceneGraph.h:80) */
  int iterations;          /* DensityPBD.varIterationNumber */
  float lambda_rest;       /* 1000.0f literal of K_ComputeLambdas (DensityPBD.cu:87) */
  float lambda_eps;        /* 0.1f literal of K_ComputeLambdas (DensityPBD.cu:87) */
  float corr_scale;        /* 10.0f literal of K_ComputeDisplacement (DensityPBD.cu:158) */
  float pair_eps;          /* EPSILON 1e-6 of the lambda/dx pair skip (cuda_utilities.h:30) */
  float scene_lo[3];       /* SceneGraph lower bound (SceneGraph.h:77) */
  float scene_hi[3];       /* SceneGraph upper bound (SceneGraph.h:78) */
  int mask_words;          /* neighbor bit-mask scratch reservation, 32-bit words
                              per particle; 0 (default) = 32 to start with,
                              grown automatically between calls (pbf_step) */
  int list_slots;          /* neighbor-list reservation, 16-bit slots per particle
                              (each 64-particle wavefront owns 64 x this; an
                              entry takes 1 slot, or 2 when the wavefront's
                              candidates span more than 65536 indices);
                              0 (default) = 96 to start with, grown automatically.
                              A value > 0 fixes the reservation: waves that do
                              not fit fall back to bit masks or re-testing,
                              with identical results */
  float viscosity;         /* ImplicitViscosity m_viscosity; the model sets 1
                              (PositionBasedFluidModel.cpp:70) */
  int viscosity_iterations;/* ImplicitViscosity m_maxInteration (5,
                              ImplicitViscosity.cu:71-76) */
  float box_lo[3];         /* box of StaticBoundary::loadCube (fluid inside,
                              StaticBoundary.cpp:72-77) */
  float box_hi[3];
  float box_spacing;       /* loadCube distance: sampling of the box field
                              (BoundaryConstraint.cu:109-117) */
  float normal_friction;   /* BoundaryConstraint m_normal_friction (0.95) */
  float tangent_friction;  /* BoundaryConstraint m_tangent_friction (0) */
  unsigned flags;          /* PBF_FLAG_* (0: the plain PBD step) */
  int stage_min_particles; /* searches over at least this many particles let
                              each wavefront's PBD passes read the neighbors in
                              its own z-plane rows from a copy in LDS (staged
                              lists); 0 (default) = 262144, < 0 = never. The
                              results are identical either way */
} pbf_params;

/* pbf_params.flags */
#define PBF_FLAG_VISCOSITY 1u /* ImplicitViscosity after the PBD solve, as the
                                 reference model's step does */
#define PBF_FLAG_BOX 2u       /* box boundary on x, v before every step (the
                                 StaticBoundary root node advances before the
                                 fluid, SceneGraph.cpp:101-114) */
#define PBF_FLAG_EAGER 4u     /* launch every kernel of a multi-step call
                                 directly. By default the steps between the
                                 first and the last of a plain pbf_step /
                                 pbf_batch_step call (no viscosity, no box)
                                 replay one captured hipGraph of the step's
                                 kernels; the results are identical */

/* Diagnostics of the last neighbor search of a context. */
typedef struct pbf_stats {
  long long neighbor_pairs;      /* sum of neighbor-list sizes (self included) */
  long long out_of_grid;         /* particles outside the grid: not hashed, still querying */
  long long nonfinite;           /* particles with a NaN/Inf coordinate (counted in out_of_grid) */
  long long mask_words_needed;   /* neighbor bit-mask words the last search wanted */
  long long mask_words_reserved; /* neighbor bit-mask words reserved */
  long long mask_fallback_waves; /* wavefronts that re-test candidates in every pass
                                    (bit-mask reservation full) */
  long long list_slots_needed;   /* neighbor-list slots the last search wanted: its
                                    wavefronts x 64 x the slots per particle of the
                                    fullest wavefront */
  long long list_slots_reserved; /* neighbor-list slots reserved */
  long long cells_used;          /* compact cells of the last search: the scenes'
                                    occupied-cell windows (sort cells: half a
                                    grid cell along x) plus one out-of-grid
                                    bucket per scene */
  long long candidates;          /* candidate pairs the last search tested (every
                                    particle of the 27 stencil cells) */
  long long staged_waves;        /* wavefronts of the last search whose PBD passes read
                                    the neighbors of their own z-plane rows from a
                                    copy in LDS */
  int grid_dims[3];              /* cells per axis (GridHash::setSpace) */
  float grid_lo[3];              /* grid origin = scene_lo - 2 cells */
  float cell_size;
  float density_factor;          /* m_factor of SummationDensity */
  float particle_mass;           /* rest_density * d^3 */
  float neighbor_threshold;      /* squared-distance bound T: sqrtf(s) < h  <=>  s < T */
  long long graph_captures;      /* hipGraph captures of a multi-step call's middle step (context lifetime) */
  long long graph_replays;       /* middle steps replayed from that graph (context lifetime) */
  int pass_lanes;                /* lanes per particle of the last step's PBD passes (pbf_tuning) */
} pbf_stats;

typedef struct pbf_ctx pbf_ctx;

int pbf_abi_version(void);
const char* pbf_status_string(int status);

/* Workload defaults (the fluid benchmark: lattice spacing 0.005 with
 * h = 2.5 spacings and 5 PBD iterations): h 0.0125, d 0.005, rho0 1000,
 * dt 0.001, g (0,-9.8,0), 5 iterations, lambda constants 1000 / 0.1,
 * correction 10, pair eps 1e-6, scene [0,1]^3; viscosity 1 with 5
 * iterations, box [0,1]^3 sampled at 0.005 with normal friction 0.95 and
 * tangential friction 0; flags 0 (the viscosity and the box are off until
 * their flag is set).
 * The reference's own defaults differ in two fields: the model's smoothing
 * length is 0.006 (PositionBasedFluidModel.cpp:25) and DensityPBD runs 3
 * iterations (DensityPBD.cu:242); set smoothing_length and iterations to
 * those values to run the reference's default model. Every other field
 * equals the reference's value or literal. */
void pbf_params_default(pbf_params* p);

/* Creates a context on the current HIP device. Replaces module construction
 * + Module::initialize (NeighborQuery::initializeImpl builds the grid once,
 * NeighborQuery.cu:107-149; SummationDensity computes m_factor,
 * SummationDensity.cu:113-138). */
int pbf_create(const pbf_params* p, pbf_ctx** out);
int pbf_destroy(pbf_ctx* ctx);

/* Replaces the VarField setters and their callbacks (FieldVar.h:110-145):
 * recomputes m_factor and the mass; rebuilds the grid when h or the scene
 * bounds change. Blocking (may reallocate the grid arrays). */
int pbf_set_params(pbf_ctx* ctx, const pbf_params* p);
int pbf_get_params(const pbf_ctx* ctx, pbf_params* out);

/* Sizes the context's scratch for up to n particles. Blocking. */
int pbf_reserve(pbf_ctx* ctx, int n);
/* Same for a batch of up to nscenes scenes holding n particles in total. */
int pbf_reserve_batch(pbf_ctx* ctx, int n, int nscenes);

/* NeighborQuery::compute (NeighborQuery.cu:152-199): neighbor lists of n
 * particles at d_x (original order, original ids, self included, strict
 * |x_i - x_j| < h) in the NeighborList layout: d_index receives the n row
 * starts (m_index), *total the element count (m_elements.size()). Rows are
 * written when the total fits in `cap` elements (d_elems may be NULL to size
 * only). Returns PBF_ERR_CAPACITY when cap < total. Blocking: the total is
 * read back, as the reference does (NeighborQuery.cu:463). */
int pbf_neighbor_query(pbf_ctx* ctx, const float* d_x, int n, int* d_index, int* d_elems,
                       long long cap, long long* total, void* stream);

/* NeighborQuery::compute with a neighbor limit (NeighborList::isLimited;
 * queryNeighborFixed, NeighborQuery.cu:514-581): row i keeps the `limit`
 * nearest of its neighbors (1 <= limit <= 32, else PBF_ERR_UNSUPPORTED):
 * d_counts[i] = kept, d_elems[limit*i + k] their ids sorted by distance
 * (ties: smaller id first; the reference's tie order follows its atomic
 * visit order). d_elems holds n*limit ints. */
int pbf_neighbor_query_fixed(pbf_ctx* ctx, const float* d_x, int n, int limit, int* d_counts, int* d_elems,
                             void* stream);

/* ParticleIntegrator::begin + integrate (ParticleIntegrator.cu:20-34,129-138):
 * v += dt*(f + g), x += dt*v. d_f may be NULL (the zeroed force of begin()). */
int pbf_integrate(pbf_ctx* ctx, float* d_x, float* d_v, const float* d_f, int n, void* stream);

/* SummationDensity::compute (SummationDensity.cu:59-110) over a caller
 * NeighborList (d_index: n row starts, d_elems: nelems ids). */
int pbf_density(pbf_ctx* ctx, const float* d_x, const int* d_index, const int* d_elems, long long nelems,
                float* d_rho, int n, void* stream);

/* DensityPBD::constrain (DensityPBD.cu:271-302) over a caller NeighborList
 * (d_index: n row starts, d_elems: nelems ids; any list, symmetric or not):
 * `iterations` Jacobi rounds of density, lambda and pairwise displacement
 * (float atomics, as the reference), then v += (x - x_old)/dt. x and v are
 * updated in place; d_rho (may be NULL) receives the last iteration's
 * density. */
int pbf_pbd_constrain(pbf_ctx* ctx, float* d_x, float* d_v, const int* d_index,
                      const int* d_elems, long long nelems, float* d_rho, int n, void* stream);

/* DensityPBD::constrain with DensityPBD::m_massInv set (the massInv variants
 * of K_ComputeLambdas and K_ComputeDisplacement, DensityPBD.cu:93-131,182-224):
 * d_mass_inv holds one inverse mass per particle. Otherwise as
 * pbf_pbd_constrain. */
int pbf_pbd_constrain_mass(pbf_ctx* ctx, float* d_x, float* d_v, const int* d_index, const int* d_elems,
                           long long nelems, const float* d_mass_inv, float* d_rho, int n, void* stream);

/* pbf_pbd_constrain without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, semantics and results (to float-atomic summation order) as
 * pbf_pbd_constrain, but its scratch (iterate positions and displacement
 * sums) is a fixed array in the library, one copy per device, instead of the
 * context's reservation: at most PBF_NOALLOC_MAX_PARTICLES particles
 * (PBF_ERR_UNSUPPORTED beyond). The array holds 28 MiB on every device the
 * library runs on until the process ends. Every call on one device shares
 * that device's array: calls that overlap on one device (several streams,
 * threads or contexts) race and must be serialized by the caller. The
 * context only supplies the parameters. */
#define PBF_NOALLOC_MAX_PARTICLES (1 << 20)
int pbf_pbd_constrain_noalloc(pbf_ctx* ctx, float* d_x, float* d_v, const int* d_index, const int* d_elems,
                              long long nelems, float* d_rho, int n, void* stream);

/* nsteps of PositionBasedFluidModel::step without viscosity
 * (PositionBasedFluidModel.cpp:83-101): begin -> NeighborQuery on x^n ->
 * integrate -> DensityPBD::constrain, fused. x and v (original order) are
 * updated in place, d_rho (may be NULL) receives the density of the last
 * iteration of the last step. Internally the particles are sorted by cell, the
 * neighbor sets are kept as compact lists over cell runs and the displacement
 * is gathered without atomics (deterministic: repeated runs give identical
 * bits). The neighbor-storage reservations grow between calls when a previous
 * call needed more (read without blocking once that call has finished). */
int pbf_step(pbf_ctx* ctx, float* d_x, float* d_v, float* d_rho, int n, int nsteps, void* stream);

/* pbf_step on a batch of independent scenes that share the parameters (each
 * scene in its own copy of the scene box; the reference runs one
 * PositionBasedFluidModel per scene). Scene s holds particles
 * d_scene_offsets[s] .. d_scene_offsets[s+1]-1 of x, v and rho
 * (d_scene_offsets: nscenes+1 non-decreasing device ints, first 0, last n).
 * Particles of different scenes never interact; every scene's result is
 * bit-identical to a pbf_step of that scene alone. All scenes advance in one
 * launch per pass. No counterpart in the reference, which has no batching. */
int pbf_batch_step(pbf_ctx* ctx, float* d_x, float* d_v, float* d_rho, const int* d_scene_offsets,
                   int nscenes, int n, int nsteps, void* stream);

/* Reverse mode (not part of the reference, which has no derivatives).
 *
 * pbf_rollout_forward runs nsteps of pbf_batch_step (d_scene_offsets may be
 * NULL for one scene, nscenes = 1) on x, v in place and keeps the state at
 * every checkpoint_every-th step in the context. pbf_rollout_backward then
 * takes the adjoint of the final state (d_xbar, d_vbar: dL/dx_N, dL/dv_N,
 * original order) and overwrites it with the adjoint of the initial state
 * (dL/dx_0, dL/dv_0). Each segment between checkpoints is recomputed, every
 * step of it re-run with its PBD iterates kept and its neighbor sets rebuilt,
 * and differentiated by hand-written adjoint kernels with the neighbor sets
 * held fixed (the step is C1 across a set change: W and the kernel gradient
 * vanish with their first derivative at r = h). The rollout stays valid for
 * repeated backward calls until the next pbf_rollout_forward or a parameter
 * change (PBF_ERR_INVALID). Returns PBF_ERR_UNSUPPORTED when a step had
 * particles outside the grid. pbf_rollout_backward is blocking. */
int pbf_rollout_forward(pbf_ctx* ctx, float* d_x, float* d_v, float* d_rho, const int* d_scene_offsets,
                        int nscenes, int n, int nsteps, int checkpoint_every, void* stream);
int pbf_rollout_backward(pbf_ctx* ctx, float* d_xbar, float* d_vbar, void* stream);

/* Reverse mode in double precision: pbf_rollout_forward / _backward with the
 * state (x, v, rho and their adjoints: packed xyz doubles, original order) and
 * every quantity of the integrator and the PBD iterations in double, the
 * constants computed from the parameters promoted to double. The neighbor
 * sets come from the single-precision search on the positions rounded to
 * float (they differ from a double-precision search only for pairs within a
 * float ulp of h, where the kernel and its derivative vanish). For gradients
 * that must match FP64 finite differences over horizons where the fp32
 * trajectory departs from the FP64 one (lambda = min(0, .) switches on
 * rounding). pbf_rollout_forward_f64 is blocking and returns
 * PBF_ERR_UNSUPPORTED when a step had particles outside the grid; otherwise
 * as pbf_rollout_forward / pbf_rollout_backward. */
int pbf_rollout_forward_f64(pbf_ctx* ctx, double* d_x, double* d_v, double* d_rho, const int* d_scene_offsets,
                            int nscenes, int n, int nsteps, int checkpoint_every, void* stream);
int pbf_rollout_backward_f64(pbf_ctx* ctx, double* d_xbar, double* d_vbar, void* stream);

/* Export of the neighbor lists built by the last pbf_step (its last step, on
 * the positions before that step's integration), original ids, NeighborList
 * layout; same contract as pbf_neighbor_query (the outNeighborhood field of
 * the fused step). Blocking. */
int pbf_step_neighbors(pbf_ctx* ctx, int* d_index, int* d_elems, long long cap, long long* total,
                       void* stream);

/* Diagnostics of the last neighbor search. Blocking (synchronises the stream
 * of the last call). */
int pbf_get_stats(pbf_ctx* ctx, pbf_stats* out);

/* ImplicitViscosity::constrain (ImplicitViscosity.cu:99-128) over a caller
 * NeighborList: viscosity_iterations Jacobi rounds on v (in place) with
 * weights from the positions d_x. */
int pbf_viscosity(pbf_ctx* ctx, const float* d_x, float* d_v, const int* d_index, const int* d_elems,
                  long long nelems, int n, void* stream);

/* BoundaryConstraint::constrain (BoundaryConstraint.cu:30-84) with the box of
 * box_lo/box_hi/box_spacing: particles at or outside the walls are projected
 * back inside (plus a jitter of at most 1e-4 per particle id, drawn by this
 * library's own counter-based generator rather than curand), their normal
 * velocity reflected and damped by the normal friction. x, v in place. */
int pbf_box_constrain(pbf_ctx* ctx, float* d_x, float* d_v, int n, void* stream);

/* Per-pass GPU timing of pbf_step (the reference's GTimer,
 * Core/Utility/GTimer.cpp:21-27, has no fluid hook). Not part of the reference
 * interface. A one-kernel pass (neighbors, density + lambda, displacement) is
 * timed by start/stop events attached to its dispatch (hipExtLaunchKernel: the
 * kernel's own duration); the sort, several kernels, by events recorded on the
 * call's stream before and after it. */
typedef enum pbf_pass {
  PBF_PASS_SORT = 0,            /* clear + bin + scan + scatter + reorder/integrate */
  PBF_PASS_NEIGHBORS = 1,       /* neighbor masks (one kernel) */
  PBF_PASS_DENSITY_LAMBDA = 2,  /* density + lambda (one kernel per iteration) */
  PBF_PASS_DISPLACEMENT = 3,    /* displacement + position update (one kernel per iteration) */
  PBF_PASS_COUNT = 4
} pbf_pass;

typedef struct pbf_profile {
  double ms[PBF_PASS_COUNT];          /* summed event time per pass */
  long long launches[PBF_PASS_COUNT]; /* timed segments per pass */
} pbf_profile;

/* enable != 0 starts (and clears) timing; 0 stops it. */
int pbf_set_profiling(pbf_ctx* ctx, int enable);
/* Accumulated timings since profiling was enabled. Blocking. */
int pbf_get_profile(pbf_ctx* ctx, pbf_profile* out);

/* Launch settings of a context. Not part of the reference (its kernels run in
 * blocks of BLOCK_SIZE = 64 threads, Core/Utility/cuda_utilities.h:34). None
 * of them changes a result: every setting gives bit-identical outputs. */
typedef struct pbf_tuning {
  int pass_block;          /* threads per workgroup of the density/lambda and
                              displacement passes: 64, 128, 256 (default) or 512 */
  int stage_min_particles; /* as pbf_params.stage_min_particles: searches over at
                              least this many particles stage their neighbor
                              lists in LDS; < 0 never */
  int persistent;          /* 1: the two passes launch as many workgroups as stay
                              resident on the device, each wave walking
                              64-particle tiles of its XCD's contiguous share in
                              a fixed order; 0 (default): one wave per tile */
  int pass_lanes;          /* lanes per particle in the two passes: 1; 2 (a lane
                              pair splits each neighbor list by the parity of
                              the entries' order and adds the two sums, twice
                              the waves for small searches; not with
                              persistent); 0 (default): 2 without persistent
                              launches when the search keeps unstaged lists or
                              covers at most PBF_TWO_LANE_MAX_PARTICLES
                              particles, else 1 */
} pbf_tuning;

#define PBF_TWO_LANE_MAX_PARTICLES 196608

int pbf_set_tuning(pbf_ctx* ctx, const pbf_tuning* t);
int pbf_get_tuning(const pbf_ctx* ctx, pbf_tuning* out);

/* Autotuner. Times every setting of pbf_tuning (4 workgroup sizes x staged
 * lists on / off x persistent or not, pass_lanes automatic: the applied
 * setting has pass_lanes 0) on the caller's workload: a copy of (d_x, d_v), `steps` plain
 * steps after a warm-up call, the better of two runs, on `stream`; keeps the
 * fastest (a setting replaces the one before it only when more than 1%
 * faster, the default first) and applies it to the context. The workload is a
 * scene (d_scene_offsets NULL, nscenes 1) or a batch (as pbf_batch_step); the
 * caller's x and v are not changed. With cache_path (may be NULL) the choice
 * is kept in a JSON file under a key of the device architecture, the particle
 * count (half octaves), the scene count (octaves), the mean neighbor count
 * (octaves, from one step of the copy), h/d, the iteration count and the
 * flags; a later call with an equal key applies the stored setting without
 * timing. Blocking; allocates two copies of the state while it runs. */
typedef struct pbf_tune_report {
  pbf_tuning best;
  double best_ms;     /* per step of the chosen setting (timed, or as stored in the cache) */
  double default_ms;  /* per step of the default setting (0 when read from the cache) */
  int candidates;     /* settings timed (0 when read from the cache) */
  int from_cache;     /* 1: the setting came from cache_path */
  char key[128];      /* the cache key of this workload */
} pbf_tune_report;

int pbf_autotune(pbf_ctx* ctx, const float* d_x, const float* d_v, const int* d_scene_offsets, int nscenes, int n,
                 int steps, const char* cache_path, pbf_tune_report* out, void* stream);

const char* pbf_last_error(const pbf_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* PBF_H_ */
