// C++ mirror of the reference's fluid modules over the C ABI of libpbf.so.
//
// Class, field and method names follow PhysIKA so that code written against
// its modules reads the same:
//   NeighborQuery            Framework/Topology/NeighborQuery.h:13-95
//                            inRadius(), inPosition(), outNeighborhood(),
//                            setNeighborSizeLimit(), compute()
//   NeighborList             Framework/Topology/NeighborList.h:9-125
//   ParticleIntegrator       Dynamics/ParticleSystem/ParticleIntegrator.h:8-54
//                            inPosition(), inVelocity(), inForceDensity(),
//                            begin(), integrate(), end()
//   DensityPBD               Dynamics/ParticleSystem/DensityPBD.h:18-88
//                            varIterationNumber(), varRestDensity(),
//                            varSamplingDistance(), varSmoothingLength(),
//                            inPosition(), inVelocity(), inNeighborIndex(),
//                            outDensity(), m_massInv, constrain()
//   ImplicitViscosity        Dynamics/ParticleSystem/ImplicitViscosity.h:10-44
//   BoundaryConstraint       Dynamics/ParticleSystem/BoundaryConstraint.h:12-40
//   PositionBasedFluidModel  Dynamics/ParticleSystem/PositionBasedFluidModel.h:25-60
//                            m_smoothingLength, m_position, m_velocity,
//                            m_forceDensity, initialize(), step(dt)
//
// Arrays are HIP device memory of packed 12-byte xyz triples (the reference's
// Vector<float,3>); connected array fields share one device array, as
// FieldArray.h:166-185 does. Every compute call goes through the C ABI
// (include/pbf.h); errors come back as physika_amd::Error exceptions on this side of
// the ABI. All modules of one model share one pbf_ctx.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pbf.h"

namespace physika_amd {

struct Vec3f {
  float x, y, z;
};
using Real = float;
using Coord = Vec3f;
static_assert(sizeof(Coord) == 12, "Coord must be a packed float triple");

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& msg) : std::runtime_error(msg), status(status) {}
  int status;
};

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(PBF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// DeviceArray<T> (Core/Array/Array.h): a device buffer; copies share it.
template <class T> class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(int n) { resize(n); }
  void resize(int n) {
    if (n == size_) return;
    T* p = nullptr;
    if (n > 0) check_hip(hipMalloc(&p, sizeof(T) * (size_t)n), "hipMalloc");
    buf_ = std::shared_ptr<T>(p, [](T* q) {
      if (q) (void)hipFree(q);
    });
    size_ = n;
  }
  int size() const { return size_; }
  T* getDataPtr() const { return buf_.get(); }
  void release() {
    buf_.reset();
    size_ = 0;
  }
  void reset() {
    if (size_) check_hip(hipMemset(buf_.get(), 0, sizeof(T) * (size_t)size_), "hipMemset");
  }
  void assign(const std::vector<T>& host) {
    resize((int)host.size());
    if (size_) check_hip(hipMemcpy(buf_.get(), host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice), "hipMemcpy");
  }
  std::vector<T> toHost() const {
    std::vector<T> h((size_t)size_);
    if (size_) check_hip(hipMemcpy(h.data(), buf_.get(), sizeof(T) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
  }

 private:
  std::shared_ptr<T> buf_;
  int size_ = 0;
};

// DeviceArrayField<T> (Framework/Framework/FieldArray.h): connect(sink) makes
// the sink read and write the source's array.
template <class T> class DeviceArrayField {
 public:
  DeviceArray<T>& getValue() { return root()->value_; }
  void setValue(const DeviceArray<T>& a) { root()->value_ = a; }
  void setElementCount(int n) { getValue().resize(n); }
  int getElementCount() { return getValue().size(); }
  bool isEmpty() { return getElementCount() == 0; }
  void connect(DeviceArrayField<T>* sink) { sink->src_ = this; }

 private:
  DeviceArrayField* root() {
    DeviceArrayField* f = this;
    while (f->src_) f = f->src_;
    return f;
  }
  DeviceArrayField* src_ = nullptr;
  DeviceArray<T> value_;
};

// NeighborList<int> (Framework/Topology/NeighborList.h:9-125): dynamic rows as
// CSR (index: the n row starts, as the reference's m_index; the last row ends
// at getElements().size()) or limited rows of getNeighborLimit() slots
// (index: row sizes).
class NeighborList {
 public:
  DeviceArray<int>& getIndex() { return index_; }
  DeviceArray<int>& getElements() { return elements_; }
  bool isLimited() const { return limit_ > 0; }
  int getNeighborLimit() const { return limit_; }
  void setNeighborLimit(int n) { limit_ = n; }
  int size() const { return index_.size(); }

 private:
  DeviceArray<int> index_, elements_;
  int limit_ = 0;
};

class NeighborField {
 public:
  NeighborList& getValue() { return root()->value_; }
  void connect(NeighborField* sink) { sink->src_ = this; }

 private:
  NeighborField* root() {
    NeighborField* f = this;
    while (f->src_) f = f->src_;
    return f;
  }
  NeighborField* src_ = nullptr;
  NeighborList value_;
};

// The pbf_ctx shared by the modules of a model; parameter writes go through
// pbf_set_params (the VarField callbacks of FieldVar.h:110-145).
class Context {
 public:
  explicit Context(const pbf_params* p = nullptr) {
    check(pbf_create(p, &ctx_), nullptr);
  }
  ~Context() {
    if (ctx_) pbf_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  pbf_ctx* get() const { return ctx_; }
  pbf_params params() const {
    pbf_params p;
    check(pbf_get_params(ctx_, &p), ctx_);
    return p;
  }
  template <class F> void update(F&& f) {
    pbf_params p = params();
    f(p);
    check(pbf_set_params(ctx_, &p), ctx_);
  }
  static void check(int st, const pbf_ctx* c) {
    if (st != PBF_OK) throw Error(st, c ? pbf_last_error(c) : pbf_status_string(st));
  }

 private:
  pbf_ctx* ctx_ = nullptr;
};

// VarField<T> (Framework/Framework/FieldVar.h): a value whose writes update the
// context's parameters.
template <class T> class VarField {
 public:
  using Setter = void (*)(pbf_params&, T);
  using Getter = T (*)(const pbf_params&);
  VarField(std::shared_ptr<Context> ctx, Getter get, Setter set) : ctx_(std::move(ctx)), get_(get), set_(set) {}
  T getValue() const { return get_(ctx_->params()); }
  void setValue(T v) {
    ctx_->update([&](pbf_params& p) { set_(p, v); });
  }
  bool isEmpty() const { return false; }

 private:
  std::shared_ptr<Context> ctx_;
  Getter get_;
  Setter set_;
};

#define PBF_VAR(ctx, T, field) \
  VarField<T>(ctx, [](const pbf_params& p) { return (T)p.field; }, [](pbf_params& p, T v) { p.field = v; })

class NeighborQuery {
 public:
  explicit NeighborQuery(std::shared_ptr<Context> ctx)
      : ctx_(ctx), m_radius(PBF_VAR(ctx, Real, smoothing_length)) {}
  VarField<Real>* inRadius() { return &m_radius; }
  DeviceArrayField<Coord>* inPosition() { return &m_position; }
  NeighborField* outNeighborhood() { return &m_neighborhood; }
  void setNeighborSizeLimit(int num) { m_maxNum = num; }

  // NeighborQuery::compute (NeighborQuery.cu:152-199)
  void compute() {
    DeviceArray<Coord>& x = m_position.getValue();
    const int n = x.size();
    NeighborList& nl = m_neighborhood.getValue();
    if (m_maxNum > 0) {
      nl.setNeighborLimit(m_maxNum);
      nl.getIndex().resize(n);
      nl.getElements().resize(n * m_maxNum);
      Context::check(pbf_neighbor_query_fixed(ctx_->get(), (const float*)x.getDataPtr(), n, m_maxNum,
                                              nl.getIndex().getDataPtr(), nl.getElements().getDataPtr(), nullptr),
                     ctx_->get());
      return;
    }
    nl.setNeighborLimit(0);
    nl.getIndex().resize(n);
    long long total = 0;
    Context::check(pbf_neighbor_query(ctx_->get(), (const float*)x.getDataPtr(), n, nl.getIndex().getDataPtr(),
                                      nullptr, 0, &total, nullptr),
                   ctx_->get());
    nl.getElements().resize((int)total);
    Context::check(pbf_neighbor_query(ctx_->get(), (const float*)x.getDataPtr(), n, nl.getIndex().getDataPtr(),
                                      nl.getElements().getDataPtr(), total, &total, nullptr),
                   ctx_->get());
  }

 private:
  std::shared_ptr<Context> ctx_;
  VarField<Real> m_radius;
  DeviceArrayField<Coord> m_position;
  NeighborField m_neighborhood;
  int m_maxNum = 0;
};

class ParticleIntegrator {
 public:
  explicit ParticleIntegrator(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) {}
  DeviceArrayField<Coord>* inPosition() { return &m_position; }
  DeviceArrayField<Coord>* inVelocity() { return &m_velocity; }
  DeviceArrayField<Coord>* inForceDensity() { return &m_force; }
  // ParticleIntegrator::begin (ParticleIntegrator.cu:20-34): the force is zeroed
  void begin() { m_force.getValue().reset(); }
  // ParticleIntegrator::integrate (ParticleIntegrator.cu:129-138)
  bool integrate() {
    const int n = m_position.getElementCount();
    if (n == 0) return true;
    Context::check(pbf_integrate(ctx_->get(), (float*)m_position.getValue().getDataPtr(),
                                 (float*)m_velocity.getValue().getDataPtr(),
                                 m_force.isEmpty() ? nullptr : (const float*)m_force.getValue().getDataPtr(), n,
                                 nullptr),
                   ctx_->get());
    return true;
  }
  void end() {}

 private:
  std::shared_ptr<Context> ctx_;
  DeviceArrayField<Coord> m_position, m_velocity, m_force;
};

class DensityPBD {
 public:
  explicit DensityPBD(std::shared_ptr<Context> ctx)
      : ctx_(ctx),
        m_iterations(PBF_VAR(ctx, int, iterations)),
        m_restDensity(PBF_VAR(ctx, Real, rest_density)),
        m_samplingDistance(PBF_VAR(ctx, Real, sampling_distance)),
        m_smoothingLength(PBF_VAR(ctx, Real, smoothing_length)) {}
  VarField<int>* varIterationNumber() { return &m_iterations; }
  VarField<Real>* varRestDensity() { return &m_restDensity; }
  VarField<Real>* varSamplingDistance() { return &m_samplingDistance; }
  VarField<Real>* varSmoothingLength() { return &m_smoothingLength; }
  DeviceArrayField<Coord>* inPosition() { return &m_position; }
  DeviceArrayField<Coord>* inVelocity() { return &m_velocity; }
  NeighborField* inNeighborIndex() { return &m_neighbors; }
  DeviceArrayField<Real>* outDensity() { return &m_density; }
  DeviceArrayField<Real> m_massInv;  // DensityPBD.h:35

  // DensityPBD::constrain (DensityPBD.cu:271-302)
  bool constrain() {
    const int n = m_position.getElementCount();
    if (n == 0) return true;
    NeighborList& nl = m_neighbors.getValue();
    if (nl.isLimited()) throw Error(PBF_ERR_UNSUPPORTED, "DensityPBD takes a dynamic (CSR) neighbor list");
    if (m_density.getElementCount() != n) m_density.setElementCount(n);
    float* x = (float*)m_position.getValue().getDataPtr();
    float* v = (float*)m_velocity.getValue().getDataPtr();
    if (m_massInv.isEmpty())
      Context::check(pbf_pbd_constrain(ctx_->get(), x, v, nl.getIndex().getDataPtr(), nl.getElements().getDataPtr(),
                                       nl.getElements().size(), m_density.getValue().getDataPtr(), n, nullptr),
                     ctx_->get());
    else
      Context::check(pbf_pbd_constrain_mass(ctx_->get(), x, v, nl.getIndex().getDataPtr(),
                                            nl.getElements().getDataPtr(), nl.getElements().size(),
                                            m_massInv.getValue().getDataPtr(), m_density.getValue().getDataPtr(), n,
                                            nullptr),
                     ctx_->get());
    return true;
  }

 private:
  std::shared_ptr<Context> ctx_;
  VarField<int> m_iterations;
  VarField<Real> m_restDensity, m_samplingDistance, m_smoothingLength;
  DeviceArrayField<Coord> m_position, m_velocity;
  NeighborField m_neighbors;
  DeviceArrayField<Real> m_density;
};

class ImplicitViscosity {
  std::shared_ptr<Context> ctx_;

 public:
  explicit ImplicitViscosity(std::shared_ptr<Context> ctx)
      : ctx_(ctx), m_viscosity(PBF_VAR(ctx, Real, viscosity)), m_smoothingLength(PBF_VAR(ctx, Real, smoothing_length)) {}
  void setViscosity(Real mu) { m_viscosity.setValue(mu); }
  void setIterationNumber(int n) {
    ctx_->update([&](pbf_params& p) { p.viscosity_iterations = n; });
  }
  // ImplicitViscosity::constrain (ImplicitViscosity.cu:99-128)
  bool constrain() {
    const int n = m_position.getElementCount();
    if (n == 0) return true;
    NeighborList& nl = m_neighborhood.getValue();
    Context::check(pbf_viscosity(ctx_->get(), (const float*)m_position.getValue().getDataPtr(),
                                 (float*)m_velocity.getValue().getDataPtr(), nl.getIndex().getDataPtr(),
                                 nl.getElements().getDataPtr(), nl.getElements().size(), n, nullptr),
                   ctx_->get());
    return true;
  }
  VarField<Real> m_viscosity;
  VarField<Real> m_smoothingLength;
  DeviceArrayField<Coord> m_position, m_velocity;
  NeighborField m_neighborhood;
};

class BoundaryConstraint {
 public:
  explicit BoundaryConstraint(std::shared_ptr<Context> ctx) : ctx_(std::move(ctx)) {}
  // BoundaryConstraint::setCube (BoundaryConstraint.cu:109-117); only the
  // container box (inverted = true) is available
  void setCube(Coord lo, Coord hi, Real distance, bool inverted = false) {
    if (!inverted) throw Error(PBF_ERR_UNSUPPORTED, "only the container box (inverted) is supported");
    ctx_->update([&](pbf_params& p) {
      p.box_lo[0] = lo.x; p.box_lo[1] = lo.y; p.box_lo[2] = lo.z;
      p.box_hi[0] = hi.x; p.box_hi[1] = hi.y; p.box_hi[2] = hi.z;
      p.box_spacing = distance;
      p.normal_friction = m_normal_friction;
      p.tangent_friction = m_tangent_friction;
    });
  }
  // BoundaryConstraint::constrain(position, velocity, dt) (BoundaryConstraint.cu:84-99)
  bool constrain(DeviceArray<Coord>& position, DeviceArray<Coord>& velocity) {
    Context::check(pbf_box_constrain(ctx_->get(), (float*)position.getDataPtr(), (float*)velocity.getDataPtr(),
                                     position.size(), nullptr),
                   ctx_->get());
    return true;
  }
  bool constrain() { return constrain(m_position.getValue(), m_velocity.getValue()); }
  DeviceArrayField<Coord> m_position, m_velocity;
  Real m_normal_friction = 0.95f;
  Real m_tangent_friction = 0.0f;

 private:
  std::shared_ptr<Context> ctx_;
};

// PositionBasedFluidModel (PositionBasedFluidModel.cpp:20-101). initialize()
// wires the modules as initializeImpl does; step(dt) runs the fused
// pbf_step, stepModules(dt) the module sequence begin -> NeighborQuery ->
// integrate -> DensityPBD -> ImplicitViscosity -> end.
class PositionBasedFluidModel {
  std::shared_ptr<Context> ctx_;  // first: the fields below are built on it

 public:
  explicit PositionBasedFluidModel(const pbf_params* p = nullptr)
      : ctx_(std::make_shared<Context>(p)), m_smoothingLength(PBF_VAR(ctx_, Real, smoothing_length)) {}

  bool initialize() {
    m_nbrQuery = std::make_shared<NeighborQuery>(ctx_);
    m_position.connect(m_nbrQuery->inPosition());
    m_pbdModule = std::make_shared<DensityPBD>(ctx_);
    m_position.connect(m_pbdModule->inPosition());
    m_velocity.connect(m_pbdModule->inVelocity());
    m_nbrQuery->outNeighborhood()->connect(m_pbdModule->inNeighborIndex());
    m_integrator = std::make_shared<ParticleIntegrator>(ctx_);
    m_position.connect(m_integrator->inPosition());
    m_velocity.connect(m_integrator->inVelocity());
    m_forceDensity.connect(m_integrator->inForceDensity());
    m_visModule = std::make_shared<ImplicitViscosity>(ctx_);
    m_visModule->setViscosity(Real(1));
    m_position.connect(&m_visModule->m_position);
    m_velocity.connect(&m_visModule->m_velocity);
    m_nbrQuery->outNeighborhood()->connect(&m_visModule->m_neighborhood);
    ctx_->update([](pbf_params& p) { p.flags |= PBF_FLAG_VISCOSITY; });
    return true;
  }

  void step(Real dt) {
    setDt(dt);
    const int n = m_position.getElementCount();
    if (m_density.getElementCount() != n) m_density.setElementCount(n);
    Context::check(pbf_step(ctx_->get(), (float*)m_position.getValue().getDataPtr(),
                            (float*)m_velocity.getValue().getDataPtr(), m_density.getValue().getDataPtr(), n, 1,
                            nullptr),
                   ctx_->get());
  }

  void stepModules(Real dt) {
    setDt(dt);
    m_integrator->begin();
    m_nbrQuery->compute();
    m_integrator->integrate();
    m_pbdModule->constrain();
    m_visModule->constrain();
    m_integrator->end();
  }

  // Launch settings (no reference counterpart): pbf_autotune on the current
  // state (not changed), optionally kept in the JSON file cache_path; the
  // chosen setting applies to every later step of this model.
  pbf_tune_report autotune(int steps = 3, const char* cache_path = nullptr) {
    pbf_tune_report r;
    Context::check(pbf_autotune(ctx_->get(), (const float*)m_position.getValue().getDataPtr(),
                                (const float*)m_velocity.getValue().getDataPtr(), nullptr, 1,
                                m_position.getElementCount(), steps, cache_path, &r, nullptr),
                   ctx_->get());
    return r;
  }
  void setTuning(const pbf_tuning& t) { Context::check(pbf_set_tuning(ctx_->get(), &t), ctx_->get()); }
  pbf_tuning getTuning() const {
    pbf_tuning t;
    Context::check(pbf_get_tuning(ctx_->get(), &t), ctx_->get());
    return t;
  }

  std::shared_ptr<Context> context() { return ctx_; }
  VarField<Real> m_smoothingLength;
  DeviceArrayField<Coord> m_position, m_velocity, m_forceDensity;
  DeviceArrayField<Real> m_density;

  std::shared_ptr<NeighborQuery> m_nbrQuery;
  std::shared_ptr<DensityPBD> m_pbdModule;
  std::shared_ptr<ParticleIntegrator> m_integrator;
  std::shared_ptr<ImplicitViscosity> m_visModule;

 private:
  void setDt(Real dt) {
    if (ctx_->params().dt != dt) ctx_->update([&](pbf_params& p) { p.dt = dt; });
  }
};

#undef PBF_VAR

}  // namespace physika_amd
