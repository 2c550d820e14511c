// GPU check of the C++ module mirror (physika-old_amd/cpp/pbf_modules.h):
// the module sequence of PositionBasedFluidModel::step against the fused
// step, the dynamic and limited NeighborQuery, DensityPBD with unit inverse
// masses. Prints one "PASS <case>" or "FAIL <case>" line per case; exit
// status 0 when all pass.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <vector>

#include "pbf_modules.h"

using namespace physika_amd;

static std::vector<Coord> lattice(int nx, int ny, int nz, float x0, float y0, float z0, float s, uint64_t seed) {
  std::vector<Coord> p;
  uint64_t st = seed;
  auto u = [&]() {  // SplitMix64 in [0, 1)
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        p.push_back({(float)(x0 + s * i + (2 * u() - 1) * 0.1 * s), (float)(y0 + s * j + (2 * u() - 1) * 0.1 * s),
                     (float)(z0 + s * k + (2 * u() - 1) * 0.1 * s)});
  return p;
}

static int failures = 0;
static void report(bool ok, const char* name, double v) {
  std::printf("%s %s (%.3e)\n", ok ? "PASS" : "FAIL", name, v);
  std::fflush(stdout);
  if (!ok) ++failures;
}

static double max_diff(const std::vector<Coord>& a, const std::vector<Coord>& b) {
  double m = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    m = std::fmax(m, std::fabs((double)a[i].x - b[i].x));
    m = std::fmax(m, std::fabs((double)a[i].y - b[i].y));
    m = std::fmax(m, std::fabs((double)a[i].z - b[i].z));
  }
  return m;
}

int main() {
  try {
    const std::vector<Coord> x0 = lattice(16, 16, 16, 0.3f, 0.5f, 0.3f, 0.005f, 7);
    const int n = (int)x0.size();
    std::vector<Coord> v0((size_t)n, Coord{0.f, 0.f, 0.f});

    // fused step vs the module sequence, viscosity on as in the reference model
    PositionBasedFluidModel fused, modules;
    fused.initialize();
    modules.initialize();
    for (auto* m : {&fused, &modules}) {
      m->m_position.getValue().assign(x0);
      m->m_velocity.getValue().assign(v0);
      m->m_forceDensity.setElementCount(n);
    }
    for (int s = 0; s < 3; ++s) {
      fused.step(0.001f);
      modules.stepModules(0.001f);
    }
    const double dx = max_diff(fused.m_position.getValue().toHost(), modules.m_position.getValue().toHost());
    const double dv = max_diff(fused.m_velocity.getValue().toHost(), modules.m_velocity.getValue().toHost());
    report(dx <= 1e-6, "fused step == module sequence, positions", dx);
    report(dv <= 1e-3, "fused step == module sequence, velocities", dv);

    // the autotuner leaves the state alone; any setting it applies gives the same bits
    PositionBasedFluidModel tuned, plain;
    tuned.initialize();
    plain.initialize();
    for (auto* m : {&tuned, &plain}) {
      m->m_position.getValue().assign(x0);
      m->m_velocity.getValue().assign(v0);
      m->m_forceDensity.setElementCount(n);
    }
    const pbf_tune_report tr = tuned.autotune(2);
    const double moved = max_diff(tuned.m_position.getValue().toHost(), x0);
    const pbf_tuning now = tuned.getTuning();
    report(moved == 0.0 && tr.candidates == 17 && now.pass_block == tr.best.pass_block &&
               now.persistent == tr.best.persistent,
           "autotune: state unchanged, setting applied", (double)tr.best.pass_block);
    for (int s = 0; s < 2; ++s) {
      tuned.step(0.001f);
      plain.step(0.001f);
    }
    const double dt_bits = max_diff(tuned.m_position.getValue().toHost(), plain.m_position.getValue().toHost());
    report(dt_bits == 0.0, "tuned setting == default setting, positions", dt_bits);

    // dynamic and limited neighbor rows
    auto ctx = std::make_shared<Context>();
    NeighborQuery nq(ctx);
    DeviceArray<Coord> pos;
    pos.assign(x0);
    nq.inPosition()->setValue(pos);
    nq.compute();
    NeighborList& nl = nq.outNeighborhood()->getValue();
    const std::vector<int> idx = nl.getIndex().toHost();
    pbf_stats st;
    Context::check(pbf_get_stats(ctx->get(), &st), ctx->get());
    const long long nel = nl.getElements().size();
    report(!nl.isLimited() && nl.size() == n && (int)idx.size() == n && nel == st.neighbor_pairs && idx[0] == 0,
           "NeighborQuery dynamic rows", (double)nel);
    nq.setNeighborSizeLimit(5);
    nq.compute();
    const std::vector<int> cnt = nq.outNeighborhood()->getValue().getIndex().toHost();
    int worst = 0;
    for (int i = 0; i < n; ++i) {
      const int end = i + 1 < n ? idx[(size_t)i + 1] : (int)nel;
      worst = cnt[(size_t)i] == std::min(5, end - idx[(size_t)i]) ? worst : 1;
    }
    report(worst == 0 && nq.outNeighborhood()->getValue().getNeighborLimit() == 5, "NeighborQuery limited rows",
           (double)worst);

    // DensityPBD with unit inverse masses == without
    nq.setNeighborSizeLimit(0);
    nq.compute();
    DensityPBD a(ctx), b(ctx);
    DeviceArray<Coord> xa, va, xb, vb;
    xa.assign(x0);
    va.assign(v0);
    xb.assign(x0);
    vb.assign(v0);
    a.inPosition()->setValue(xa);
    a.inVelocity()->setValue(va);
    b.inPosition()->setValue(xb);
    b.inVelocity()->setValue(vb);
    nq.outNeighborhood()->connect(a.inNeighborIndex());
    nq.outNeighborhood()->connect(b.inNeighborIndex());
    b.m_massInv.setValue(DeviceArray<Real>(n));
    b.m_massInv.getValue().assign(std::vector<Real>((size_t)n, 1.0f));
    a.varIterationNumber()->setValue(3);
    a.constrain();
    b.constrain();
    const double dm = max_diff(xa.toHost(), xb.toHost());
    report(dm <= 1e-6 && a.varIterationNumber()->getValue() == 3, "DensityPBD m_massInv = 1", dm);
  } catch (const Error& e) {
    std::printf("FAIL exception: %s (status %d)\n", e.what(), e.status);
    return 2;
  }
  return failures ? 1 : 0;
}
