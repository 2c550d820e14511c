// Host-side checks of libpbf.so that need no GPU, built with AddressSanitizer
// and UndefinedBehaviorSanitizer on the host code of the library and of this
// check program (run by tests/test_sanitizers.py): parameter validation, the derivation of
// the grid and the kernel constants inside pbf_create (it runs before the
// first HIP call), the status strings, and every entry point's refusal of a
// null context. Prints one PASS line per check; exits 1 on a failed check (a
// sanitizer report aborts the process).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "pbf.h"

static int failures = 0;

static void check(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}

// pbf_create on parameters `p`: the status, with out left NULL on failure.
static int create_status(const pbf_params& p, bool* out_null) {
  pbf_ctx* c = reinterpret_cast<pbf_ctx*>(0x1);
  const int st = pbf_create(&p, &c);
  *out_null = c == nullptr;
  if (st == PBF_OK && c) pbf_destroy(c);
  return st;
}

int main() {
  check(pbf_abi_version() == PBF_ABI_VERSION, "abi version");
  const int codes[] = {PBF_OK, PBF_ERR_INVALID, PBF_ERR_HIP, PBF_ERR_CAPACITY, PBF_ERR_ALLOC, PBF_ERR_UNSUPPORTED, 7};
  bool strings = true;
  for (int s : codes) strings = strings && pbf_status_string(s) && std::strlen(pbf_status_string(s)) > 0;
  check(strings, "status strings");

  pbf_params_default(nullptr);
  pbf_params d;
  std::memset(&d, 0xff, sizeof(d));
  pbf_params_default(&d);
  check(d.smoothing_length == 0.0125f && d.iterations == 5 && d.flags == 0u && d.scene_hi[2] == 1.0f,
        "defaults");

  // invalid parameters are refused before anything is allocated
  const float nan = std::numeric_limits<float>::quiet_NaN();
  struct Bad {
    const char* what;
    void (*set)(pbf_params&, float);
    float v;
  } bad[] = {
      {"h <= 0", [](pbf_params& p, float v) { p.smoothing_length = v; }, 0.0f},
      {"h NaN", [](pbf_params& p, float v) { p.smoothing_length = v; }, nan},
      {"d < 0", [](pbf_params& p, float v) { p.sampling_distance = v; }, -1.0f},
      {"dt inf", [](pbf_params& p, float v) { p.dt = v; }, INFINITY},
      {"iterations < 0", [](pbf_params& p, float v) { p.iterations = (int)v; }, -1.0f},
      {"iterations > 1000", [](pbf_params& p, float v) { p.iterations = (int)v; }, 1001.0f},
      {"mask_words < 0", [](pbf_params& p, float v) { p.mask_words = (int)v; }, -3.0f},
      {"list_slots < 0", [](pbf_params& p, float v) { p.list_slots = (int)v; }, -3.0f},
      {"unknown flag", [](pbf_params& p, float) { p.flags = 64u; }, 0.0f},
      {"viscosity < 0", [](pbf_params& p, float v) { p.viscosity = v; }, -1.0f},
      {"box hi <= lo", [](pbf_params& p, float v) { p.box_hi[1] = v; }, 0.0f},
      {"scene hi <= lo", [](pbf_params& p, float v) { p.scene_hi[0] = v; }, -1.0f},
      {"gravity NaN", [](pbf_params& p, float v) { p.gravity[2] = v; }, nan},
      {"lambda_eps NaN", [](pbf_params& p, float v) { p.lambda_eps = v; }, nan},
  };
  for (const Bad& b : bad) {
    pbf_params p = d;
    b.set(p, b.v);
    bool null_out = false;
    const int st = create_status(p, &null_out);
    char msg[96];
    std::snprintf(msg, sizeof(msg), "create refuses %s", b.what);
    check(st == PBF_ERR_INVALID && null_out, msg);
  }
  {  // a grid too large for 32-bit cell indices
    pbf_params p = d;
    p.smoothing_length = 1e-5f;
    bool null_out = false;
    const int st = create_status(p, &null_out);
    check(st != PBF_OK && null_out, "create refuses a grid beyond 32-bit cell indices");
  }
  {  // valid parameters: the grid and constants are derived on the host; with
     // no GPU the first HIP call fails and the partial context is released
    bool null_out = false;
    const int st = create_status(d, &null_out);
    check(st == PBF_OK || (st == PBF_ERR_HIP && null_out), "create with valid parameters");
    pbf_params p = d;
    p.smoothing_length = 0.006f;  // the reference model's default
    p.iterations = 3;
    const int st2 = create_status(p, &null_out);
    check(st2 == PBF_OK || (st2 == PBF_ERR_HIP && null_out), "create with the reference defaults");
  }
  check(pbf_create(&d, nullptr) == PBF_ERR_INVALID, "create needs an out pointer");

  // a null context is refused by every entry point
  float f = 0.0f;
  int i = 0;
  long long t = 0;
  pbf_stats stats;
  pbf_profile prof;
  pbf_params got;
  bool refused = true;
  refused &= pbf_destroy(nullptr) == PBF_ERR_INVALID;
  refused &= pbf_set_params(nullptr, &d) == PBF_ERR_INVALID;
  refused &= pbf_get_params(nullptr, &got) == PBF_ERR_INVALID;
  refused &= pbf_reserve(nullptr, 10) == PBF_ERR_INVALID;
  refused &= pbf_reserve_batch(nullptr, 10, 2) == PBF_ERR_INVALID;
  refused &= pbf_neighbor_query(nullptr, &f, 1, &i, &i, 1, &t, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_neighbor_query_fixed(nullptr, &f, 1, 4, &i, &i, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_integrate(nullptr, &f, &f, nullptr, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_density(nullptr, &f, &i, &i, 1, &f, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_pbd_constrain(nullptr, &f, &f, &i, &i, 1, &f, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_pbd_constrain_mass(nullptr, &f, &f, &i, &i, 1, &f, &f, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_step(nullptr, &f, &f, &f, 1, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_batch_step(nullptr, &f, &f, &f, &i, 1, 1, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_rollout_forward(nullptr, &f, &f, &f, nullptr, 1, 1, 1, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_rollout_backward(nullptr, &f, &f, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_step_neighbors(nullptr, &i, &i, 1, &t, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_get_stats(nullptr, &stats) == PBF_ERR_INVALID;
  refused &= pbf_viscosity(nullptr, &f, &f, &i, &i, 1, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_box_constrain(nullptr, &f, &f, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_set_profiling(nullptr, 1) == PBF_ERR_INVALID;
  refused &= pbf_get_profile(nullptr, &prof) == PBF_ERR_INVALID;
  refused &= pbf_pbd_constrain_noalloc(nullptr, &f, &f, &i, &i, 1, &f, 1, nullptr) == PBF_ERR_INVALID;
  refused &= pbf_rollout_forward_f64(nullptr, nullptr, nullptr, nullptr, nullptr, 1, 1, 1, 1, nullptr) ==
             PBF_ERR_INVALID;
  refused &= pbf_rollout_backward_f64(nullptr, nullptr, nullptr, nullptr) == PBF_ERR_INVALID;
  pbf_tuning tun{256, 0, 0};
  pbf_tune_report rep;
  refused &= pbf_set_tuning(nullptr, &tun) == PBF_ERR_INVALID;
  refused &= pbf_get_tuning(nullptr, &tun) == PBF_ERR_INVALID;
  refused &= pbf_autotune(nullptr, &f, &f, nullptr, 1, 1, 1, nullptr, &rep, nullptr) == PBF_ERR_INVALID;
  check(refused, "null context refused by every entry point");
  check(pbf_last_error(nullptr) != nullptr, "last error of a null context");

  std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
  return failures ? 1 : 0;
}
