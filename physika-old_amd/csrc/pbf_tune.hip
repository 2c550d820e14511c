// The launch-setting autotuner (pbf_autotune): times every pbf_tuning setting
// on the caller's workload through the library's own entry points, applies
// the fastest to the context and keeps it in a JSON cache file. Host code
// only. No reference counterpart: the reference launches every kernel in
// blocks of BLOCK_SIZE = 64 (Core/Utility/cuda_utilities.h:34).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "pbf.h"
#include "pbf_internal.h"

namespace {

constexpr int kBlocks[4] = {64, 128, 256, 512};
constexpr int kStageAlways = 1, kStageNever = -1;

struct DeviceCopy {
  float* x = nullptr;
  float* v = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~DeviceCopy() {
    if (x) (void)hipFree(x);
    if (v) (void)hipFree(v);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// The cache key: device architecture, particle count in half octaves, scene
// count in octaves, mean neighbors per particle in octaves (a lattice and a
// dense blob of one size prefer different settings), h/d, iterations and the
// module flags.
std::string tune_key(const pbf_params& p, int n, int nscenes, double mean_nbrs, int device) {
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
  }
  char buf[128];
  std::snprintf(buf, sizeof buf, "%s n2^%.1f s2^%d k2^%d h/d%.2f K%d f%u", arch.c_str(),
                std::floor(std::log2(n) * 2.0) / 2.0, (int)std::floor(std::log2(nscenes)),
                (int)std::floor(std::log2(std::max(mean_nbrs, 1.0))), p.smoothing_length / p.sampling_distance,
                p.iterations, p.flags & (PBF_FLAG_VISCOSITY | PBF_FLAG_BOX));
  return buf;
}

// One entry per line: "key": {"pass_block": B, "stage_min_particles": S, "persistent": P, "ms_per_step": T}
struct Entry {
  std::string key, body;
};

std::vector<Entry> read_cache(const char* path) {
  std::vector<Entry> out;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t a = line.find('"');
    if (a == std::string::npos) continue;
    const size_t b = line.find('"', a + 1);
    const size_t c = line.find('{', b == std::string::npos ? a : b);
    const size_t d = line.rfind('}');
    if (b == std::string::npos || c == std::string::npos || d == std::string::npos || d < c) continue;
    out.push_back({line.substr(a + 1, b - a - 1), line.substr(c, d - c + 1)});
  }
  return out;
}

bool field(const std::string& body, const char* name, double* v) {
  const std::string tag = std::string("\"") + name + "\":";
  const size_t at = body.find(tag);
  if (at == std::string::npos) return false;
  char* end = nullptr;
  const char* p = body.c_str() + at + tag.size();
  *v = std::strtod(p, &end);
  return end != p;
}

bool write_cache(const char* path, const std::vector<Entry>& entries) {
  // a file of this process, renamed over the cache (rename is atomic: a reader
  // sees the old or the new file; two processes tuning at once keep the last)
  const std::string tmp = std::string(path) + ".tmp" + std::to_string((long)getpid());
  {
    std::ofstream f(tmp, std::ios::trunc);
    if (!f) return false;
    f << "{\n";
    for (size_t i = 0; i < entries.size(); ++i)
      f << "  \"" << entries[i].key << "\": " << entries[i].body << (i + 1 < entries.size() ? ",\n" : "\n");
    f << "}\n";
    if (!f) return false;
  }
  return std::rename(tmp.c_str(), path) == 0;
}

// Per-step time of the context's current setting: a warm-up call, then the
// better of two timed calls, each from a fresh copy of the workload.
int time_steps(pbf_ctx* c, const float* x, const float* v, const int* off, int nscenes, int n, int steps,
               DeviceCopy& w, hipStream_t st, double* ms) {
  const size_t bytes = sizeof(float) * 3 * (size_t)n;
  double best = 1e300;
  for (int rep = 0; rep < 3; ++rep) {
    if (hipMemcpyAsync(w.x, x, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(w.v, v, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return pbf::ctx_fail(c, PBF_ERR_HIP, "pbf_autotune: copy of the workload failed");
    if (hipEventRecord(w.e0, st) != hipSuccess) return pbf::ctx_fail(c, PBF_ERR_HIP, "pbf_autotune: event record");
    const int s = off ? pbf_batch_step(c, w.x, w.v, nullptr, off, nscenes, n, steps, st)
                      : pbf_step(c, w.x, w.v, nullptr, n, steps, st);
    if (s) return s;
    if (hipEventRecord(w.e1, st) != hipSuccess || hipEventSynchronize(w.e1) != hipSuccess)
      return pbf::ctx_fail(c, PBF_ERR_HIP, "pbf_autotune: event wait");
    float t = 0.f;
    if (hipEventElapsedTime(&t, w.e0, w.e1) != hipSuccess) return pbf::ctx_fail(c, PBF_ERR_HIP, "pbf_autotune: timing");
    if (rep > 0 && t < best) best = t;
  }
  *ms = best / steps;
  return PBF_OK;
}

}  // namespace

extern "C" int pbf_autotune(pbf_ctx* c, const float* d_x, const float* d_v, const int* d_scene_offsets, int nscenes,
                            int n, int steps, const char* cache_path, pbf_tune_report* out, void* stream) {
  if (!c) return PBF_ERR_INVALID;
  if (!out || !d_x || !d_v || n < 1 || steps < 1 || nscenes < 1 || (nscenes > 1 && !d_scene_offsets))
    return pbf::ctx_fail(c, PBF_ERR_INVALID, "pbf_autotune: bad argument");
  std::memset(out, 0, sizeof(*out));
  hipStream_t st = (hipStream_t)stream;
  pbf_params p;
  pbf_tuning start;
  int device = 0;
  if (pbf_get_params(c, &p) || pbf_get_tuning(c, &start) || hipGetDevice(&device) != hipSuccess)
    return pbf::ctx_fail(c, PBF_ERR_HIP, "pbf_autotune: context query failed");
  DeviceCopy w;
  const size_t bytes = sizeof(float) * 3 * (size_t)n;
  if (hipMalloc(&w.x, bytes) != hipSuccess || hipMalloc(&w.v, bytes) != hipSuccess ||
      hipEventCreate(&w.e0) != hipSuccess || hipEventCreate(&w.e1) != hipSuccess) {
    (void)hipGetLastError();
    return pbf::ctx_fail(c, PBF_ERR_ALLOC, "pbf_autotune: no memory for a copy of the workload");
  }
  const int* off = nscenes > 1 || d_scene_offsets ? d_scene_offsets : nullptr;
  // one step of a copy: the neighbor count of the workload for the key
  double mean_nbrs = 0;
  {
    if (hipMemcpyAsync(w.x, d_x, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(w.v, d_v, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return pbf::ctx_fail(c, PBF_ERR_HIP, "pbf_autotune: copy of the workload failed");
    const int s = off ? pbf_batch_step(c, w.x, w.v, nullptr, off, nscenes, n, 1, st)
                      : pbf_step(c, w.x, w.v, nullptr, n, 1, st);
    if (s) return s;
    pbf_stats stats;
    if (pbf_get_stats(c, &stats)) return pbf::ctx_fail(c, PBF_ERR_HIP, "pbf_autotune: statistics");
    mean_nbrs = (double)stats.neighbor_pairs / n;
  }
  const std::string key = tune_key(p, n, nscenes, mean_nbrs, device);
  std::snprintf(out->key, sizeof out->key, "%s", key.c_str());

  std::vector<Entry> cache;
  if (cache_path) {
    cache = read_cache(cache_path);
    for (const Entry& e : cache) {
      double b, sm, ps = 0, ms = 0;
      if (e.key != key || !field(e.body, "pass_block", &b) || !field(e.body, "stage_min_particles", &sm)) continue;
      (void)field(e.body, "persistent", &ps);
      (void)field(e.body, "ms_per_step", &ms);
      const pbf_tuning t{(int)b, (int)sm, (int)ps};
      if (pbf_set_tuning(c, &t)) break;  // not a valid setting: time again
      out->best = t;
      out->best_ms = ms;
      out->from_cache = 1;
      return PBF_OK;
    }
  }

  // the default first, then the rest; a setting must beat the best so far by 1%
  std::vector<pbf_tuning> cand{{256, 0, 0}};
  for (int persistent : {0, 1})
    for (int b : kBlocks)
      for (int stage : {kStageAlways, kStageNever}) cand.push_back({b, stage, persistent});
  pbf_tuning best = cand[0];
  double best_ms = 0;
  for (size_t k = 0; k < cand.size(); ++k) {
    double ms = 0;
    int s = pbf_set_tuning(c, &cand[k]);
    if (!s) s = time_steps(c, d_x, d_v, off, nscenes, n, steps, w, st, &ms);
    if (s) {
      (void)pbf_set_tuning(c, &start);
      return s;
    }
    if (k == 0) {
      out->default_ms = ms;
      best_ms = ms;
    } else if (ms < 0.99 * best_ms) {
      best = cand[k];
      best_ms = ms;
    }
  }
  out->candidates = (int)cand.size();
  out->best = best;
  out->best_ms = best_ms;
  int s = pbf_set_tuning(c, &best);
  if (s) return s;
  if (cache_path) {
    char body[192];
    std::snprintf(body, sizeof body,
                  "{\"pass_block\": %d, \"stage_min_particles\": %d, \"persistent\": %d, \"ms_per_step\": %.5f}",
                  best.pass_block, best.stage_min_particles, best.persistent, best_ms);
    bool found = false;
    for (Entry& e : cache)
      if (e.key == key) {
        e.body = body;
        found = true;
      }
    if (!found) cache.push_back({key, body});
    if (!write_cache(cache_path, cache))
      return pbf::ctx_fail(c, PBF_ERR_INVALID, std::string("pbf_autotune: cannot write ") + cache_path);
  }
  return PBF_OK;
}
