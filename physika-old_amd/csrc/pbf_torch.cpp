// pybind11 layer of the reverse mode (a PyTorch C++ extension, built by
// build_torch_ext.py into ../_build): the differentiable rollout as a
// torch::autograd::Function whose forward and backward call the library's C
// ABI (pbf_rollout_forward / pbf_rollout_backward and their double-precision
// forms), and the two calls on their own. The reference's Python layer is a
// pybind11 module too (Python/PyPhysIKA.cpp:12); it exposes no fluid.
//
// A context handle is the pbf_ctx* as an integer, a stream the hipStream_t as
// an integer (torch.cuda.current_stream().cuda_stream), so this file needs no
// HIP header. Each context keeps one rollout per precision: every forward
// through this module takes a new generation number, and a backward whose
// forward is no longer the context's latest of that precision is refused.
#include <torch/extension.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>

#include "pbf.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

std::mutex g_mu;
std::unordered_map<uintptr_t, int64_t> g_generation;  // (context, precision) -> latest forward

uintptr_t gen_key(int64_t handle, bool f64) { return (uintptr_t)handle * 2u + (f64 ? 1u : 0u); }

// Releases the GIL around a blocking library call when this thread holds it
// (a Python call does; torch's autograd engine threads run backward without it).
struct ReleaseIfHeld {
  PyThreadState* saved = nullptr;
  ReleaseIfHeld() {
    if (PyGILState_Check()) saved = PyEval_SaveThread();
  }
  ~ReleaseIfHeld() {
    if (saved) PyEval_RestoreThread(saved);
  }
};

pbf_ctx* ctx_of(int64_t handle) {
  TORCH_CHECK(handle != 0, "null pbf context");
  return reinterpret_cast<pbf_ctx*>((uintptr_t)handle);
}

// A failed call raises the Python binding's PBFError(status, message).
void check(int status, pbf_ctx* c) {
  if (status == PBF_OK) return;
  pybind11::gil_scoped_acquire gil;
  const std::string msg = c ? pbf_last_error(c) : pbf_status_string(status);
  pybind11::object cls = pybind11::module_::import("physika_old_amd._capi").attr("PBFError");
  pybind11::object err = cls(status, msg);
  PyErr_SetObject(cls.ptr(), err.ptr());
  throw pybind11::error_already_set();
}

void check_state(const Tensor& t, const char* what, int64_t n, const Tensor& like) {
  TORCH_CHECK(t.dim() == 2 && t.size(1) == 3 && (n < 0 || t.size(0) == n), what, ": expected an (n, 3) tensor");
  TORCH_CHECK(t.is_contiguous() && t.device() == like.device() && t.scalar_type() == like.scalar_type(), what,
              ": expected a contiguous tensor of the positions' device and dtype");
}

// pbf_rollout_forward(_f64) on x, v (in place) and rho; returns the forward's generation.
int64_t forward_call(int64_t handle, Tensor& x, Tensor& v, const c10::optional<Tensor>& rho, int64_t nsteps,
                     int64_t every, const c10::optional<Tensor>& offsets, int64_t stream) {
  pbf_ctx* c = ctx_of(handle);
  TORCH_CHECK(x.is_cuda(), "the rollout runs on a HIP device");
  TORCH_CHECK(x.scalar_type() == torch::kFloat32 || x.scalar_type() == torch::kFloat64,
              "positions: float32 or float64");
  check_state(x, "position", -1, x);
  check_state(v, "velocity", x.size(0), x);
  const bool has_rho = rho && rho->defined();
  TORCH_CHECK(!has_rho || (rho->numel() == x.size(0) && rho->scalar_type() == x.scalar_type() &&
                           rho->is_contiguous() && rho->device() == x.device()),
              "rho: expected n contiguous values of the positions' dtype and device");
  const bool f64 = x.scalar_type() == torch::kFloat64;
  const int* off = nullptr;
  int nscenes = 1;
  if (offsets && offsets->defined()) {
    TORCH_CHECK(offsets->scalar_type() == torch::kInt32 && offsets->dim() == 1 && offsets->numel() >= 2 &&
                    offsets->is_contiguous() && offsets->device() == x.device(),
                "scene offsets: expected nscenes + 1 contiguous int32 on the positions' device");
    off = offsets->data_ptr<int>();
    nscenes = (int)offsets->numel() - 1;
  }
  void* st = reinterpret_cast<void*>((uintptr_t)stream);
  const int n = (int)x.size(0);
  int s;
  {
    ReleaseIfHeld nogil;
    s = f64 ? pbf_rollout_forward_f64(c, x.data_ptr<double>(), v.data_ptr<double>(),
                                      has_rho ? rho->data_ptr<double>() : nullptr, off, nscenes, n, (int)nsteps,
                                      (int)every, st)
            : pbf_rollout_forward(c, x.data_ptr<float>(), v.data_ptr<float>(),
                                  has_rho ? rho->data_ptr<float>() : nullptr, off, nscenes, n, (int)nsteps,
                                  (int)every, st);
  }
  check(s, c);
  std::lock_guard<std::mutex> lock(g_mu);
  return ++g_generation[gen_key(handle, f64)];
}

// pbf_rollout_backward(_f64): (dL/dx_N, dL/dv_N) -> (dL/dx_0, dL/dv_0) in place.
void backward_call(int64_t handle, Tensor& xb, Tensor& vb, int64_t generation, int64_t stream) {
  pbf_ctx* c = ctx_of(handle);
  TORCH_CHECK(xb.scalar_type() == torch::kFloat32 || xb.scalar_type() == torch::kFloat64, "xbar: float32 or float64");
  check_state(xb, "xbar", -1, xb);
  check_state(vb, "vbar", xb.size(0), xb);
  const bool f64 = xb.scalar_type() == torch::kFloat64;
  if (generation > 0) {
    std::lock_guard<std::mutex> lock(g_mu);
    TORCH_CHECK(g_generation[gen_key(handle, f64)] == generation,
                "the context ran another rollout of this precision since this one; its checkpoints are gone");
  }
  void* st = reinterpret_cast<void*>((uintptr_t)stream);
  int s;
  {
    ReleaseIfHeld nogil;
    s = f64 ? pbf_rollout_backward_f64(c, xb.data_ptr<double>(), vb.data_ptr<double>(), st)
            : pbf_rollout_backward(c, xb.data_ptr<float>(), vb.data_ptr<float>(), st);
  }
  check(s, c);
}

// The differentiable rollout: x_N, v_N, rho_N from x0, v0; gradients flow to
// x0 and v0 through the library's adjoint kernels (rho_N carries none).
struct Rollout : public torch::autograd::Function<Rollout> {
  static variable_list forward(AutogradContext* fc, Tensor x0, Tensor v0, int64_t handle, int64_t nsteps,
                               int64_t every, c10::optional<Tensor> offsets, int64_t stream) {
    Tensor x = x0.detach().clone().contiguous();
    Tensor v = v0.detach().clone().contiguous();
    Tensor rho = torch::empty({x.size(0)}, x.options());
    const int64_t gen = forward_call(handle, x, v, rho, nsteps, every, offsets, stream);
    fc->saved_data["handle"] = handle;
    fc->saved_data["generation"] = gen;
    fc->saved_data["stream"] = stream;
    fc->mark_non_differentiable({rho});
    return {x, v, rho};
  }

  static variable_list backward(AutogradContext* fc, variable_list g) {
    const Tensor& gx = g[0];
    const Tensor& gv = g[1];
    const Tensor& like = gx.defined() ? gx : gv;
    Tensor xb = gx.defined() ? gx.detach().clone().contiguous() : torch::zeros_like(like);
    Tensor vb = gv.defined() ? gv.detach().clone().contiguous() : torch::zeros_like(like);
    backward_call(fc->saved_data["handle"].toInt(), xb, vb, fc->saved_data["generation"].toInt(),
                  fc->saved_data["stream"].toInt());
    return {xb, vb, Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
  }
};

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "Reverse mode of the PBD fluid step (libpbf.so) for torch.autograd";
  m.def("abi_version", &pbf_abi_version);
  m.def("rollout",
        [](Tensor x0, Tensor v0, int64_t handle, int64_t nsteps, int64_t every, c10::optional<Tensor> offsets,
           int64_t stream) { return Rollout::apply(x0, v0, handle, nsteps, every, offsets, stream); },
        "differentiable nsteps of the fluid step: (x_N, v_N, rho_N)");
  m.def("rollout_forward",
        [](int64_t handle, Tensor x, Tensor v, c10::optional<Tensor> rho, int64_t nsteps, int64_t every,
           c10::optional<Tensor> offsets, int64_t stream) {
          return forward_call(handle, x, v, rho, nsteps, every, offsets, stream);
        },
        "pbf_rollout_forward(_f64) in place; returns the rollout's generation");
  m.def("rollout_backward",
        [](int64_t handle, Tensor xb, Tensor vb, int64_t generation, int64_t stream) {
          backward_call(handle, xb, vb, generation, stream);
        },
        "pbf_rollout_backward(_f64) in place; generation 0 skips the check");
}
