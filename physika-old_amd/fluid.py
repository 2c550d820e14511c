"""Python mirror of the reference's fluid modules over the C ABI.

Class and field names follow the reference so that code written against
PhysIKA's modules reads the same:

  NeighborQuery        Framework/Topology/NeighborQuery.h:13-95
                       inRadius, inPosition, outNeighborhood; compute()
  NeighborList         Framework/Topology/NeighborList.h:9-125
  ParticleIntegrator   Dynamics/ParticleSystem/ParticleIntegrator.h:8-54
                       inPosition, inVelocity, inForceDensity; begin/integrate/end
  DensityPBD           Dynamics/ParticleSystem/DensityPBD.h:18-88
                       varIterationNumber, varRestDensity, varSamplingDistance,
                       varSmoothingLength, inPosition, inVelocity,
                       inNeighborIndex, outDensity; constrain()
  PositionBasedFluidModel
                       Dynamics/ParticleSystem/PositionBasedFluidModel.h:25-60
                       m_smoothingLength, m_position, m_velocity, m_forceDensity;
                       step(dt)

Array fields hold torch tensors on a HIP device ((n, 3) float32 for
positions/velocities/forces, (n,) float32 for density); connected fields share
one tensor, as the reference's fields share one device array
(Framework/Framework/FieldArray.h:166-185). Every compute call goes through
libpbf.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Optional

import numpy as np
import torch

from . import _capi


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _stream_int(device) -> int:
    """The current stream's hipStream_t as an integer (0: the null stream)."""
    return int(torch.cuda.current_stream(device).cuda_stream)


def _check_vec3(t: torch.Tensor, name: str, n: Optional[int] = None, device=None,
                dtype=torch.float32) -> torch.Tensor:
    """A (n, 3) contiguous device tensor of `dtype`, checked, not copied: the
    in-place entry points write into it, so a copy would leave the caller's
    tensor unchanged."""
    if not isinstance(t, torch.Tensor) or t.dtype != dtype or t.dim() != 2 or t.shape[1] != 3 \
            or not t.is_cuda:
        shape = tuple(t.shape) if isinstance(t, torch.Tensor) else type(t).__name__
        dt = t.dtype if isinstance(t, torch.Tensor) else ""
        raise ValueError(f"{name}: expected a (n, 3) {dtype} device tensor, got {shape} {dt}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor (the call writes it in place)")
    if n is not None and t.shape[0] != n:
        raise ValueError(f"{name}: expected {n} rows, got {t.shape[0]}")
    if device is not None and t.device != device:
        raise ValueError(f"{name}: expected a tensor on {device}, got {t.device}")
    return t


def _check_scalar(t: Optional[torch.Tensor], name: str, n: int, device, dtype=torch.float32):
    """None, or a contiguous 1-D tensor of `dtype` on `device` with at least n
    entries."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or t.dtype != dtype or t.dim() != 1 or t.device != device \
            or not t.is_contiguous() or t.shape[0] < n:
        raise ValueError(f"{name}: expected a contiguous 1-D {dtype} tensor on {device} with at least {n} entries")
    return t


def _check_scene_offsets(off: torch.Tensor, n: int, device, validate: bool = True) -> torch.Tensor:
    """nscenes + 1 int32 offsets on `device`: first 0, last n, non-decreasing
    (checked on the host when `validate`; one small device-to-host copy)."""
    if not isinstance(off, torch.Tensor) or off.dtype != torch.int32 or off.dim() != 1 or off.device != device \
            or off.shape[0] < 2 or not off.is_contiguous():
        raise ValueError("scene_offsets: expected a contiguous int32 device tensor of nscenes+1 offsets")
    if validate:
        h = off.cpu()
        if int(h[0]) != 0 or int(h[-1]) != n or bool((h[1:] < h[:-1]).any()):
            raise ValueError(f"scene_offsets: must start at 0, end at {n} and never decrease")
    return off


def _check_nbr(nbr: "NeighborList", n: int, device) -> None:
    if nbr.isLimited():
        raise ValueError("the module passes take a dynamic (CSR) neighbor list, not a limited one")
    if nbr.index.dtype != torch.int32 or nbr.index.device != device or nbr.index.shape != (n,) \
            or not nbr.index.is_contiguous():
        raise ValueError(f"neighbor index: expected {n} int32 row starts on {device}")
    if nbr.elements.dtype != torch.int32 or nbr.elements.device != device or not nbr.elements.is_contiguous():
        raise ValueError(f"neighbor elements: expected a contiguous int32 tensor on {device}")


class Context:
    """A pbf_ctx (include/pbf.h) bound to one HIP device."""

    # scene offsets are checked on the host before each batched call (one small
    # copy); a caller that builds them once may turn this off
    validate_offsets = True

    def __init__(self, device="cuda:0", **params):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the PBD fluid step runs on a HIP device only")
        self._lib = _capi.lib()
        self._ctx = C.c_void_p()
        p = _capi.params_from(**params)
        with torch.cuda.device(self.device):
            _capi.check(self._lib.pbf_create(C.byref(p), C.byref(self._ctx)))

    # -- parameters (the VarField setters of the reference) --------------------
    def params(self) -> dict:
        p = _capi.pbf_params()
        _capi.check(self._lib.pbf_get_params(self._ctx, C.byref(p)), self._ctx)
        return _capi.params_dict(p)

    def set_params(self, **kw) -> None:
        cur = self.params()
        cur.update(kw)
        p = _capi.params_from(**cur)
        with torch.cuda.device(self.device):
            _capi.check(self._lib.pbf_set_params(self._ctx, C.byref(p)), self._ctx)

    def reserve(self, n: int, nscenes: int = 1) -> None:
        with torch.cuda.device(self.device):
            if nscenes == 1:
                _capi.check(self._lib.pbf_reserve(self._ctx, int(n)), self._ctx)
            else:
                _capi.check(self._lib.pbf_reserve_batch(self._ctx, int(n), int(nscenes)), self._ctx)

    def stats(self) -> dict:
        s = _capi.pbf_stats()
        _capi.check(self._lib.pbf_get_stats(self._ctx, C.byref(s)), self._ctx)
        out = {}
        for name, _ in _capi.pbf_stats._fields_:
            v = getattr(s, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out

    def set_profiling(self, enable: bool) -> None:
        _capi.check(self._lib.pbf_set_profiling(self._ctx, 1 if enable else 0), self._ctx)

    # -- launch settings and the autotuner (no reference counterpart) ----------
    def tuning(self) -> dict:
        t = _capi.pbf_tuning()
        _capi.check(self._lib.pbf_get_tuning(self._ctx, C.byref(t)), self._ctx)
        return {"pass_block": t.pass_block, "stage_min_particles": t.stage_min_particles,
                "persistent": t.persistent, "pass_lanes": t.pass_lanes}

    def set_tuning(self, pass_block: int = 256, stage_min_particles: int = 0, persistent: int = 0,
                   pass_lanes: int = 0) -> None:
        t = _capi.pbf_tuning(int(pass_block), int(stage_min_particles), int(persistent), int(pass_lanes))
        _capi.check(self._lib.pbf_set_tuning(self._ctx, C.byref(t)), self._ctx)

    def autotune(self, x: torch.Tensor, v: torch.Tensor, scene_offsets: Optional[torch.Tensor] = None,
                 steps: int = 3, cache: Optional[str] = None) -> dict:
        """pbf_autotune: times every launch setting on (x, v) (not modified),
        applies the fastest; with `cache` (a JSON file path) the choice is
        stored and reused for workloads of the same key."""
        x = _check_vec3(x, "position")
        v = _check_vec3(v, "velocity", x.shape[0], x.device)
        nscenes, off = 1, None
        if scene_offsets is not None:
            off = _check_scene_offsets(scene_offsets, x.shape[0], x.device, self.validate_offsets)
            nscenes = off.shape[0] - 1
        r = _capi.pbf_tune_report()
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_autotune(self._ctx, _ptr(x), _ptr(v), _ptr(off), nscenes, x.shape[0],
                                               int(steps), None if cache is None else str(cache).encode(),
                                               C.byref(r), _stream(x.device)), self._ctx)
        return {"pass_block": r.best.pass_block, "stage_min_particles": r.best.stage_min_particles,
                "persistent": r.best.persistent, "pass_lanes": r.best.pass_lanes, "best_ms": r.best_ms,
                "default_ms": r.default_ms, "candidates": r.candidates,
                "from_cache": bool(r.from_cache), "key": r.key.decode()}

    def profile(self) -> dict:
        """Per pass: (summed ms, timed launches) since profiling was enabled."""
        p = _capi.pbf_profile()
        _capi.check(self._lib.pbf_get_profile(self._ctx, C.byref(p)), self._ctx)
        return {name: (p.ms[i], p.launches[i]) for i, name in enumerate(_capi.PASS_NAMES)}

    @property
    def handle(self):
        return self._ctx

    def close(self) -> None:
        if self._ctx:
            self._lib.pbf_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- C ABI calls on tensors -------------------------------------------------
    def neighbor_query(self, x: torch.Tensor) -> "NeighborList":
        x = _check_vec3(x, "position")
        n = x.shape[0]
        off = torch.empty(max(n, 1), dtype=torch.int32, device=x.device)[:n]
        total = C.c_longlong()
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_neighbor_query(self._ctx, _ptr(x), n, _ptr(off), None, 0, C.byref(total),
                                                     _stream(x.device)), self._ctx)
            el = torch.empty(max(total.value, 1), dtype=torch.int32, device=x.device)
            _capi.check(self._lib.pbf_step_neighbors(self._ctx, _ptr(off), _ptr(el), total.value, C.byref(total),
                                                     _stream(x.device)), self._ctx)
        return NeighborList(off, el[: total.value])

    def step_neighbors(self, n: int) -> "NeighborList":
        off = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)[:n]
        total = C.c_longlong()
        with torch.cuda.device(self.device):
            _capi.check(self._lib.pbf_step_neighbors(self._ctx, _ptr(off), None, 0, C.byref(total),
                                                     _stream(self.device)), self._ctx)
            el = torch.empty(max(total.value, 1), dtype=torch.int32, device=self.device)
            _capi.check(self._lib.pbf_step_neighbors(self._ctx, _ptr(off), _ptr(el), total.value, C.byref(total),
                                                     _stream(self.device)), self._ctx)
        return NeighborList(off, el[: total.value])

    def neighbor_query_fixed(self, x: torch.Tensor, limit: int) -> "NeighborList":
        """pbf_neighbor_query_fixed: rows of at most `limit` nearest neighbors
        (the reference's limited NeighborList layout)."""
        x = _check_vec3(x, "position")
        n = x.shape[0]
        counts = torch.zeros(n, dtype=torch.int32, device=x.device)
        elems = torch.full((n * limit,), -1, dtype=torch.int32, device=x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_neighbor_query_fixed(self._ctx, _ptr(x), n, int(limit), _ptr(counts),
                                                           _ptr(elems), _stream(x.device)), self._ctx)
        return NeighborList(counts, elems, limit=int(limit))

    def integrate(self, x, v, f=None) -> None:
        x = _check_vec3(x, "position")
        v = _check_vec3(v, "velocity", x.shape[0], x.device)
        if f is not None:
            f = _check_vec3(f, "force", x.shape[0], x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_integrate(self._ctx, _ptr(x), _ptr(v), _ptr(f), x.shape[0], _stream(x.device)),
                        self._ctx)

    def density(self, x, nbr: "NeighborList", rho: torch.Tensor) -> None:
        x = _check_vec3(x, "position")
        n = x.shape[0]
        _check_nbr(nbr, n, x.device)
        if rho is None:
            raise ValueError("rho: expected an output tensor")
        rho = _check_scalar(rho, "rho", n, x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_density(self._ctx, _ptr(x), _ptr(nbr.index), _ptr(nbr.elements),
                                              nbr.elements.numel(), _ptr(rho), n, _stream(x.device)), self._ctx)

    def pbd_constrain_mass(self, x, v, nbr: "NeighborList", mass_inv: torch.Tensor,
                           rho: Optional[torch.Tensor]) -> None:
        """pbf_pbd_constrain_mass: DensityPBD::constrain with inverse masses."""
        x = _check_vec3(x, "position")
        n = x.shape[0]
        v = _check_vec3(v, "velocity", n, x.device)
        _check_nbr(nbr, n, x.device)
        rho = _check_scalar(rho, "rho", n, x.device)
        if mass_inv.dtype != torch.float32 or mass_inv.shape != (n,) or mass_inv.device != x.device:
            raise ValueError("mass_inv: expected an (n,) float32 tensor on the positions' device")
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_pbd_constrain_mass(self._ctx, _ptr(x), _ptr(v), _ptr(nbr.index),
                                                         _ptr(nbr.elements), nbr.elements.numel(),
                                                         _ptr(mass_inv.contiguous()), _ptr(rho), n,
                                                         _stream(x.device)), self._ctx)

    def pbd_constrain(self, x, v, nbr: "NeighborList", rho: Optional[torch.Tensor]) -> None:
        x = _check_vec3(x, "position")
        n = x.shape[0]
        v = _check_vec3(v, "velocity", n, x.device)
        _check_nbr(nbr, n, x.device)
        rho = _check_scalar(rho, "rho", n, x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_pbd_constrain(self._ctx, _ptr(x), _ptr(v), _ptr(nbr.index),
                                                    _ptr(nbr.elements), nbr.elements.numel(), _ptr(rho), n,
                                                    _stream(x.device)), self._ctx)

    def pbd_constrain_noalloc(self, x, v, nbr: "NeighborList", rho: Optional[torch.Tensor]) -> None:
        """pbf_pbd_constrain_noalloc: pbd_constrain with the library's fixed
        per-device scratch (at most 2^20 particles; calls on one device must
        not overlap)."""
        x = _check_vec3(x, "position")
        n = x.shape[0]
        v = _check_vec3(v, "velocity", n, x.device)
        _check_nbr(nbr, n, x.device)
        rho = _check_scalar(rho, "rho", n, x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_pbd_constrain_noalloc(self._ctx, _ptr(x), _ptr(v), _ptr(nbr.index),
                                                            _ptr(nbr.elements), nbr.elements.numel(), _ptr(rho),
                                                            n, _stream(x.device)), self._ctx)

    def step(self, x, v, rho: Optional[torch.Tensor] = None, nsteps: int = 1) -> None:
        x = _check_vec3(x, "position")
        v = _check_vec3(v, "velocity", x.shape[0], x.device)
        rho = _check_scalar(rho, "rho", x.shape[0], x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_step(self._ctx, _ptr(x), _ptr(v), _ptr(rho), x.shape[0], int(nsteps),
                                           _stream(x.device)), self._ctx)

    def batch_step(self, x, v, scene_offsets: torch.Tensor, rho: Optional[torch.Tensor] = None,
                   nsteps: int = 1) -> None:
        """nsteps of every scene of a batch; scene s holds rows
        scene_offsets[s]:scene_offsets[s+1] (int32 device tensor, nscenes+1)."""
        x = _check_vec3(x, "position")
        v = _check_vec3(v, "velocity", x.shape[0], x.device)
        rho = _check_scalar(rho, "rho", x.shape[0], x.device)
        off = _check_scene_offsets(scene_offsets, x.shape[0], x.device, self.validate_offsets)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_batch_step(self._ctx, _ptr(x), _ptr(v), _ptr(rho), _ptr(off),
                                                 off.shape[0] - 1, x.shape[0], int(nsteps), _stream(x.device)),
                        self._ctx)


    def viscosity(self, x, v, nbr: "NeighborList") -> None:
        """pbf_viscosity: ImplicitViscosity::constrain on v (in place)."""
        x = _check_vec3(x, "position")
        n = x.shape[0]
        v = _check_vec3(v, "velocity", n, x.device)
        _check_nbr(nbr, n, x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_viscosity(self._ctx, _ptr(x), _ptr(v), _ptr(nbr.index), _ptr(nbr.elements),
                                                nbr.elements.numel(), n, _stream(x.device)), self._ctx)

    def box_constrain(self, x, v) -> None:
        """pbf_box_constrain: the box boundary on x, v (in place)."""
        x = _check_vec3(x, "position")
        v = _check_vec3(v, "velocity", x.shape[0], x.device)
        with torch.cuda.device(x.device):
            _capi.check(self._lib.pbf_box_constrain(self._ctx, _ptr(x), _ptr(v), x.shape[0], _stream(x.device)),
                        self._ctx)

    # -- reverse mode (through the pybind11 layer, csrc/pbf_torch.cpp) --------
    def rollout_forward(self, x, v, rho: Optional[torch.Tensor] = None, nsteps: int = 1,
                        checkpoint_every: int = 10, scene_offsets: Optional[torch.Tensor] = None) -> int:
        """pbf_rollout_forward (float32 state) or pbf_rollout_forward_f64
        (float64 state): nsteps in place, checkpoints kept in the context.
        Returns a token that rollout_backward checks."""
        f64 = isinstance(x, torch.Tensor) and x.dtype == torch.float64
        dt = torch.float64 if f64 else torch.float32
        x = _check_vec3(x, "position", dtype=dt)
        v = _check_vec3(v, "velocity", x.shape[0], x.device, dtype=dt)
        rho = _check_scalar(rho, "rho", x.shape[0], x.device, dtype=dt)
        off = None
        if scene_offsets is not None:
            off = _check_scene_offsets(scene_offsets, x.shape[0], x.device, self.validate_offsets)
        with torch.cuda.device(x.device):
            return _capi.native().rollout_forward(int(self._ctx.value or 0), x, v, rho, int(nsteps),
                                                  int(checkpoint_every), off, _stream_int(x.device))

    def rollout_backward(self, xbar, vbar, token: Optional[int] = None) -> None:
        """pbf_rollout_backward (float32) or pbf_rollout_backward_f64 (float64):
        (dL/dx_N, dL/dv_N) -> (dL/dx_0, dL/dv_0) in place. A token from a
        rollout that is no longer the context's latest of its precision is
        refused (RuntimeError)."""
        dt = torch.float64 if isinstance(xbar, torch.Tensor) and xbar.dtype == torch.float64 else torch.float32
        xbar = _check_vec3(xbar, "xbar", dtype=dt)
        vbar = _check_vec3(vbar, "vbar", xbar.shape[0], xbar.device, dtype=dt)
        with torch.cuda.device(xbar.device):
            _capi.native().rollout_backward(int(self._ctx.value or 0), xbar, vbar, int(token or 0),
                                            _stream_int(xbar.device))


class NeighborList:
    """Neighbor lists (NeighborList.h:9-125). Dynamic (CSR): ``index`` holds
    the n exclusive row starts (the reference's m_index) and ``elements`` the
    neighbor ids; row i ends at index[i+1], the last row at elements.numel()
    (getNeighborSize). Limited (``limit`` > 0, isLimited()): ``index`` holds
    the n row sizes and row i sits at elements[limit*i:]."""

    def __init__(self, index: torch.Tensor, elements: torch.Tensor, limit: int = 0):
        self.index = index
        self.elements = elements
        self.limit = limit

    def isLimited(self) -> bool:
        return self.limit > 0

    def getNeighborLimit(self) -> int:
        return self.limit

    def size(self) -> int:
        return self.index.numel()

    def getNeighborSize(self, i: int) -> int:
        if self.limit:
            return int(self.index[i])
        end = self.elements.numel() if i + 1 == self.index.numel() else int(self.index[i + 1])
        return end - int(self.index[i])

    def offsets(self) -> torch.Tensor:
        """The n + 1 row offsets (row starts plus the element count)."""
        tail = torch.tensor([self.elements.numel()], dtype=torch.int32, device=self.index.device)
        return torch.cat([self.index, tail])

    def getElement(self, i: int, j: int) -> int:
        if self.limit:
            return int(self.elements[self.limit * i + j])
        return int(self.elements[int(self.index[i]) + j])

    def rows(self):
        """Host copy as a list of numpy arrays."""
        idx = self.index.cpu().numpy()
        el = self.elements.cpu().numpy()
        if self.limit:
            return [el[self.limit * i:self.limit * i + idx[i]] for i in range(len(idx))]
        ends = np.append(idx[1:], len(el))
        return [el[idx[i]:ends[i]] for i in range(len(idx))]


class VarField:
    """A scalar parameter field (Framework/Framework/FieldVar.h): setValue
    forwards to the context, as the reference's callbacks recompute m_factor
    (SummationDensity.cu:47-51)."""

    def __init__(self, ctx: Context, name: str, cast: Callable = float):
        self._ctx, self._name, self._cast = ctx, name, cast

    def setValue(self, v) -> None:
        self._ctx.set_params(**{self._name: self._cast(v)})

    def getValue(self):
        return self._ctx.params()[self._name]


class ArrayField:
    """A device-array field; connect() makes the sink share the source's tensor
    (FieldArray.h:166-185)."""

    def __init__(self, value: Optional[torch.Tensor] = None):
        self._src: Optional[ArrayField] = None
        self._value = value

    def _root(self) -> "ArrayField":
        f = self
        while f._src is not None:
            f = f._src
        return f

    def setValue(self, t: Optional[torch.Tensor]) -> None:
        self._root()._value = t

    def getValue(self) -> Optional[torch.Tensor]:
        return self._root()._value

    def connect(self, sink: "ArrayField") -> None:
        sink._src = self

    def isEmpty(self) -> bool:
        v = self.getValue()
        return v is None or v.numel() == 0

    def getElementCount(self) -> int:
        v = self.getValue()
        return 0 if v is None else v.shape[0]


class NeighborQuery:
    """Uniform-grid neighbor search (NeighborQuery.cu:152-199)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.inRadius = VarField(ctx, "smoothing_length")
        self.inPosition = ArrayField()
        self.outNeighborhood = ArrayField()

        self.m_maxNum = 0

    def setNeighborSizeLimit(self, num: int) -> None:
        """NeighborQuery::setNeighborSizeLimit (NeighborQuery.h:34): rows keep
        at most `num` nearest neighbors."""
        self.m_maxNum = int(num)

    def compute(self) -> None:
        x = self.inPosition.getValue()
        if self.m_maxNum > 0:
            self.outNeighborhood.setValue(self.ctx.neighbor_query_fixed(x, self.m_maxNum))
        else:
            self.outNeighborhood.setValue(self.ctx.neighbor_query(x))


class ParticleIntegrator:
    """Symplectic Euler (ParticleIntegrator.cu:20-138)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.inPosition = ArrayField()
        self.inVelocity = ArrayField()
        self.inForceDensity = ArrayField()

    def begin(self) -> None:
        f = self.inForceDensity.getValue()
        if f is not None:
            f.zero_()

    def integrate(self) -> bool:
        if not self.inPosition.isEmpty():
            self.ctx.integrate(self.inPosition.getValue(), self.inVelocity.getValue(), self.inForceDensity.getValue())
        return True

    def end(self) -> None:
        pass


class DensityPBD:
    """Position-based incompressibility (DensityPBD.cu:271-302)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.varIterationNumber = VarField(ctx, "iterations", int)
        self.varRestDensity = VarField(ctx, "rest_density")
        self.varSamplingDistance = VarField(ctx, "sampling_distance")
        self.varSmoothingLength = VarField(ctx, "smoothing_length")
        self.inPosition = ArrayField()
        self.inVelocity = ArrayField()
        self.inNeighborIndex = ArrayField()
        self.outDensity = ArrayField()
        self.m_massInv = ArrayField()  # optional inverse masses (DensityPBD.h:77)

    def constrain(self) -> bool:
        x = self.inPosition.getValue()
        rho = self.outDensity.getValue()
        if rho is None or rho.shape[0] != x.shape[0]:
            rho = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
            self.outDensity.setValue(rho)
        if self.m_massInv.isEmpty():
            self.ctx.pbd_constrain(x, self.inVelocity.getValue(), self.inNeighborIndex.getValue(), rho)
        else:
            self.ctx.pbd_constrain_mass(x, self.inVelocity.getValue(), self.inNeighborIndex.getValue(),
                                        self.m_massInv.getValue(), rho)
        return True


class ImplicitViscosity:
    """XSPH-like implicit viscosity (ImplicitViscosity.h:10-44,
    ImplicitViscosity.cu:71-128): m_viscosity, m_smoothingLength, m_position,
    m_velocity, m_neighborhood; constrain()."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.m_viscosity = VarField(ctx, "viscosity")
        self.m_smoothingLength = VarField(ctx, "smoothing_length")
        self.m_position = ArrayField()
        self.m_velocity = ArrayField()
        self.m_neighborhood = ArrayField()

    def setViscosity(self, mu: float) -> None:
        self.m_viscosity.setValue(mu)

    def setIterationNumber(self, n: int) -> None:
        self.ctx.set_params(viscosity_iterations=int(n))

    def constrain(self) -> bool:
        if not self.m_position.isEmpty():
            self.ctx.viscosity(self.m_position.getValue(), self.m_velocity.getValue(), self.m_neighborhood.getValue())
        return True


class BoundaryConstraint:
    """Particles kept inside a box (BoundaryConstraint.h:12-40): setCube with
    inverted=True (fluid inside), m_normal_friction, m_tangent_friction,
    constrain()."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.m_position = ArrayField()
        self.m_velocity = ArrayField()
        self.m_normal_friction = VarField(ctx, "normal_friction")
        self.m_tangent_friction = VarField(ctx, "tangent_friction")

    def setCube(self, lo, hi, distance: float = 0.005, inverted: bool = True) -> None:
        if not inverted:
            raise NotImplementedError("only the container box (particles inside) is supported")
        self.ctx.set_params(box_lo=list(lo), box_hi=list(hi), box_spacing=float(distance))

    def constrain(self, position=None, velocity=None) -> bool:
        x = self.m_position.getValue() if position is None else position
        v = self.m_velocity.getValue() if velocity is None else velocity
        if x is not None and x.shape[0]:
            self.ctx.box_constrain(x, v)
        return True


class StaticBoundary:
    """Root node holding box obstacles (StaticBoundary.cpp:45-77): advance(dt)
    constrains every registered particle system before it steps."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.m_obstacles = []
        self.m_particleSystems = []

    def loadCube(self, lo, hi, distance: float = 0.005, bOutBoundary: bool = True) -> BoundaryConstraint:
        b = BoundaryConstraint(self.ctx)
        b.setCube(lo, hi, distance, bOutBoundary)
        self.m_obstacles.append(b)
        return b

    def addParticleSystem(self, model: "PositionBasedFluidModel") -> None:
        self.m_particleSystems.append(model)

    def advance(self, dt: Optional[float] = None) -> None:
        for b in self.m_obstacles:
            for m in self.m_particleSystems:
                b.constrain(m.m_position.getValue(), m.m_velocity.getValue())


class PositionBasedFluidModel:
    """NumericalModel of the PBD fluid (PositionBasedFluidModel.cpp:41-101).

    As in the reference, the model runs ImplicitViscosity (viscosity 1) after
    the PBD solve; viscosity=False gives the plain PBD step. box=(lo, hi,
    spacing) adds the container box of a StaticBoundary root node, applied
    before every step.

    step(dt) runs the fused MI355X step (one library call);
    step_modules(dt) runs the module sequence begin -> NeighborQuery::compute ->
    integrate -> DensityPBD::constrain -> ImplicitViscosity::constrain -> end
    through the module-granular entry points, as the reference's step does
    (the box, when set, first)."""

    def __init__(self, smoothing_length: float = 0.0125, device="cuda:0", viscosity: bool = True, box=None,
                 **params):
        flags = params.pop("flags", 0)
        if viscosity:
            flags |= _capi.PBF_FLAG_VISCOSITY
        if box is not None:
            flags |= _capi.PBF_FLAG_BOX
            lo, hi, spacing = box
            params.update(box_lo=list(lo), box_hi=list(hi), box_spacing=float(spacing))
        self.ctx = Context(device, smoothing_length=smoothing_length, flags=flags, **params)
        self.m_smoothingLength = VarField(self.ctx, "smoothing_length")
        self.m_position = ArrayField()
        self.m_velocity = ArrayField()
        self.m_forceDensity = ArrayField()
        self.m_density = ArrayField()
        self.nbrQuery = NeighborQuery(self.ctx)
        self.pbdModule = DensityPBD(self.ctx)
        self.integrator = ParticleIntegrator(self.ctx)
        self.visModule = ImplicitViscosity(self.ctx) if viscosity else None
        self.boundary = None
        if box is not None:
            self.boundary = BoundaryConstraint(self.ctx)
            self.m_position.connect(self.boundary.m_position)
            self.m_velocity.connect(self.boundary.m_velocity)
        self.m_position.connect(self.nbrQuery.inPosition)
        self.m_position.connect(self.pbdModule.inPosition)
        self.m_velocity.connect(self.pbdModule.inVelocity)
        self.m_position.connect(self.integrator.inPosition)
        self.m_velocity.connect(self.integrator.inVelocity)
        self.m_forceDensity.connect(self.integrator.inForceDensity)
        self.nbrQuery.outNeighborhood.connect(self.pbdModule.inNeighborIndex)
        self.pbdModule.outDensity.connect(self.m_density)
        if self.visModule is not None:
            self.m_position.connect(self.visModule.m_position)
            self.m_velocity.connect(self.visModule.m_velocity)
            self.nbrQuery.outNeighborhood.connect(self.visModule.m_neighborhood)

    def _set_dt(self, dt: Optional[float]) -> None:
        if dt is not None and abs(self.ctx.params()["dt"] - float(np.float32(dt))) > 0:
            self.ctx.set_params(dt=float(dt))

    def step(self, dt: Optional[float] = None, nsteps: int = 1) -> None:
        self._set_dt(dt)
        x = self.m_position.getValue()
        rho = self.m_density.getValue()
        if rho is None or rho.shape[0] != x.shape[0]:
            rho = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
            self.m_density.setValue(rho)
        self.ctx.step(x, self.m_velocity.getValue(), rho, nsteps)

    def step_modules(self, dt: Optional[float] = None) -> None:
        self._set_dt(dt)
        if self.boundary is not None:
            self.boundary.constrain()
        self.integrator.begin()
        self.nbrQuery.compute()
        self.integrator.integrate()
        self.pbdModule.constrain()
        if self.visModule is not None:
            self.visModule.constrain()
        self.integrator.end()

    def step_arrays(self, x: np.ndarray, v: np.ndarray, nsteps: int = 1, modules: bool = False):
        """Convenience for numpy callers: copies in, steps, copies out
        (positions, velocities, densities)."""
        dev = self.ctx.device
        self.m_position.setValue(torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev))
        self.m_velocity.setValue(torch.from_numpy(np.ascontiguousarray(v, np.float32)).to(dev))
        self.m_forceDensity.setValue(torch.zeros_like(self.m_position.getValue()))
        if modules:
            for _ in range(nsteps):
                self.step_modules()
        else:
            self.step(nsteps=nsteps)
        torch.cuda.synchronize(dev)
        return (self.m_position.getValue().cpu().numpy(), self.m_velocity.getValue().cpu().numpy(),
                self.m_density.getValue().cpu().numpy())


class BatchedPositionBasedFluidModel(PositionBasedFluidModel):
    """Many independent PositionBasedFluidModel scenes with shared parameters,
    stepped together: the scenes' particles are concatenated in m_position /
    m_velocity / m_density and m_sceneOffsets (int32, nscenes+1) delimits them.
    Each scene's result equals a PositionBasedFluidModel step of that scene
    alone (the reference runs one model per scene)."""

    def __init__(self, smoothing_length: float = 0.0125, device="cuda:0", **params):
        super().__init__(smoothing_length, device, **params)
        self.m_sceneOffsets = ArrayField()

    def setScenes(self, positions, velocities=None) -> None:
        """Concatenates per-scene (n_s, 3) arrays or tensors."""
        dev = self.ctx.device
        xs = [torch.as_tensor(np.asarray(p, np.float32) if not torch.is_tensor(p) else p, dtype=torch.float32)
              for p in positions]
        if velocities is None:
            vs = [torch.zeros_like(t) for t in xs]
        else:
            vs = [torch.as_tensor(np.asarray(q, np.float32) if not torch.is_tensor(q) else q, dtype=torch.float32)
                  for q in velocities]
        counts = [int(t.shape[0]) for t in xs]
        off = np.zeros(len(counts) + 1, np.int64)
        off[1:] = np.cumsum(counts)
        self.m_position.setValue(torch.cat([t.reshape(-1, 3) for t in xs]).to(dev).contiguous())
        self.m_velocity.setValue(torch.cat([t.reshape(-1, 3) for t in vs]).to(dev).contiguous())
        self.m_sceneOffsets.setValue(torch.from_numpy(off.astype(np.int32)).to(dev))
        self.ctx.reserve(int(off[-1]), len(counts))

    def scene(self, s: int):
        off = self.m_sceneOffsets.getValue().cpu().numpy()
        sl = slice(int(off[s]), int(off[s + 1]))
        rho = self.m_density.getValue()
        return (self.m_position.getValue()[sl], self.m_velocity.getValue()[sl],
                None if rho is None else rho[sl])

    def step(self, dt: Optional[float] = None, nsteps: int = 1) -> None:
        self._set_dt(dt)
        x = self.m_position.getValue()
        rho = self.m_density.getValue()
        if rho is None or rho.shape[0] != x.shape[0]:
            rho = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
            self.m_density.setValue(rho)
        self.ctx.batch_step(x, self.m_velocity.getValue(), self.m_sceneOffsets.getValue(), rho, nsteps)


class ParticleFluid:
    """ParticleFluid node (ParticleFluid.cpp:14-129): a PositionBasedFluidModel
    whose particles may come from particle emitters (emitters.py).

    currentPosition / currentVelocity / currentForce are the model's
    m_position / m_velocity / m_forceDensity. advance(dt) follows
    ParticleFluid::advance: with emitters, the fluid's state is copied back
    into each emitter's segment (emitters own consecutive segments, in the
    order added), every emitter appends its new particles (advance2), the
    fluid's arrays become the emitters' arrays concatenated, and the model
    steps when there is at least one particle and self_update is set. Without
    emitters the fluid steps the particles it holds. The particle count can
    change every frame; the library grows its buffers as needed."""

    def __init__(self, name: str = "fluid", model: Optional[PositionBasedFluidModel] = None, **model_args):
        self.name = name
        self.model = model if model is not None else PositionBasedFluidModel(**model_args)
        self.currentPosition = self.model.m_position
        self.currentVelocity = self.model.m_velocity
        self.currentForce = self.model.m_forceDensity
        self.self_update = True
        self.dt = self.model.ctx.params()["dt"]
        self._emitters = []

    def addParticleEmitter(self, emitter) -> None:
        self._emitters.append(emitter)

    def getParticleEmitters(self):
        return list(self._emitters)

    def setDt(self, dt: float) -> None:
        self.dt = float(dt)

    def getDt(self) -> float:
        return self.dt

    def resetStatus(self) -> bool:
        """With emitters the fluid starts empty (ParticleFluid.cpp:132-145)."""
        if self._emitters:
            dev = self.model.ctx.device
            for f in (self.currentPosition, self.currentVelocity, self.currentForce):
                f.setValue(torch.empty((0, 3), dtype=torch.float32, device=dev))
        return True

    def _copy_out(self) -> None:
        x, v, f = self.currentPosition.getValue(), self.currentVelocity.getValue(), self.currentForce.getValue()
        start = 0
        for e in self._emitters:
            num = e.currentPosition.getElementCount()
            if num == 0:
                continue
            if start + num > x.shape[0]:
                raise ValueError(f"emitter {e.name!r} owns {num} particles past the fluid's {x.shape[0]}")
            e.currentPosition.setValue(x[start:start + num].clone())
            e.currentVelocity.setValue(v[start:start + num].clone())
            e.currentForce.setValue(f[start:start + num].clone() if f is not None and f.shape[0] >= start + num
                                    else torch.zeros_like(x[start:start + num]))
            start += num

    def advance(self, dt: Optional[float] = None) -> int:
        """One frame; returns the particle count stepped (0 when nothing ran)."""
        if dt is not None:
            self.dt = float(dt)
        if self._emitters:
            if self.currentPosition.getElementCount() > 0:
                self._copy_out()
            for e in self._emitters:
                e.advance2(self.dt)
            parts = [e for e in self._emitters if e.currentPosition.getElementCount() > 0]
            total = sum(e.currentPosition.getElementCount() for e in parts)
            if total > 0:
                self.currentPosition.setValue(torch.cat([e.currentPosition.getValue() for e in parts]).contiguous())
                self.currentVelocity.setValue(torch.cat([e.currentVelocity.getValue() for e in parts]).contiguous())
                self.currentForce.setValue(torch.cat([e.currentForce.getValue() for e in parts]).contiguous())
        else:
            total = self.currentPosition.getElementCount()
        if total > 0 and self.self_update:
            f = self.currentForce.getValue()
            if f is not None:
                f.zero_()  # ParticleIntegrator::begin resets the force density
            self.model.step(self.dt)
            return total
        return 0


def rollout(ctx: Context, x0: torch.Tensor, v0: torch.Tensor, nsteps: int, checkpoint_every: int = 10,
            scene_offsets: Optional[torch.Tensor] = None):
    """Differentiable nsteps of the fluid step (one scene, or a batch delimited
    by scene_offsets): returns (x_N, v_N, rho_N); gradients flow to x0 and v0
    through the library's adjoint kernels (rho_N carries none). float32 x0, v0
    run the single-precision kernels, float64 the double-precision path
    (pbf_rollout_forward_f64). The context keeps one rollout per precision:
    run backward before the next rollout of that precision on it."""
    off = None
    if scene_offsets is not None:
        off = _check_scene_offsets(scene_offsets, x0.shape[0], x0.device, ctx.validate_offsets)
    with torch.cuda.device(x0.device):
        out = tuple(_capi.native().rollout(x0, v0, int(ctx.handle.value or 0), int(nsteps),
                                           int(checkpoint_every), off, _stream_int(x0.device)))
    if out[0].grad_fn is not None:
        # the backward node holds the context's handle only: keep the context
        # (and its checkpoints) alive as long as the node
        out[0].grad_fn.metadata["pbf_context"] = ctx
    return out
