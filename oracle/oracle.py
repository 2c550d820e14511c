"""ctypes front end of the CPU oracle (oracle/_build/liborc.so).

Test infrastructure: the checker. Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module; the product never does.
Arrays are numpy, positions and velocities (n, 3) packed xyz.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
# PBF_ORACLE_LIB selects another build of the same sources (the sanitizer
# build, _build/liborc_san.so)
_LIB_PATH = Path(os.environ.get("PBF_ORACLE_LIB", _DIR / "_build" / "liborc.so"))


def build(force: bool = False) -> Path:
    if "PBF_ORACLE_LIB" in os.environ:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"PBF_ORACLE_LIB={_LIB_PATH} does not exist")
        return _LIB_PATH
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-C", str(_DIR)] + (["-B"] if force else []), check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def _params_struct(real):
    class P(C.Structure):
        _fields_ = [("h", real), ("d", real), ("rho0", real), ("dt", real), ("g", real * 3),
                    ("iters", C.c_int), ("lam_rest", real), ("lam_eps", real), ("corr", real),
                    ("pair_eps", real), ("lo", real * 3), ("hi", real * 3), ("dx_mode", C.c_int),
                    ("mu", real), ("vis_iters", C.c_int), ("box_lo", real * 3), ("box_hi", real * 3),
                    ("box_spacing", real), ("nfric", real), ("tfric", real), ("flags", C.c_int)]
    return P


ParamsF = _params_struct(C.c_float)
ParamsD = _params_struct(C.c_double)

DEFAULTS = dict(h=0.0125, d=0.005, rho0=1000.0, dt=0.001, g=(0.0, -9.8, 0.0), iters=5,
                lam_rest=1000.0, lam_eps=0.1, corr=10.0, pair_eps=1e-6,
                lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), dx_mode=0,
                mu=1.0, vis_iters=5, box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0), box_spacing=0.005,
                nfric=0.95, tfric=0.0, flags=0)

FLAG_VISCOSITY = 1
FLAG_BOX = 2


def make_params(double: bool = False, **kw):
    v = dict(DEFAULTS)
    v.update(kw)
    P = ParamsD if double else ParamsF
    p = P()
    for k, val in v.items():
        if k in ("g", "lo", "hi", "box_lo", "box_hi"):
            arr = getattr(p, k)
            for i in range(3):
                arr[i] = val[i]
        else:
            setattr(p, k, val)
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(str(_LIB_PATH))
        fp, dp, ip, llp = (C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int),
                           C.POINTER(C.c_longlong))
        PF, PD = C.POINTER(ParamsF), C.POINTER(ParamsD)
        sig = {
            "orc_abi_version": (C.c_int, []),
            "orc_set_threads": (C.c_int, [C.c_int]),
            "orc_grid_f": (None, [PF, fp, fp, ip]),
            "orc_cell3_f": (None, [PF, fp, C.c_int, ip, llp]),
            "orc_spiky_w_f": (C.c_float, [C.c_float, C.c_float]),
            "orc_spiky_grad_f": (C.c_float, [C.c_float, C.c_float]),
            "orc_spiky_w_d": (C.c_double, [C.c_double, C.c_double]),
            "orc_spiky_grad_d": (C.c_double, [C.c_double, C.c_double]),
            "orc_scaling_factor_f": (C.c_float, [C.c_float, C.c_float]),
            "orc_scaling_factor_d": (C.c_double, [C.c_double, C.c_double]),
            "orc_mass_f": (C.c_float, [PF]),
            "orc_neighbor_query_f": (C.c_longlong, [PF, fp, C.c_int, ip, ip, C.c_longlong]),
            "orc_neighbor_query_d": (C.c_longlong, [PD, dp, C.c_int, ip, ip, C.c_longlong]),
            "orc_density_f": (None, [PF, fp, C.c_int, ip, ip, fp]),
            "orc_lambdas_f": (None, [PF, fp, fp, C.c_int, ip, ip, fp]),
            "orc_displacement_f": (None, [PF, fp, fp, C.c_int, ip, ip, fp, C.c_int]),
            "orc_pbd_constrain_f": (None, [PF, fp, fp, C.c_int, ip, ip, fp]),
            "orc_pbd_constrain_d": (None, [PD, dp, dp, C.c_int, ip, ip, dp]),
            "orc_integrate_f": (None, [PF, fp, fp, fp, C.c_int]),
            "orc_integrate_d": (None, [PD, dp, dp, dp, C.c_int]),
            "orc_step_f": (C.c_longlong, [PF, fp, fp, fp, C.c_int, C.c_int]),
            "orc_step_d": (C.c_longlong, [PD, dp, dp, dp, C.c_int, C.c_int]),
            "orc_pbd_constrain_minv_f": (None, [PF, fp, fp, C.c_int, ip, ip, fp, fp]),
            "orc_pbd_constrain_minv_d": (None, [PD, dp, dp, C.c_int, ip, ip, dp, dp]),
            "orc_neighbor_query_fixed_f": (None, [PF, fp, C.c_int, C.c_int, ip, ip]),
            "orc_viscosity_f": (None, [PF, fp, fp, C.c_int, ip, ip]),
            "orc_vis_weight_f": (C.c_float, [C.c_float, C.c_float]),
            "orc_box_sdf_f": (C.c_longlong, [fp, fp, C.c_float, C.c_int, ip, fp, fp, fp, C.c_longlong]),
            "orc_sdf_distance_f": (None, [fp, fp, C.c_float, fp, C.c_int, fp, fp]),
            "orc_box_constrain_f": (None, [PF, fp, fp, C.c_int]),
            "orc_box_jitter": (C.c_float, [C.c_int]),
            "orc_rollout_adjoint_f": (None, [PF, fp, fp, C.c_int, C.c_int, fp, fp]),
            "orc_rollout_adjoint_d": (None, [PD, dp, dp, C.c_int, C.c_int, dp, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _ptr(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def set_threads(t: int) -> int:
    return lib().orc_set_threads(int(t))


def grid(params):
    ds = C.c_float()
    lo = (C.c_float * 3)()
    dims = (C.c_int * 3)()
    lib().orc_grid_f(C.byref(params), C.byref(ds), lo, dims)
    return ds.value, np.array(lo[:], dtype=np.float32), tuple(dims[:])


def cell3(params, x):
    x = _f32(x)
    n = len(x)
    c3 = np.zeros((n, 3), dtype=np.int32)
    cid = np.zeros(n, dtype=np.int64)
    lib().orc_cell3_f(C.byref(params), _ptr(x, C.c_float), n, _ptr(c3, C.c_int), _ptr(cid, C.c_longlong))
    return c3, cid


def neighbor_query(params, x):
    """CSR (offsets[n+1], elems) of NeighborQuery::compute on x."""
    dbl = isinstance(params, ParamsD)
    x = np.ascontiguousarray(x, dtype=np.float64 if dbl else np.float32)
    n = len(x)
    off = np.zeros(n + 1, dtype=np.int32)
    fn = lib().orc_neighbor_query_d if dbl else lib().orc_neighbor_query_f
    ct = C.c_double if dbl else C.c_float
    total = fn(C.byref(params), _ptr(x, ct), n, _ptr(off, C.c_int), None, 0)
    el = np.zeros(max(total, 1), dtype=np.int32)
    fn(C.byref(params), _ptr(x, ct), n, _ptr(off, C.c_int), _ptr(el, C.c_int), total)
    return off, el[:total]


def mass(params) -> float:
    return lib().orc_mass_f(C.byref(params))


def density(params, x, off, el):
    x = _f32(x)
    rho = np.zeros(len(x), dtype=np.float32)
    lib().orc_density_f(C.byref(params), _ptr(x, C.c_float), len(x), _ptr(off, C.c_int), _ptr(el, C.c_int),
                        _ptr(rho, C.c_float))
    return rho


def lambdas(params, rho, x, off, el):
    x = _f32(x)
    rho = _f32(rho)
    lam = np.zeros(len(x), dtype=np.float32)
    lib().orc_lambdas_f(C.byref(params), _ptr(rho, C.c_float), _ptr(x, C.c_float), len(x), _ptr(off, C.c_int),
                        _ptr(el, C.c_int), _ptr(lam, C.c_float))
    return lam


def displacement(params, lam, x, off, el, mode=1):
    x = _f32(x)
    lam = _f32(lam)
    dpos = np.zeros_like(x)
    lib().orc_displacement_f(C.byref(params), _ptr(lam, C.c_float), _ptr(x, C.c_float), len(x),
                             _ptr(off, C.c_int), _ptr(el, C.c_int), _ptr(dpos, C.c_float), int(mode))
    return dpos


def pbd_constrain(params, x, v, off, el):
    dbl = isinstance(params, ParamsD)
    dt = np.float64 if dbl else np.float32
    ct = C.c_double if dbl else C.c_float
    x = np.array(x, dtype=dt, copy=True, order="C")
    v = np.array(v, dtype=dt, copy=True, order="C")
    rho = np.zeros(len(x), dtype=dt)
    fn = lib().orc_pbd_constrain_d if dbl else lib().orc_pbd_constrain_f
    fn(C.byref(params), _ptr(x, ct), _ptr(v, ct), len(x), _ptr(np.ascontiguousarray(off, np.int32), C.c_int),
       _ptr(np.ascontiguousarray(el, np.int32), C.c_int), _ptr(rho, ct))
    return x, v, rho


def integrate(params, x, v, f=None):
    dbl = isinstance(params, ParamsD)
    dt = np.float64 if dbl else np.float32
    ct = C.c_double if dbl else C.c_float
    x = np.array(x, dtype=dt, copy=True, order="C")
    v = np.array(v, dtype=dt, copy=True, order="C")
    f = None if f is None else np.ascontiguousarray(f, dtype=dt)
    fn = lib().orc_integrate_d if dbl else lib().orc_integrate_f
    fn(C.byref(params), _ptr(x, ct), _ptr(v, ct), _ptr(f, ct), len(x))
    return x, v


def step(params, x, v, nsteps=1):
    """nsteps of the fluid step; returns (x, v, rho_of_last_step, pairs_of_last_step)."""
    dbl = isinstance(params, ParamsD)
    dt = np.float64 if dbl else np.float32
    ct = C.c_double if dbl else C.c_float
    x = np.array(x, dtype=dt, copy=True, order="C")
    v = np.array(v, dtype=dt, copy=True, order="C")
    rho = np.zeros(len(x), dtype=dt)
    fn = lib().orc_step_d if dbl else lib().orc_step_f
    pairs = fn(C.byref(params), _ptr(x, ct), _ptr(v, ct), _ptr(rho, ct), len(x), int(nsteps))
    return x, v, rho, pairs


def rollout_adjoint(params, x0, v0, nsteps, xbar, vbar):
    """Adjoint of nsteps steps from (x0, v0) with the neighbor sets held
    fixed: given the adjoint (xbar, vbar) of the final state, returns the
    adjoint of (x0, v0)."""
    dbl = isinstance(params, ParamsD)
    dt = np.float64 if dbl else np.float32
    ct = C.c_double if dbl else C.c_float
    x0 = np.array(x0, dtype=dt, copy=True, order="C")
    v0 = np.array(v0, dtype=dt, copy=True, order="C")
    xb = np.array(xbar, dtype=dt, copy=True, order="C")
    vb = np.array(vbar, dtype=dt, copy=True, order="C")
    fn = lib().orc_rollout_adjoint_d if dbl else lib().orc_rollout_adjoint_f
    fn(C.byref(params), _ptr(x0, ct), _ptr(v0, ct), len(x0), int(nsteps), _ptr(xb, ct), _ptr(vb, ct))
    return xb, vb


def viscosity(params, x, v, off, el):
    """ImplicitViscosity::constrain on the list (off, el); returns the new v."""
    x = _f32(x)
    v = np.array(v, dtype=np.float32, copy=True, order="C")
    lib().orc_viscosity_f(C.byref(params), _ptr(x, C.c_float), _ptr(v, C.c_float), len(x),
                          _ptr(np.ascontiguousarray(off, np.int32), C.c_int),
                          _ptr(np.ascontiguousarray(el, np.int32), C.c_int))
    return v


def vis_weight(r, h):
    return lib().orc_vis_weight_f(float(r), float(h))


def box_sdf(lo, hi, spacing, inverted=True):
    """(dims, left, h, samples) of the sampled box field of setCube."""
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    dims = np.zeros(3, np.int32)
    left = np.zeros(3, np.float32)
    h = np.zeros(3, np.float32)
    total = lib().orc_box_sdf_f(_ptr(lo, C.c_float), _ptr(hi, C.c_float), float(spacing), int(inverted),
                                _ptr(dims, C.c_int), _ptr(left, C.c_float), _ptr(h, C.c_float), None, 0)
    samples = np.zeros(total, np.float32)
    lib().orc_box_sdf_f(_ptr(lo, C.c_float), _ptr(hi, C.c_float), float(spacing), int(inverted),
                        _ptr(dims, C.c_int), _ptr(left, C.c_float), _ptr(h, C.c_float),
                        _ptr(samples, C.c_float), total)
    return dims, left, h, samples.reshape(dims[2], dims[1], dims[0])


def sdf_distance(lo, hi, spacing, x):
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    x = _f32(x)
    d = np.zeros(len(x), np.float32)
    nrm = np.zeros((len(x), 3), np.float32)
    lib().orc_sdf_distance_f(_ptr(lo, C.c_float), _ptr(hi, C.c_float), float(spacing), _ptr(x, C.c_float), len(x),
                             _ptr(d, C.c_float), _ptr(nrm, C.c_float))
    return d, nrm


def box_constrain(params, x, v):
    x = np.array(x, dtype=np.float32, copy=True, order="C")
    v = np.array(v, dtype=np.float32, copy=True, order="C")
    lib().orc_box_constrain_f(C.byref(params), _ptr(x, C.c_float), _ptr(v, C.c_float), len(x))
    return x, v


def box_jitter(i):
    return lib().orc_box_jitter(int(i))


def pbd_constrain_minv(params, x, v, off, el, minv):
    """DensityPBD::constrain with inverse masses (the massInv variant);
    returns (x, v, rho)."""
    dbl = isinstance(params, ParamsD)
    dt = np.float64 if dbl else np.float32
    ct = C.c_double if dbl else C.c_float
    x = np.array(x, dtype=dt, copy=True, order="C")
    v = np.array(v, dtype=dt, copy=True, order="C")
    minv = np.ascontiguousarray(minv, dtype=dt)
    rho = np.zeros(len(x), dtype=dt)
    fn = lib().orc_pbd_constrain_minv_d if dbl else lib().orc_pbd_constrain_minv_f
    fn(C.byref(params), _ptr(x, ct), _ptr(v, ct), len(x), _ptr(np.ascontiguousarray(off, np.int32), C.c_int),
       _ptr(np.ascontiguousarray(el, np.int32), C.c_int), _ptr(minv, ct), _ptr(rho, ct))
    return x, v, rho


def neighbor_query_fixed(params, x, limit):
    """Fixed-capacity neighbor rows: (counts (n,), elems (n, limit))."""
    x = _f32(x)
    counts = np.zeros(len(x), np.int32)
    elems = np.full((len(x), limit), -1, np.int32)
    lib().orc_neighbor_query_fixed_f(C.byref(params), _ptr(x, C.c_float), len(x), int(limit),
                                     _ptr(counts, C.c_int), _ptr(elems, C.c_int))
    return counts, elems
