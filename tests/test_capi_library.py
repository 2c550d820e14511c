"""libpbf.so loads without a GPU and exports every function include/pbf.h
declares; the Python binding's signature table covers exactly those."""
import ctypes
import re

import pytest
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "pbf.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ["pbf_create", "pbf_destroy", "pbf_set_params", "pbf_neighbor_query", "pbf_integrate",
                     "pbf_density", "pbf_pbd_constrain", "pbf_step", "pbf_step_neighbors", "pbf_get_stats",
                     "pbf_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol(capi):
    lib = ctypes.CDLL(str(capi.library_path(True)))
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_table_matches_header(capi):
    assert sorted(capi.SIGNATURES) == declared_functions()


def declared_signatures():
    """{name: [parameter type strings]} of every function include/pbf.h
    declares (comments removed, one declaration per match)."""
    text = (ROOT / "include" / "pbf.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(pbf_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(1)] = args
    return out


def _kind(decl):
    """Coarse C type class of one parameter declaration."""
    d = re.sub(r"\s+", " ", decl)
    if "*" in d:
        return "ptr"
    if "long long" in d:
        return "i64"
    if re.search(r"\b(int|unsigned)\b", d):
        return "i32"
    return "other"


def test_binding_argument_lists_match_header(capi):
    """Every ctypes signature has the header's parameter count and kinds
    (pointer, 32-bit or 64-bit integer): a call with a stale table would hand
    the library shifted arguments."""
    sigs = declared_signatures()
    assert set(sigs) == set(capi.SIGNATURES)
    for name, args in sigs.items():
        _, argtypes = capi.SIGNATURES[name]
        assert len(argtypes) == len(args), (name, len(argtypes), args)
        for decl, t in zip(args, argtypes):
            k = _kind(decl)
            if k == "ptr":
                assert t in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(t, ctypes._Pointer), (name, decl, t)
            elif k == "i64":
                assert t is ctypes.c_longlong, (name, decl, t)
            elif k == "i32":
                assert t in (ctypes.c_int, ctypes.c_uint), (name, decl, t)


def test_host_only_calls(capi):
    """Calls that touch no device: version, status strings, defaults."""
    L = capi.lib()
    assert L.pbf_abi_version() == capi.ABI_VERSION == 9
    assert L.pbf_status_string(0) == b"ok"
    assert L.pbf_status_string(-3) == b"buffer capacity exceeded"
    p = capi.default_params()
    assert abs(p.smoothing_length - 0.0125) < 1e-9 and p.iterations == 5
    assert list(p.gravity) == [0.0, -9.800000190734863, 0.0]
    assert abs(p.lambda_rest - 1000.0) < 1e-9 and abs(p.lambda_eps - 0.1) < 1e-8
    assert abs(p.corr_scale - 10.0) < 1e-9 and p.flags == 0


def test_struct_layout_matches_c_compiler(capi, tmp_path):
    """Field offsets and sizes of the ctypes mirrors equal what gcc lays out
    for include/pbf.h."""
    import subprocess
    structs = {"pbf_params": capi.pbf_params, "pbf_stats": capi.pbf_stats, "pbf_profile": capi.pbf_profile,
               "pbf_tuning": capi.pbf_tuning, "pbf_tune_report": capi.pbf_tune_report}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pbf.h"', "int main(void){"]
    for sname, cls in structs.items():
        lines.append(f'printf("{sname} %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname}.{fname} %zu\\n", offsetof({sname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                text=True).stdout.splitlines())
    for sname, cls in structs.items():
        assert int(out[sname]) == ctypes.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(out[f"{sname}.{fname}"]) == getattr(cls, fname).offset, (sname, fname)


def test_null_context_rejected(capi):
    L = capi.lib()
    assert L.pbf_destroy(None) == capi.PBF_ERR_INVALID
    assert L.pbf_step(None, None, None, None, 0, 1, None) == capi.PBF_ERR_INVALID
    assert L.pbf_last_error(None) == b"null context"


def test_header_and_binding_agree_on_the_abi_version(capi):
    m = re.search(r"#define PBF_ABI_VERSION (\d+)", (ROOT / "include" / "pbf.h").read_text())
    assert int(m.group(1)) == capi.ABI_VERSION


def test_pybind11_layer_loads(capi):
    """The pybind11 layer (_pbf_torch) links the same library version and
    refuses a null context and host tensors before any device call."""
    import torch
    m = capi.native()
    assert m.abi_version() == capi.ABI_VERSION
    x = torch.zeros((4, 3))
    with pytest.raises(RuntimeError, match="null pbf context"):
        m.rollout(x, x, 0, 1, 1, None, 0)
    with pytest.raises(RuntimeError, match="HIP device"):
        m.rollout_forward(1, x, x.clone(), None, 1, 1, None, 0)
