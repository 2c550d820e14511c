"""Builds the pybind11 layer pbf_torch.cpp (a PyTorch C++ extension, host code
only) into ../_build/_pbf_torch<EXT_SUFFIX>, linked against libpbf.so in the
same folder. One g++ call with the include paths, library paths and defines
that torch.utils.cpp_extension uses for its own builds."""
import subprocess
import sys
import sysconfig
from pathlib import Path

import torch
import torch.utils.cpp_extension as ce

HERE = Path(__file__).resolve().parent
OUT = HERE.parent / "_build"
NAME = "_pbf_torch"


def command():
    target = OUT / (NAME + sysconfig.get_config_var("EXT_SUFFIX"))
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", f"-DTORCH_EXTENSION_NAME={NAME}", "-DTORCH_API_INCLUDE_EXTENSION_H",
           "-I", str(HERE.parent.parent / "include"), "-I", sysconfig.get_paths()["include"]]
    for inc in ce.include_paths():
        cmd += ["-isystem", inc]
    cmd += [str(HERE / "pbf_torch.cpp"), "-o", str(target)]
    for lib in ce.library_paths():
        cmd += ["-L", lib, "-Wl,-rpath," + lib]
    cmd += ["-L", str(OUT), "-lpbf", "-Wl,-rpath,$ORIGIN", "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python"]
    return cmd, target


def main():
    cmd, target = command()
    OUT.mkdir(exist_ok=True)
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return target


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
