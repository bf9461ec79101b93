"""Build the reference's own C++ force kernel as a test reference (test infrastructure).

Compiles "CPU - HOOMD simulation test/dipole_forces_ext/_dipole_forces_ext.cpp"
from the reference tree where it lies (pybind11 headers + Python headers, the
same flags as the reference's setup script: -O3 -std=c++17 -fopenmp) into
oracle/_ref/.  The output is git-ignored but is kept with the tree, so test
machines without the reference sources still have it.  Only tests import it.

    python oracle/build_ref.py [/root/reference]
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")


def build(ref_root: str = "/root/reference") -> str | None:
    src = os.path.join(ref_root, "CPU - HOOMD simulation test", "dipole_forces_ext", "_dipole_forces_ext.cpp")
    if not os.path.exists(src):
        return None
    try:
        import pybind11
    except ImportError:
        return None
    os.makedirs(OUT, exist_ok=True)
    target = os.path.join(OUT, "_dipole_forces_ext" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-shared", "-fPIC", "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"], src, "-o", target]
    subprocess.run(cmd, check=True)
    return target


if __name__ == "__main__":
    print(build(*(sys.argv[1:2] or [])))
