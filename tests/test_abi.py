"""The C ABI library: it loads, exports every symbol include/mpol.h declares,
and its host-side table routine matches the reference's golden tables.
No GPU needed."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, REPO, gpu_available

HEADER = os.path.join(REPO, "include", "mpol.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mpol_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from mpol_hip import _lib
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(syms) <= bound, set(syms) - bound


def test_version_and_error_string():
    from mpol_hip import _lib
    lib = _lib.load()
    assert lib.mpol_version() >= 100
    assert isinstance(lib.mpol_last_error(), bytes)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "tables_*.npz"))), ids=os.path.basename)
def test_library_tables_match_reference(path):
    from mpol_hip import real_space_table
    d = np.load(path)
    r = d["r"]
    tabs = real_space_table(r, float(d["xi"]))
    for k in range(6):
        assert tabs[k][0] == pytest.approx(d["self_terms"][k], rel=1e-14, abs=1e-300)
        ref = d["tables"][k]
        big = r >= 0.1
        assert np.max(np.abs(tabs[k][1:][big] - ref[big])) / np.max(np.abs(ref)) < 1e-10


def test_table_argument_errors():
    from mpol_hip import real_space_table
    with pytest.raises(ValueError):
        real_space_table(np.array([0.0, 1.0]), 0.5)
    with pytest.raises(ValueError):
        real_space_table(np.array([1.0]), -1.0)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    from mpol_hip import Context
    with pytest.raises(RuntimeError):
        Context()
