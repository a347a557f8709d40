"""The C-ABI library loads (no GPU needed), exports every function
include/lcurve_amd.h declares, and its struct layouts match the ctypes
mirror.  Host-only entry points (model parsing, grid sizes) run here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lcurve_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(lc_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_exports_every_declared_symbol():
    from lcurve_re_amd import _abi
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.SIGNATURES), set(names) ^ set(_abi.SIGNATURES)


def test_struct_layouts():
    from lcurve_re_amd import _abi
    lib = _abi.load()
    for name, cls in _abi.struct_sizes().items():
        assert lib.lc_sizeof(name.encode()) == C.sizeof(cls), name
    assert lib.lc_sizeof(b"nope") == 0


def test_version_without_gpu():
    from lcurve_re_amd import _abi
    assert b"gfx950" in _abi.load().lc_version()


def test_grid_sizes_match_oracle():
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    lib = P.oracle()
    for w in (W.detached_small(), W.semidetached_large()):
        m = L.Model.from_json(w["model"])
        for star in (0, 1):
            for fine in (0, 1):
                n = C.c_size_t()
                assert lib.orc_star_grid_size(C.byref(m.pod), star, fine, C.byref(n)) == 0
                assert m.grid_size(star, fine) == n.value


def test_numface_known_sizes():
    """numface(40) = 2058, numface(200) = 50872, numface(400) = 203560
    (numface.cpp:44-51 evaluated for the benchmark grids)."""
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    for nlat, n in ((40, 2058), (200, 50872), (400, 203560)):
        m = L.Model.from_json(W.model_json(comp={"nlat1f": nlat, "nlat1c": nlat}))
        assert m.grid_size(0, True) == n
