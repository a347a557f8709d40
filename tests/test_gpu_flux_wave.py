"""The per-wave phase tiles of the mixed-precision flux sums (k_flux_wave:
grids without beaming or lensing) against the FP64 path of the same
engine, against the workgroup-tile kernel (k_flux_mixed, selected with
LCURVE_AMD_FLUX_WAVE=0 when a context is created) and against the FP64
oracle, over the tile shapes the autotuner chooses from (sub-phases per
lane R = 1, 2, 4: wave tiles of 64, 128, 256 sub-phases; face chunks of
64, 256 and 4096).  Batched draws of the batched workload's prior with both
limb-darkening laws on both stars and a mu cut-off, and a flat sum whose
sub-phases mix two inclinations (the moment sums need one inclination per
wave tile; the other faces are summed directly).
"""
import json
import os

import numpy as np
import pytest

from conftest import CASES, golden_faces, load_golden

pytestmark = pytest.mark.gpu

NPHASE = 2500


def _engine(env):
    import lcurve_re_amd as L
    keys = ("LCURVE_AMD_FLUX_WAVE", "LCURVE_AMD_FLUX_R", "LCURVE_AMD_FLUX_CHUNK")
    old = {k: os.environ.get(k) for k in keys}
    try:
        for k in keys:
            os.environ.pop(k, None)
        os.environ.update(env)
        return L.Engine(0)  # the context reads the tile settings when it is created
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _models():
    import lcurve_re_amd as L
    from lcurve_re_amd import workloads as W
    w = W.batch_prior(m=24, nlat_lo=40, nlat_hi=120)
    out = []
    for k, cfg in enumerate(w["models"][:12]):
        mp = dict(cfg["model_parameters"])
        if k % 3 == 1:  # laws exchanged
            mp["limb1"], mp["limb2"] = "Poly", "Claret"
        if k % 4 == 2:  # a mu cut-off on both stars
            mp["mucrit1"], mp["mucrit2"] = "0.05", "0.02"
        out.append(L.Model.from_json(json.dumps({"model_parameters": mp})))
    return out


def errors(calc, ref):
    d = np.abs(calc - ref)
    scale = np.max(np.abs(ref), axis=1, keepdims=True)
    nz = np.abs(ref) > 1e-12 * scale
    return float(np.max(d / scale)), float(np.max(d[nz] / np.abs(ref[nz])))


@pytest.fixture(scope="module")
def batch():
    import lcurve_re_amd as L
    models = _models()
    data = L.make_data(np.arange(NPHASE) / NPHASE, ferr=1e-3, flux=1.0)
    eng = _engine({})
    try:
        _, _, st = eng.light_curve_comp_batch(models, data, scale=False, rdata=True, precision=L.LC_PREC_FP64)
        models = [m for m, s in zip(models, st) if s == 0]  # draws the engine rejects are left out
        assert len(models) >= 8, st
        fp64, _, st = eng.light_curve_comp_batch(models, data, scale=False, rdata=True, precision=L.LC_PREC_FP64)
    finally:
        eng.close()
    assert (st == 0).all(), st
    return models, data, fp64


@pytest.mark.parametrize("R", [1, 2, 4])
@pytest.mark.parametrize("chunk", [64, 256, 4096])
def test_wave_tiles_vs_fp64_path(batch, R, chunk):
    import lcurve_re_amd as L
    models, data, fp64 = batch
    eng = _engine({"LCURVE_AMD_FLUX_R": str(R), "LCURVE_AMD_FLUX_CHUNK": str(chunk)})
    try:
        calc, _, st = eng.light_curve_comp_batch(models, data, scale=False, rdata=True, precision=L.LC_PREC_MIXED)
    finally:
        eng.close()
    assert (st == 0).all(), st
    e_max, e_pt = errors(calc, fp64)
    print(f"wave tiles R={R} chunk={chunk}: vs FP64 path max-normalised {e_max:.3e}, pointwise {e_pt:.3e}",
          flush=True)
    assert e_max <= 3e-7 and e_pt <= 5e-7, (e_max, e_pt)


def test_wave_tiles_vs_workgroup_tiles_and_oracle(batch):
    import lcurve_re_amd as L
    from oracle import pyoracle as P
    models, data, fp64 = batch
    got = {}
    for wave in ("1", "0"):
        eng = _engine({"LCURVE_AMD_FLUX_WAVE": wave, "LCURVE_AMD_FLUX_CHUNK": "128"})
        try:
            calc, _, st = eng.light_curve_comp_batch(models, data, scale=False, rdata=True,
                                                     precision=L.LC_PREC_MIXED)
        finally:
            eng.close()
        assert (st == 0).all(), st
        got[wave] = calc
    e_max, e_pt = errors(got["1"], got["0"])
    print(f"wave vs workgroup tiles: max-normalised {e_max:.3e}, pointwise {e_pt:.3e}", flush=True)
    assert e_max <= 3e-7 and e_pt <= 5e-7, (e_max, e_pt)
    for k in (0, 1, 2, 5):  # a few curves against the FP64 oracle
        st, ref, _ = P.light_curve(models[k].pod, data, scale=False, rdata=True)
        assert st == 0
        e_max, e_pt = errors(got["1"][k:k + 1], ref[None, :])
        print(f"instance {k} wave tiles vs oracle: max-normalised {e_max:.3e}, pointwise {e_pt:.3e}", flush=True)
        assert e_max <= 1e-5 and e_pt <= 1e-5, (k, e_max, e_pt)


@pytest.mark.parametrize("case", CASES[:3])
def test_flat_sum_two_inclinations(case, engine):
    """Sub-phases alternating between two inclinations: no wave tile has a
    single e_z, so every face seen is summed directly; mixed against the
    FP64 path of the same entry point."""
    import lcurve_re_amd as L
    from oracle import pyoracle as P
    from lcurve_re_amd import _abi
    g = load_golden(case)
    faces = golden_faces(g, "s0f")
    faces.flux = np.abs(faces.area).astype(np.float32) * 1e3
    flat = P.Flat(faces)
    S = 3000
    phase = np.arange(S) / S
    inc = np.where(np.arange(S) % 2 == 0, np.radians(84.0), np.radians(86.5))
    ph = P.Phases(np.sin(inc) * np.cos(2 * np.pi * phase), -np.sin(inc) * np.sin(2 * np.pi * phase),
                  np.cos(inc), phase)
    for lt, ldc_c in ((0, (0.5, 0.1, -0.05, 0.02)), (1, (0.6, -0.2, 0.3, -0.1))):
        ldc = _abi.lc_ldc(*ldc_c, 0.0, lt, 0)
        out = {}
        for prec in (L.LC_PREC_FP64, L.LC_PREC_MIXED):
            out[prec] = engine.flat_sum(0, flat.view(), ph.view(), ldc, 0.0, 1.0, 0.0, 0.0, 0, 0.0, precision=prec)
        ref = out[L.LC_PREC_FP64]
        e = float(np.max(np.abs(out[L.LC_PREC_MIXED] - ref)) / np.max(np.abs(ref)))
        print(f"{case} law {lt} two inclinations: mixed vs FP64 {e:.3e}", flush=True)
        assert e <= 1e-6, (case, lt, e)
