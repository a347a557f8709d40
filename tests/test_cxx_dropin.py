"""The C++ drop-in header (include/lcurve_re/lcurve.h) through a program
built from tests/cxx/dropin_check.cpp, which calls Lcurve::Model,
Lcurve::Data and Lcurve::light_curve_comp the way the reference's fitting
programs do.

CPU tests: the parameter API, the ASCII data format (lcurve.cpp:14-100)
and that an evaluation without a GPU raises Lcurve_Error (no CPU path).
GPU tests: light_curve_comp / _fast / _batch against the FP64 oracle.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cxx", "_build", "dropin_check")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="tests/cxx not built (run __graft_entry__.build())")

DATA_TEXT = """# time expose flux ferr weight ndiv
0.10 0.0 1.0 0.01 1 1

 0.2 this line starts with a space and is skipped
\t0.25 so is this one
0.30 0.001 1.5 -0.02 1 3
0.40 0 2 0.03 0.5 1"""


def run(*args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout


def parse_lines(out):
    d = {}
    for line in out.splitlines():
        k, *v = line.split()
        d[k] = v
    return d


@pytest.fixture()
def fit_config(tmp_path):
    cfg = W.model_json(vary=W.FIT_PARAMS, comp={"use_radii": "0"}, phys={"cphi3": 0.02, "cphi4": 0.06})
    p = tmp_path / "cfg.json"
    p.write_text(json.dumps(cfg))
    return str(p), cfg


def test_model_api_through_cxx(fit_config):
    path, cfg = fit_config
    rc, out = run("model", path)
    assert rc == 0, out
    d = parse_lines(out)
    m = L.Model.from_json(json.dumps(cfg))
    names = [n for n in W.FIT_PARAMS if n not in ("r1", "r2")]
    assert d["names"] == names
    assert int(d["nvary"][0]) == len(names)
    np.testing.assert_array_equal(np.array(d["param"], float), m.get_param())
    np.testing.assert_array_equal(np.array(d["range"], float), m.get_range())
    np.testing.assert_array_equal(np.array(d["dstep"], float), m.get_dstep())
    # Model::get_r1r2 for the cphi parameterisation (lcurve.h:399-409)
    sini = math.sin(math.pi * 85.0 / 180.0)
    rp = math.sqrt(1.0 - (sini * math.cos(2.0 * math.pi * 0.06)) ** 2)
    rm = math.sqrt(1.0 - (sini * math.cos(2.0 * math.pi * 0.02)) ** 2)
    r1, r2 = (float(x) for x in d["r1r2"])
    assert abs(r1 - (rp - rm) / 2) <= 1e-15 and abs(r2 - (rp + rm) / 2) <= 1e-15
    assert d["q"][0] == "0.5"          # implicit Pparam -> double conversion
    assert d["legal"] == ["1"]
    np.testing.assert_array_equal(np.array(d["param2"], float), np.array(d["param"], float) * (1 + 1e-3))


def test_data_ascii_round_trip(tmp_path):
    src, dst = tmp_path / "in.dat", tmp_path / "out.dat"
    src.write_text(DATA_TEXT)
    rc, out = run("data", str(src), str(dst))
    assert rc == 0 and out.strip() == "n 3", out
    # the negative uncertainty zeroes ferr and weight (lcurve.cpp:26-29)
    expect = ["       0.10000000 0.00000000    1.00000 0.01000 1.00000 1",
              "       0.30000000 0.00100000    1.50000 0.00000 0.00000 3",
              "       0.40000000 0.00000000    2.00000 0.03000 0.50000 1"]
    assert dst.read_text().splitlines() == expect
    py = L.read_data(str(src))
    L.write_data(str(tmp_path / "py.dat"), py)
    assert (tmp_path / "py.dat").read_text() == dst.read_text()


def test_data_written_by_wrasc_reads_back_empty(tmp_path):
    # the reference's own quirk, kept: wrasc pads the time to 17 columns and
    # rasc skips lines starting with a space (lcurve.cpp:35-44, 66-69)
    src, dst = tmp_path / "in.dat", tmp_path / "out.dat"
    src.write_text(DATA_TEXT)
    assert run("data", str(src), str(dst))[0] == 0
    rc, out = run("data", str(dst), str(tmp_path / "again.dat"))
    assert rc == 0 and out.strip() == "n 0"
    assert len(L.read_data(str(dst))) == 0


def test_data_malformed_line_raises(tmp_path):
    src = tmp_path / "bad.dat"
    src.write_text("0.1 0 1 0.01 1 1\n0.2 0 x 0.01 1 1\n")
    rc, out = run("data", str(src), str(tmp_path / "o.dat"))
    assert rc == 3 and "failed to read t,e,f,fe,wgt,ndiv" in out
    with pytest.raises(RuntimeError):
        L.read_data(str(src))


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except Exception:
        return True


@pytest.mark.skipif(not _no_gpu(), reason="checks the no-GPU error path")
def test_evaluation_without_gpu_raises(tmp_path, fit_config):
    path, _ = fit_config
    src = tmp_path / "in.dat"
    src.write_text(DATA_TEXT)
    rc, out = run("error", path, str(src))
    assert rc == 0 and out.startswith("Lcurve_Error: "), out


@pytest.mark.gpu
def test_light_curve_through_cxx_vs_oracle(tmp_path):
    from oracle import pyoracle as P
    w = W.detached_small(nphase=400)
    cfg = tmp_path / "cfg.json"
    cfg.write_text(w["model"] if isinstance(w["model"], str) else json.dumps(w["model"]))
    data = L.make_data(w["time"], expose=w["expose"], ndiv=w["ndiv"], ferr=w["ferr"], weight=w["weight"],
                       flux=w["flux"])
    # Data::wrasc right-aligns the time in 17 columns, and Data::rasc skips
    # lines that start with a space (lcurve.cpp:35-44, 66-69), so the input
    # is written without leading blanks
    v = L.data_view(data)
    with open(tmp_path / "in.dat", "w") as f:
        for r in v:
            f.write(" ".join(repr(float(r[k])) for k in ("time", "expose", "flux", "ferr", "weight"))
                    + f" {int(r['ndiv'])}\n")
    data = L.read_data(str(tmp_path / "in.dat"))
    assert len(data) == 400
    print("cxx dropin eval detached_small nphase=400", flush=True)
    rc, out = run("eval", str(cfg), str(tmp_path / "in.dat"), str(tmp_path / "out.txt"))
    assert rc == 0, out
    d = parse_lines((tmp_path / "out.txt").read_text())
    m = L.Model.from_json(cfg.read_text())
    st, ref, _ = P.light_curve(m.pod, data, scale=False, rdata=True, diagnostics=False)
    assert st == 0
    calc = np.array(d["comp"], float)
    assert np.max(np.abs(calc - ref)) / np.max(np.abs(ref)) <= 1e-5
    for i in range(3):
        row = d[f"batch{i}"]
        assert row[0] == "0"
    np.testing.assert_allclose(np.array(d["batch0"][1:], float), calc, rtol=0, atol=1e-12 * np.max(np.abs(calc)))
    st, fast, _ = P.light_curve(m.pod, data, scale=False, rdata=True, diagnostics=False, max_model_points=50)
    assert st == 0
    assert np.max(np.abs(np.array(d["fast"], float) - fast)) / np.max(np.abs(fast)) <= 1e-5
    assert int(d["scaled"][3]) == 5


@pytest.mark.gpu
def test_roche_overflow_raises_lcurve_error(tmp_path):
    cfg = W.model_json(phys={"r1": 0.9})
    p = tmp_path / "cfg.json"
    p.write_text(json.dumps(cfg))
    src = tmp_path / "in.dat"
    src.write_text(DATA_TEXT)
    print("cxx dropin roche overflow", flush=True)
    rc, out = run("error", str(p), str(src))
    assert rc == 0 and out.startswith("Lcurve_Error: "), out
