"""Result files of the solvers: the configuration with the fitted
parameters written back (new_helpers.h:470-541) and the MCMC chain
(mcmc_solver.cpp:1488-1506).

Pinned by: the parameter strings and chain rows equal what C++ iostreams
print for the same numbers (std::fixed with precision 10 / 6, and the
default six-significant-digit format), checked by a small program compiled
here with g++; the written configuration reads back through Model.from_json
to the same parameters and keeps every other key.  CPU only."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import lcurve_re_amd as L
from lcurve_re_amd import data_io as D
from lcurve_re_amd import workloads as W

CXX = r'''
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>
int main() {
    double v;
    while (std::scanf("%la", &v) == 1) {
        std::ostringstream a, b, c;
        a << std::setprecision(10) << std::fixed << v;
        b << std::setprecision(6) << std::fixed << v;
        c << v;
        std::cout << a.str() << " " << b.str() << " " << c.str() << "\n";
    }
}
'''

VALUES = [0.0, -0.0, 0.5, 1.0 / 3.0, -2.0 / 7.0, 1e-7, 123456.789, 1e6, 9.9999995e-5, 2.5e-5, 0.1234565,
          0.12345650000000001, 20000.0, -1234567.125, 7e21, 1e-300, float("inf")]


@pytest.fixture(scope="module")
def cxx_format(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fmt")
    src, exe = d / "fmt.cpp", d / "fmt"
    src.write_text(CXX)
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], input="\n".join(float.hex(v) for v in VALUES), capture_output=True,
                         text=True, check=True).stdout.split("\n")
    return [tuple(line.split()) for line in out[:len(VALUES)]]  # by position (0.0 == -0.0 as keys)


def test_number_formats_match_iostreams(cxx_format):
    for k, v in enumerate(VALUES):
        f10, f6, g = cxx_format[k]
        assert "%.10f" % v == f10, v
        assert "%.6f" % v == f6, v
        assert format(v, "g") == g, v


def test_config_round_trip(tmp_path):
    c = W.fit_batch()["model"]
    cfg = json.loads(c) if isinstance(c, str) else json.loads(json.dumps(c))
    cfg["note"] = {"b": [1, 2], "a": "ü"}
    m = L.Model.from_json(json.dumps(cfg))
    p = m.get_param()
    p[0] *= 1.2345
    m.set_param(p)
    out = tmp_path / "fit.json"
    D.write_config_and_model(m, cfg, str(out))
    text = out.read_text()
    back = json.loads(text)
    assert back["note"] == cfg["note"]
    assert back["model_parameters"]["q"] == "%.10f %.6f %.6f %d %d" % (m.pod.q.value, m.pod.q.range, m.pod.q.dstep,
                                                                      m.pod.q.vary, m.pod.q.defined)
    assert text.startswith("{\n  \"") and text.endswith("}\n")
    assert list(back) == sorted(back)
    m2 = L.Model.from_json(text)
    np.testing.assert_allclose(m2.get_param(), m.get_param(), rtol=1e-10, atol=1e-10)
    assert [m2.get_name(i) for i in range(m2.nvary())] == [m.get_name(i) for i in range(m.nvary())]


def test_chain_file(tmp_path, cxx_format):
    pars = np.array([[VALUES[2], VALUES[3]], [VALUES[5], VALUES[6]], [VALUES[9], VALUES[10]]])
    chi = np.array([VALUES[12], VALUES[13], VALUES[8]])
    out = tmp_path / "chain.txt"
    D.write_chain(str(out), ["q", "r1"], pars, chi, thin=2)
    lines = out.read_text().splitlines()
    assert lines[0] == "step,q,r1,chisq,log_prior"
    g = lambda k: cxx_format[k][2]
    assert lines[1] == f"0,{g(2)},{g(3)},{g(12)},0"
    assert lines[2] == f"2,{g(9)},{g(10)},{g(8)},0"
    assert len(lines) == 3


# ---------------------------------------------------------- plot frames
FRAME_EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "_build", "json_frame_check")


@pytest.mark.skipif(not os.path.exists(FRAME_EXE), reason="tests/cxx/json_frame_check not built (no json.hpp)")
@pytest.mark.parametrize("seed,base_time", [(1, 0.0), (2, 58000.25), (3, -3.5)])
def test_plot_frame_matches_nlohmann(seed, base_time):
    """The stream plot device's frame (new_helpers.h:71-109) byte for byte
    against the same frame built with nlohmann::json: sorted keys, shortest
    round-trip numbers, the time origin rule, the error floor and meta."""
    import subprocess

    import lcurve_re_amd as L
    from lcurve_re_amd import data_io as D
    rng = np.random.default_rng(seed)
    n = 40
    t = base_time + np.sort(rng.uniform(0, 0.2, n))
    flux = np.concatenate([rng.normal(1, 0.1, n - 4), [0.0, 1e-12, 3e20, -2.5]])
    ferr = np.concatenate([rng.uniform(1e-4, 1e-2, n - 3), [0.0, -1.0, np.inf]])
    fit = flux * (1 + rng.normal(0, 1e-3, n))
    meta = {"iteration": int(seed), "chisq": float(rng.uniform(1, 1e4)), "label": "lm"}
    data = L.make_data(t, ferr=ferr, flux=flux)
    got = "@@LCURVE_PLOT@@" + D.model_frame(data, fit, no_file=True, metadata=meta) + "\n"
    text = f"{n}\n" + "".join(f"{float(a)!r} {float(b)!r} {float(c)!r} {float(d)!r}\n"
                              for a, b, c, d in zip(t, flux, ferr, fit)) + json.dumps(meta) + "\n"
    r = subprocess.run([FRAME_EXE], input=text, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert got == r.stdout
    # the writer adds the marker and flushes
    import io
    buf = io.StringIO()
    D.stream_model_frame(data, fit, no_file=True, metadata=meta, out=buf)
    assert buf.getvalue() == r.stdout
    assert D.plotting_streamed(' "Stream" ') and D.plotting_disabled("'None'") and not D.plotting_disabled("x11")


def test_plot_model_devices():
    import io

    import lcurve_re_amd as L
    from lcurve_re_amd import data_io as D
    data = L.make_data(np.array([0.1, 0.2]), ferr=np.array([0.01, 0.02]), flux=np.array([1.0, 1.1]))
    buf = io.StringIO()
    assert D.plot_model(data, [1.0, 1.0], False, data, " Stream ", out=buf)
    frame = json.loads(buf.getvalue()[len("@@LCURVE_PLOT@@"):])
    assert frame["meta"] == {"phase": "final"} and frame["protocol"] == "lcurve.plot.v1"
    assert frame["residual"] == [0.0, (1.1 - 1.0) / 0.02]
    buf = io.StringIO()
    assert not D.plot_model(data, [1.0, 1.0], False, data, "NULL", out=buf) and buf.getvalue() == ""
