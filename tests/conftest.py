import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["detached", "semidetached", "async", "spheres", "lowq"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))


def golden_model(g):
    import json

    import lcurve_re_amd as L
    return L.Model.from_json(json.loads(bytes(g["model_json"]).decode()))


def golden_faces(g, key):
    from oracle import pyoracle as P
    f = P.Faces(len(g[f"{key}_px"]))
    for k in ("px", "py", "pz", "dx", "dy", "dz", "area", "gravity", "necl", "ingress", "egress"):
        setattr(f, k, np.ascontiguousarray(g[f"{key}_{k}"]))
    if f"{key}_flux" in g:
        f.flux = np.ascontiguousarray(g[f"{key}_flux"])
    return f


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle as P
    return P.oracle()


@pytest.fixture(scope="session")
def engine():
    import lcurve_re_amd as L
    e = L.Engine(0)
    yield e
    e.close()
