"""The batched adaptive-Metropolis sampler on the GPU engine against the
same sampler on the FP64 oracle (test_mcmc.py pins the sampler): three
chains of an eclipsing binary (inclination, both radii) with the same
seeds.  In FP64 mode every accept/reject decision compares the same
chi^2 values to ~1e-12, so the chains are the same states and their
chi^2 agree to 1e-9."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mcmc_chains_engine_vs_oracle(engine):
    import lcurve_re_amd as L
    from lcurve_re_amd import fit as F
    from lcurve_re_amd import mcmc as M
    from lcurve_re_amd import workloads as W
    from oracle import pyoracle as P
    phys = {"q": 0.4, "iangle": 84.0, "r1": 0.12, "r2": 0.25}
    comp = {"nlat1f": 20, "nlat2f": 20, "nlat1c": 20, "nlat2c": 20}
    truth = L.Model.from_json(json.dumps(W.model_json(phys=phys, comp=comp)))
    t = np.arange(400) / 400.0
    st, clean, _ = P.light_curve(truth.pod, L.make_data(t, ferr=1.0, flux=1.0), scale=False, rdata=False)
    assert st == 0
    rng = np.random.default_rng(8)
    ferr = 3e-3 * np.abs(clean).max() * np.ones_like(t)
    data = L.make_data(t, ferr=ferr, flux=clean + ferr * rng.standard_normal(len(t)))
    model = L.Model.from_json(json.dumps(W.model_json(phys=phys, comp=comp, vary=("iangle", "r1", "r2"))))
    # no covariance reset, a short warm-up: the Cholesky proposal is used from step ~40
    kw = dict(nchains=3, nsteps=90, burn_in=30, seed=21, cov_warmup=40, adapt_interval=20, cov_reset_step=0)
    ref = M.adaptive_metropolis(model, data, P.evaluator(), **kw)
    got = M.adaptive_metropolis(model, data, F.engine_evaluator(engine), **kw)
    print(f"mcmc: acceptance engine {got['acceptance']} oracle {ref['acceptance']}, "
          f"best chi2 {got['best_chisq']} / {ref['best_chisq']}", flush=True)
    assert np.array_equal(got["chain"], ref["chain"])
    np.testing.assert_allclose(got["chain_chisq"], ref["chain_chisq"], rtol=1e-9)
    assert np.array_equal(got["acceptance"], ref["acceptance"]) and ref["acceptance"].min() > 0
    assert ref["cov_ready"].all()
