"""GPU: the reference's serial runners -- SensitivityAnalysis.run
(jaxabm/analysis.py:97-165), the calibrators' robust and population
evaluations (jaxabm/analysis.py:434-476), run_parallel_simulations
(jaxabm/utils.py:128-176) and Model.batch_run (jaxabm/agentpy.py:1116-1152)
-- submit library models to the continuous-batching server as one batch.
Each result must equal the model's own run() bit for bit."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd.programs import library

    return library


def same_results(a, b):
    da, db = a._data, b._data
    assert set(da) == set(db)
    for k in da:
        assert da[k] == db[k], k


def sir_factory(L, solo=False):
    cls = L.NetworkSIRModel
    if solo:
        class Solo(L.NetworkSIRModel):   # no server request: runs on its own
            serve_request = None
        cls = Solo

    def factory(params=None, config=None):
        p = {"n": 3000, "m": 3, "n_infected": 10, "graph_seed": 5, "steps": 40}
        p.update(params or {})
        if "n" in p:
            p["n"] = int(p["n"])
        return cls(p, config=config)

    return factory


def test_sensitivity_analysis_batched_equals_serial(L, capsys):
    from jaxabm_amd.analysis import SensitivityAnalysis
    from jaxabm_amd.core import ModelConfig

    f = sir_factory(L)
    sa = SensitivityAnalysis(f, {"beta": (0.02, 0.2), "gamma": (0.05, 0.3)}, ["I", "R"],
                             num_samples=6, seed=3)
    res = sa.run(verbose=False)
    assert "simulations executed together" in capsys.readouterr().out
    for i in range(6):
        params = {"beta": float(sa.samples[i, 0]), "gamma": float(sa.samples[i, 1])}
        r = f(params=params, config=ModelConfig(seed=i + 1000)).run()
        assert res["I"][i] == np.float32(r._data["I"][-1])
        assert res["R"][i] == np.float32(r._data["R"][-1])


def test_wealth_batch_run_and_parallel_simulations(L, capsys):
    from jaxabm_amd.utils import run_parallel_simulations

    base = L.WealthModel({"n": 800, "steps": 25}, seed=4)
    got = base.batch_run({"n": [300, 1201]}, repetitions=2)
    assert "simulations executed together" in capsys.readouterr().out
    for n in (300, 1201):
        for rep in range(2):
            want = L.WealthModel({"n": n, "steps": 25}, seed=4 + rep).run()
            same_results(got[(n,)][rep], want)

    f = sir_factory(L)
    sets = [{"beta": 0.05}, {"beta": 0.15, "n": 5000}]
    out = run_parallel_simulations(f, sets, num_runs=2, seed_offset=9)
    from jaxabm_amd.core import ModelConfig

    assert [r["seed"] for r in out] == [9, 10, 11, 12]
    for r in out:
        want = f(params=r["params"], config=ModelConfig(seed=r["seed"])).run()
        for k in ("S", "I", "R", "step"):
            assert r[k] == want._data[k], k


def test_calibrator_population_and_gradient_batched_equal_serial(L):
    from jaxabm_amd.analysis import ModelCalibrator

    def make(solo):
        return ModelCalibrator(sir_factory(L, solo), {"beta": 0.08, "gamma": 0.12},
                               {"R": 2000.0, "I": 50.0},
                               param_bounds={"beta": (0.01, 0.3), "gamma": (0.02, 0.4)},
                               evaluation_steps=30, num_evaluation_runs=3, seed=11)

    a, b = make(False), make(True)
    pop = np.array([[0.05, 0.1], [0.1, 0.2], [0.2, 0.05], [0.3, 0.4]])
    t0 = time.time()
    la = a._population_loss(pop)
    t1 = time.time()
    lb = b._population_loss(pop)
    t2 = time.time()
    print(f"population of 4 x 3 runs: batched {t1 - t0:.3f} s, one by one {t2 - t1:.3f} s")
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(a.key, b.key)
    p = {"beta": 0.07, "gamma": 0.11}
    np.testing.assert_array_equal(a.grad_fn(p), b.grad_fn(p))
    ea, eb = a._evaluate_params_robust(p), b._evaluate_params_robust(p)
    assert ea == eb
