"""Serving trace generator (host only): deterministic, and distributed as the
workload specifies (Poisson arrivals at 200/s, log-uniform agent counts in
[1k, 256k], uniform parameters and step counts, half wealth, half SIR)."""
import math

import numpy as np


def test_trace_is_deterministic_and_well_formed():
    from jaxabm_amd.serving import make_trace, trace_checksum

    a = make_trace(2000, seed=0)
    b = make_trace(2000, seed=0)
    assert trace_checksum(a) == trace_checksum(b)
    assert trace_checksum(make_trace(2000, seed=1)) != trace_checksum(a)
    t = np.array([r.arrival for r in a])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert (gaps > 0).all()
    # exponential inter-arrival times with mean 1/200 s (3 sigma on 2000 draws)
    assert abs(gaps.mean() - 1 / 200) < 3 * (1 / 200) / math.sqrt(2000)
    n = np.array([r.n for r in a])
    assert n.min() >= 1000 and n.max() <= 256000
    # log-uniform: the median is the geometric mean of the bounds
    assert abs(math.log(np.median(n)) - 0.5 * (math.log(1000) + math.log(256000))) < 0.25
    steps = np.array([r.steps for r in a])
    assert steps.min() >= 100 and steps.max() <= 365
    kinds = [r.kind for r in a]
    assert abs(kinds.count("wealth") / 2000 - 0.5) < 0.05
    beta = np.array([r.beta for r in a if r.kind == "sir"])
    gamma = np.array([r.gamma for r in a if r.kind == "sir"])
    assert beta.min() >= 0.01 and beta.max() <= 0.1
    assert gamma.min() >= 0.05 and gamma.max() <= 0.2
