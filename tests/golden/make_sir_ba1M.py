"""Writes sir_ba1M_m4_365steps.json: the checksum of the Barabasi-Albert
(n = 1,000,000, m = 4, graph seed 0) CSR and the oracle's S/I/R counts of the
365-step network SIR run on it (beta 0.05, gamma 0.1, 100 seeded infections,
seed 0), in both threefry stream modes.  The measured bench workload and the
parity tests are pinned to these values.

Run from the repository root: python tests/golden/make_sir_ba1M.py"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402

N, M, GRAPH_SEED, STEPS, BETA, GAMMA, N_SEED, SEED = 1_000_000, 4, 0, 365, 0.05, 0.1, 100, 0


def csr_sha256(row_ptr, col) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(row_ptr, "<i4").tobytes())
    h.update(np.ascontiguousarray(col, "<i4").tobytes())
    return h.hexdigest()


def main():
    rp, col = O.ba_csr(N, M, GRAPH_SEED)
    out = {"n": N, "m": M, "graph_seed": GRAPH_SEED, "csr_nnz": int(rp[-1]),
           "csr_sha256": csr_sha256(rp, col), "steps": STEPS, "beta": BETA, "gamma": GAMMA,
           "seed_infected": N_SEED, "seed": SEED, "counts": {}}
    for mode, name in ((O.MODE_PARTITIONABLE, "partitionable"), (O.MODE_ORIGINAL, "original")):
        st, rng = O.sir_init(N, N_SEED, SEED, mode)
        c = O.sir_run_threads(rp, col, st, rng, np.float32(BETA), np.float32(GAMMA), STEPS,
                              os.cpu_count() or 1, mode)
        out["counts"][name] = c.tolist()
    with open(os.path.join(HERE, "sir_ba1M_m4_365steps.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
