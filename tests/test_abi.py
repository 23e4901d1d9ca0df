"""C-ABI checks that need no GPU: the HIP library loads, exports every symbol
declared in include/jaxabm_amd.h (and the ctypes table covers them), and its
host-side functions (network construction, host key schedule) agree with the
oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(REPO, "include", "jaxabm_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(abm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_header_symbol():
    from jaxabm_amd._lib import SIGNATURES, lib

    L = lib()
    syms = header_symbols()
    assert len(syms) > 30
    for name in syms:
        assert hasattr(L, name), f"{name} declared in the header but not exported"
    assert set(syms) == set(SIGNATURES), (set(syms) ^ set(SIGNATURES))


def test_abi_version_and_error_channel():
    from jaxabm_amd._lib import AbmError, check, lib

    assert lib().abm_version() == 1
    # an argument error is reported without touching a device
    rc = lib().abm_host_split(np.zeros(2, np.uint32), -1, np.zeros(2, np.uint32), 0)
    assert rc != 0
    with pytest.raises(AbmError, match="abm_host_split"):
        check(rc, "host_split")


@pytest.mark.parametrize("n,m,seed", [(10, 4, 0), (5000, 4, 3), (2000, 9, 1), (4, 4, 0)])
def test_network_builders_match_oracle(n, m, seed):
    from jaxabm_amd import network as N

    a = N.ba_edges(n, m, seed)
    b = O.ba_edges(n, m, O.prng_key(seed))
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    rp, col = N.csr_undirected(n, *a)
    rp2, col2 = O.csr_undirected(n, *b)
    np.testing.assert_array_equal(rp, rp2)
    np.testing.assert_array_equal(col, col2)


def test_network_builder_at_the_measured_size_matches_oracle_and_checksum():
    """BA(1,000,000, m = 4, graph seed 0), the measured SIR graph: the
    product's host builder equals the oracle's builder entry for entry, and
    both match the committed checksum that bench.py asserts."""
    import json

    from jaxabm_amd import network as N

    gold = json.load(open(os.path.join(REPO, "tests", "golden", "sir_ba1M_m4_365steps.json")))
    rp, col = N.ba_csr(1_000_000, 4, 0)
    orp, ocol = O.ba_csr(1_000_000, 4, 0)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    assert int(rp[-1]) == gold["csr_nnz"] == 2 * 4 * (1_000_000 - 4)
    assert N.csr_checksum(rp, col) == gold["csr_sha256"]


def test_csr_rejects_out_of_range_edges():
    from jaxabm_amd import network as N
    from jaxabm_amd._lib import AbmError

    with pytest.raises(AbmError, match="out of range"):
        N.csr_undirected(3, np.array([0, 1], np.int32), np.array([1, 3], np.int32))


def test_node_chunks_cover_and_align():
    from jaxabm_amd.network import node_chunks

    rp, _ = O.ba_csr(20000, 4, 2)
    cp = node_chunks(rp, 64, 4, 5000, 8192, 32, 128)
    assert cp[0] == 0 and cp[-1] == 20000
    assert (np.diff(cp) > 0).all() and (np.diff(cp) <= 8192).all()
    assert all(x % 64 == 0 for x in cp[:-1])


@pytest.mark.parametrize("mode", [0, 1])
def test_host_key_schedule_matches_oracle(mode):
    """abm_step_keys restates jaxabm/model.py:156,164,183 on the host."""
    from jaxabm_amd import random as jr

    rng = np.array([7, 11], np.uint32)
    tab = jr.step_key_table(rng.copy(), 5, 2, True, True, mode)
    r = rng.copy()
    for t in range(5):
        r, sk = O.split(r, 2, mode)
        sk, c0 = O.split(sk, 2, mode)
        sk, c1 = O.split(sk, 2, mode)
        sk, uk = O.split(sk, 2, mode)
        a, b = O.split(uk, 2, mode)
        np.testing.assert_array_equal(tab[t], np.concatenate([c0, c1, a, b]))
    np.testing.assert_array_equal(jr.host_split([3, 4], 5, mode), O.split([3, 4], 5, mode))


def test_unit_threshold_is_exact():
    """m < ceil(p * 2^23) <=> (1 + m 2^-23) - 1 < p for float32 p."""
    from jaxabm_amd._lib import lib

    for p in [0.05, 0.1, 0.0, 1.0, 1e-7, 0.999999, 0.375, 0.2]:
        thr = lib().abm_unit_threshold(C.c_float(p))
        pf = np.float32(p)
        ms = np.arange(max(0, thr - 3), min(1 << 23, thr + 3))
        u = (np.uint32(0x3F800000) | ms.astype(np.uint32)).view(np.float32) - np.float32(1.0)
        np.testing.assert_array_equal(ms < thr, u < pf)


def test_schelling_entry_points_reject_bad_arguments_without_a_device():
    """The Schelling entry points validate on the host before any launch: a
    continuing run needs the previous step's counters, the settled records
    need a grid, and grids above 15,360,000 cells are refused (the empty-cell
    block prefix and the agents kernel's static LDS stay within 64 KB)."""
    from jaxabm_amd._lib import AbmError, SchellingGrid, check, lib

    fake = 0x1000   # never dereferenced: every call fails before a launch
    L = lib()

    def grid(h, w):
        return SchellingGrid(h, w, 0, *([fake] * 8))

    g = grid(64, 64)
    with pytest.raises(AbmError, match="abm_schelling_run_from: prev_counters is null"):
        check(L.abm_schelling_run_from(C.byref(g), fake, 4, 0, None, fake, fake, fake, None),
              "run_from")
    with pytest.raises(AbmError, match="abm_schelling_settled: bad grid descriptor"):
        check(L.abm_schelling_settled(None, fake, 4, fake, fake, fake, None), "settled")
    big = grid(4000, 3841)   # 15,364,000 cells: 15,004 blocks of 1,024
    for rc in (L.abm_schelling_run(C.byref(big), fake, 4, 0, fake, fake, fake, None),
               L.abm_schelling_run_from(C.byref(big), fake, 4, 0, fake, fake, fake, fake, None)):
        with pytest.raises(AbmError, match="15,360,000 cells"):
            check(rc, "run")
