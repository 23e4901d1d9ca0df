"""Pins of the oracle's model programs (CPU).

Each program is checked against an independent dense numpy re-evaluation of
one step built from the oracle's PRNG pieces (pinned in test_oracle_rng.py),
and against invariants a wrong index or term would break:

* BA/CSR: E = m(n-m), symmetric, sorted rows, no self loops or duplicates;
* SIR: dense any-over-edges step vs the early-exit oracle; S+I+R = n,
  S non-increasing, R non-decreasing;
* wealth: dense bincount step; sum(wealth) = n, mean = 1, wealth >= 0; torus
  positions in [0, 1];
* Gini: pairwise |xi - xj| identity and the f32 rounding of the reference
  formula;
* Schelling: dense Moore-8 step; type counts conserved, one agent per cell;
* Deffuant: sum of opinions conserved, dense pair step; adjoint vs central
  finite differences of the float64 forward.  The differences converge to
  the adjoint as h shrinks (7 digits at h = 1e-6); at h = 1e-3 the loss is too
  curved (sigmoid gates of scale 0.02 compounded over the steps) for a
  central difference to be accurate to better than ~20 %, so the pin uses
  h = 1e-6 and checks the convergence sequence.
"""
import numpy as np
import pytest

import oracle as O

MODES = [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE]


def step_keys(rng, mode):
    rng2, sk = O.split(rng, 2, mode)
    sk2, ck = O.split(sk, 2, mode)
    _, uk = O.split(sk2, 2, mode)
    return rng2, ck, uk


# ---------------------------------------------------------------- network
@pytest.mark.parametrize("n,m", [(2, 1), (5, 4), (6, 4), (1000, 4), (3000, 7)])
def test_ba_csr_structure(n, m):
    src, dst = O.ba_edges(n, m, O.prng_key(5))
    assert len(src) == (m * (n - m) if n > m else n - 1)
    assert (src != dst).all()
    pairs = {(min(a, b), max(a, b)) for a, b in zip(src, dst)}
    assert len(pairs) == len(src)
    rp, col = O.csr_undirected(n, src, dst)
    assert rp[-1] == 2 * len(src)
    for v in range(n):
        row = col[rp[v]:rp[v + 1]]
        assert (np.diff(row) > 0).all()
        for u in row:
            assert v in col[rp[u]:rp[u + 1]]
    if n > m:
        assert (np.diff(rp)[m + 1:] >= m).all()


# ---------------------------------------------------------------- SIR
def dense_sir_step(rp, col, state, rng, beta, gamma, mode):
    n, nnz = len(state), len(col)
    rng, ck, uk = step_keys(rng, mode)
    kinf, krec = O.split(uk, 2, mode)
    u = O.uniform(kinf, nnz, mode=mode)
    r = O.uniform(krec, n, mode=mode)
    row = np.repeat(np.arange(n), np.diff(rp))
    hit = (state[col] == 1) & (u < np.float32(beta))
    any_hit = np.zeros(n, bool)
    np.logical_or.at(any_hit, row, hit)
    new = state.copy()
    new[(state == 0) & any_hit] = 1
    new[(state == 1) & (r < np.float32(gamma))] = 2
    return new, rng


@pytest.mark.parametrize("mode", MODES)
def test_sir_matches_dense_step(mode):
    n = 3000
    rp, col = O.ba_csr(n, 4, 11)
    st, rng = O.sir_init(n, 30, 2, mode)
    assert (st == 1).sum() == 30
    ref, ref_rng = st.copy(), rng.copy()
    for t in range(25):
        ref, ref_rng = dense_sir_step(rp, col, ref, ref_rng, 0.08, 0.1, mode)
        c = O.sir_run(rp, col, st, rng, 0.08, 0.1, 1, mode)[0]
        np.testing.assert_array_equal(st, ref)
        np.testing.assert_array_equal(rng, ref_rng)
        np.testing.assert_array_equal(c, np.bincount(ref, minlength=3))


def test_sir_invariants():
    n = 5000
    rp, col = O.ba_csr(n, 3, 1)
    st, rng = O.sir_init(n, 50, 7, O.MODE_PARTITIONABLE)
    c = O.sir_run(rp, col, st, rng, 0.05, 0.1, 200, O.MODE_PARTITIONABLE)
    assert (c.sum(1) == n).all()
    assert (np.diff(c[:, 0]) <= 0).all() and (np.diff(c[:, 2]) >= 0).all()
    assert c[-1, 2] > 0


def test_sir_seeding_is_permutation_prefix():
    for mode in MODES:
        st, _ = O.sir_init(1000, 17, 4, mode)
        ik = O.split(O.prng_key(4), 2, mode)[1]
        np.testing.assert_array_equal(np.sort(np.nonzero(st)[0]), np.sort(O.permutation(ik, 1000, mode)[:17]))


# ---------------------------------------------------------------- wealth
def dense_wealth_step(x, y, w, rng, mode):
    n = len(w)
    rng, ck, uk = step_keys(rng, mode)
    ak = O.split(ck, n, mode)
    dx = np.array([O.normal(O.split(k, 2, mode)[0], 1, mode)[0] for k in ak], np.float32)
    dy = np.array([O.normal(O.split(k, 2, mode)[1], 1, mode)[0] for k in ak], np.float32)
    x = np.mod(x + np.float32(0.01) * dx, np.float32(1.0)).astype(np.float32)
    y = np.mod(y + np.float32(0.01) * dy, np.float32(1.0)).astype(np.float32)
    t = O.randint(uk, n, 0, n - 1, mode)
    t = t + (t >= np.arange(n))
    give = (w > 0).astype(np.int32)
    w = w - give + np.bincount(t, weights=give, minlength=n).astype(np.int32)
    return x, y, w, rng


@pytest.mark.parametrize("mode", MODES)
def test_wealth_matches_dense_step(mode):
    n = 300
    x, y, w, rng = O.wealth_init(n, 0, mode)
    ak = O.split(O.split(O.prng_key(0), 2, mode)[1], n, mode)
    np.testing.assert_array_equal(x, [O.uniform(O.split(k, 2, mode)[0], 1, mode=mode)[0] for k in ak])
    rx, ry, rw, rr = x.copy(), y.copy(), w.copy(), rng.copy()
    for _ in range(5):
        rx, ry, rw, rr = dense_wealth_step(rx, ry, rw, rr, mode)
        mean, gini = O.wealth_run(x, y, w, rng, 1, mode)
        np.testing.assert_array_equal(w, rw)
        np.testing.assert_array_equal(rng, rr)
        # numpy's float32 mod equals jnp.mod (fmod plus divisor on sign change)
        np.testing.assert_array_equal(x, rx)
        np.testing.assert_array_equal(y, ry)
        assert mean[0] == np.float32(1.0)
        assert gini[0] == O.gini_i32(rw)


def test_wealth_invariants():
    n = 1000
    x, y, w, rng = O.wealth_init(n, 0, O.MODE_PARTITIONABLE)
    mean, gini = O.wealth_run(x, y, w, rng, 100, O.MODE_PARTITIONABLE)
    assert w.sum() == n and (w >= 0).all()
    assert (mean == np.float32(1.0)).all()
    assert (gini > 0).all() and (gini < 1).all() and gini[-1] > gini[0]
    assert (x >= 0).all() and (x <= 1).all() and (y >= 0).all() and (y <= 1).all()


@pytest.mark.parametrize("v", [[1, 1, 1, 1], [0, 0, 0, 5], [3, 1, 4, 1, 5, 9, 2, 6], list(range(100)), [0, 0]])
def test_gini_pairwise_identity(v):
    v = np.array(v, np.int32)
    n, s = len(v), int(v.sum())
    if s == 0:
        assert O.gini_i32(v) == 0.0
        return
    exact = np.abs(v[:, None].astype(np.int64) - v[None, :]).sum() / (2.0 * n * s)
    assert abs(float(O.gini_i32(v)) - exact) < 5e-7
    # the reference's float32 form: f32(2 S) / f32(n T) - f32((n + 1) / n)
    xs = np.sort(v)
    S = int((np.arange(1, n + 1) * xs).sum())
    want = np.float32(np.float32(2 * S) / np.float32(n * s)) - np.float32((n + 1) / n)
    assert O.gini_i32(v) == want


# ---------------------------------------------------------------- Schelling
def dense_schelling_step(H, W, cells, pos, typ, rng, mode):
    rng, ck, uk = step_keys(rng, mode)
    na = len(pos)
    ak = O.split(ck, na, mode)
    grid = cells.reshape(H, W)
    occ = np.zeros((H, W), int)
    same = {1: np.zeros((H, W), int), 2: np.zeros((H, W), int)}
    for dr in (-1, 0, 1):
        for dc in (-1, 0, 1):
            if dr or dc:
                sh = np.roll(np.roll(grid, -dr, 0), -dc, 1)
                occ += sh != 0
                for ty in (1, 2):
                    same[ty] += sh == ty
    r, c = pos // W, pos % W
    n_occ = occ[r, c]
    n_same = np.where(typ == 1, same[1][r, c], same[2][r, c])
    happy = 8 * n_same >= 3 * n_occ
    empty = np.nonzero(cells == 0)[0]
    target = np.full(na, -1)
    for k in np.nonzero(~happy)[0]:
        j = O.randint(ak[k], 1, 0, len(empty), mode)[0]
        target[k] = empty[j]
    claim = {}
    for k in range(na):
        if target[k] >= 0 and target[k] not in claim:
            claim[target[k]] = k
    cells = cells.copy()
    pos = pos.copy()
    for cell, k in claim.items():
        cells[pos[k]] = 0
        cells[cell] = typ[k]
        pos[k] = cell
    return cells, pos, rng, np.float32(np.float32(happy.sum()) / np.float32(na)), len(claim)


@pytest.mark.parametrize("mode", MODES)
def test_schelling_matches_dense_step(mode):
    H, W = 24, 20
    cells, pos, typ, rng = O.schelling_init(H, W, 3, mode)
    u = O.uniform(O.prng_key(3), H * W, mode=mode)
    np.testing.assert_array_equal(cells, np.where(u < np.float32(0.45), 1, np.where(u < np.float32(0.9), 2, 0)))
    rc, rp, rr = cells.copy(), pos.copy(), rng.copy()
    for _ in range(6):
        rc, rp, rr, hf, mv = dense_schelling_step(H, W, rc, rp, typ, rr, mode)
        h, m = O.schelling_run(H, W, cells, pos, typ, rng, 1, mode)
        np.testing.assert_array_equal(cells, rc)
        np.testing.assert_array_equal(pos, rp)
        assert h[0] == hf and m[0] == mv


def test_schelling_invariants():
    H = W = 64
    cells, pos, typ, rng = O.schelling_init(H, W, 0, O.MODE_PARTITIONABLE)
    n1, n2 = (cells == 1).sum(), (cells == 2).sum()
    hf, mv = O.schelling_run(H, W, cells, pos, typ, rng, 50, O.MODE_PARTITIONABLE)
    assert (cells == 1).sum() == n1 and (cells == 2).sum() == n2
    assert len(np.unique(pos)) == len(pos)
    np.testing.assert_array_equal(cells[pos], typ)
    assert hf[-1] > hf[0] and (mv >= 0).all()


# ---------------------------------------------------------------- Deffuant
def test_deffuant_dense_step_and_conservation():
    n, B = 1000, 3
    x0, rng = O.deffuant_init(n, B, 1, O.MODE_PARTITIONABLE)
    eps = np.array([0.1, 0.2, 0.5], np.float32)
    mu = np.array([0.3, 0.5, 0.1], np.float32)
    x = x0.copy()
    r2 = rng.copy()
    O.deffuant_forward(x, eps, mu, r2, 1, O.MODE_PARTITIONABLE)
    _, ck, uk = step_keys(rng, O.MODE_PARTITIONABLE)
    perm = O.permutation(uk, n, O.MODE_PARTITIONABLE)
    a, b = perm[0::2], perm[1::2]
    d = x0[b] - x0[a]
    g = 1.0 / (1.0 + np.exp(-((eps - np.abs(d)) / np.float32(0.02)).astype(np.float32)))
    delta = (mu * g.astype(np.float32)) * d
    want = x0.copy()
    want[a] = x0[a] + delta
    want[b] = x0[b] - delta
    np.testing.assert_allclose(x, want, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(x.sum(0), x0.sum(0), rtol=1e-5)


@pytest.mark.parametrize("mode", [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE])
@pytest.mark.parametrize("n,seed", [(1000, 0), (7, 3), (1, 5)])
def test_deffuant_init_is_uniform_of_split_agent_keys(mode, n, seed):
    """orc_deffuant_init composes pinned primitives as the reference does:
    JaxModel.initialize splits PRNGKey(seed) into (rng, collection key)
    (jaxabm/model.py:118-144), AgentCollection.init splits the collection key
    into N agent keys (jaxabm/agent.py:92-130) and each opinion is
    uniform(agent key); every instance column starts from the same opinions."""
    x, rng = O.deffuant_init(n, 3, seed, mode)
    rk, ck = O.split(O.prng_key(seed), 2, mode)
    ak = O.split(ck, n, mode)
    want = np.array([O.uniform(k, 1, mode=mode)[0] for k in ak], np.float32)
    for b in range(3):
        np.testing.assert_array_equal(x[:, b], want)
    np.testing.assert_array_equal(rng, rk)


def _expit(z):
    return 0.5 * (1.0 + np.tanh(0.5 * z))   # logistic function, numerically stable


def test_deffuant_stats_against_numpy_and_closed_forms():
    """orc_deffuant_stats = [mean, population variance, 32-bin soft histogram
    with slope 128]: mean and variance against numpy; the histogram against
    its telescoping sum, the reflection x -> 1 - x (bins reversed), a single
    agent at a bin centre (closed form tanh(1) in its bin), and a shift of
    all opinions by exactly one bin width (bins move by one)."""
    rs = np.random.default_rng(5)
    x = rs.uniform(0, 1, (4001, 3)).astype(np.float32)
    st = O.deffuant_stats(x)
    x64 = x.astype(np.float64)
    np.testing.assert_allclose(st[:, 0], x64.mean(0), rtol=1e-13)
    np.testing.assert_allclose(st[:, 1], x64.var(0), rtol=1e-11)
    h = st[:, 2:]
    assert h.shape == (3, 32) and (h >= 0).all()
    # sum_k [s(128(x - k/32)) - s(128(x - (k+1)/32))] telescopes
    tel = (_expit(128.0 * x64) - _expit(128.0 * (x64 - 1.0))).mean(0)
    np.testing.assert_allclose(h.sum(1), tel, rtol=1e-12)
    # reflection: bin k of 1 - x is bin 31 - k of x (x and 1 - x exact in f32
    # only for dyadic values: use multiples of 2^-12)
    q = (rs.integers(0, 4097, (2048, 1)) / 4096.0).astype(np.float32)
    hq = O.deffuant_stats(q)[0, 2:]
    hr = O.deffuant_stats((1.0 - q).astype(np.float32))[0, 2:]
    np.testing.assert_allclose(hr, hq[::-1], rtol=1e-12, atol=1e-15)
    # one agent at the centre of bin k: sigma(2) - sigma(-2) = tanh(1)
    for k in (0, 13, 31):
        c = np.array([[(k + 0.5) / 32.0]], np.float32)
        hk = O.deffuant_stats(c)[0]
        assert hk[0] == np.float64(c[0, 0]) and hk[1] == 0.0
        assert abs(hk[2 + k] - np.tanh(1.0)) < 1e-14
        if k > 0:   # neighbouring bin: sigma(6) - sigma(2)
            assert abs(hk[2 + k - 1] - (_expit(6.0) - _expit(2.0))) < 1e-14
    # a shift by one bin width (dyadic, exact) moves every bin by one
    base = (rs.integers(512, 3584, (1500, 1)) / 4096.0).astype(np.float32)
    h0 = O.deffuant_stats(base)[0, 2:]
    h1 = O.deffuant_stats((base + np.float32(1.0 / 32)).astype(np.float32))[0, 2:]
    np.testing.assert_allclose(h1[1:], h0[:-1], rtol=1e-12, atol=1e-15)


def test_deffuant_adjoint_matches_finite_differences():
    n, steps = 512, 40
    x0, rng0 = O.deffuant_init(n, 1, 0, O.MODE_PARTITIONABLE)
    xt = x0.copy()
    O.deffuant_forward(xt, np.float32([0.2]), np.float32([0.3]), rng0.copy(), steps,
                       O.MODE_PARTITIONABLE)
    target = O.deffuant_stats(xt)[0]
    eps, mu = np.array([0.27]), np.array([0.21])
    loss, ge, gm = O.deffuant_grad(x0, eps, mu, rng0, steps, target, O.MODE_PARTITIONABLE)
    def fd(h):
        lp = O.deffuant_grad(x0, eps + h, mu, rng0, steps, target, O.MODE_PARTITIONABLE)[0]
        lm = O.deffuant_grad(x0, eps - h, mu, rng0, steps, target, O.MODE_PARTITIONABLE)[0]
        mp = O.deffuant_grad(x0, eps, mu + h, rng0, steps, target, O.MODE_PARTITIONABLE)[0]
        mm = O.deffuant_grad(x0, eps, mu - h, rng0, steps, target, O.MODE_PARTITIONABLE)[0]
        return (lp - lm) / (2 * h), (mp - mm) / (2 * h)

    assert loss[0] > 0
    fd_e, fd_m = fd(1e-6)
    np.testing.assert_allclose(ge, fd_e, rtol=1e-5)
    np.testing.assert_allclose(gm, fd_m, rtol=1e-5)
    # the central difference converges to the adjoint as h shrinks
    errs = [max(abs(fd(h)[0][0] - ge[0]), abs(fd(h)[1][0] - gm[0])) for h in (1e-3, 1e-4, 1e-5)]
    assert errs[2] < errs[1] < errs[0]
    # at the target parameters the loss (same seed, same pairing) is ~0
    l0 = O.deffuant_grad(x0, np.array([0.2]), np.array([0.3]), rng0, steps, target,
                         O.MODE_PARTITIONABLE)[0]
    assert l0[0] < 1e-10


def test_sir_fixture_is_the_oracle_run():
    """tests/golden/sir_ba1M_m4_365steps.json (written by make_sir_ba1M.py,
    asserted by bench.py) is the oracle's 365-step run on BA(1M, m=4) in both
    stream modes, and satisfies S + I + R = N, S non-increasing, R
    non-decreasing."""
    import json
    import os

    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "sir_ba1M_m4_365steps.json")))
    n = gold["n"]
    rp, col = O.ba_csr(n, gold["m"], gold["graph_seed"])
    for mode, name in ((O.MODE_PARTITIONABLE, "partitionable"), (O.MODE_ORIGINAL, "original")):
        want = np.array(gold["counts"][name], np.int32)
        assert want.shape == (365, 3)
        assert (want.sum(1) == n).all()
        assert (np.diff(want[:, 0]) <= 0).all() and (np.diff(want[:, 2]) >= 0).all()
        st, rng = O.sir_init(n, gold["seed_infected"], gold["seed"], mode)
        got = O.sir_run_threads(rp, col, st, rng, np.float32(gold["beta"]),
                                np.float32(gold["gamma"]), gold["steps"], os.cpu_count() or 1, mode)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("mode", [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE])
@pytest.mark.parametrize("n,steps", [(1000, 30), (3001, 12), (2, 5)])
def test_wealth_threaded_run_is_identical(mode, n, steps):
    """The threaded wealth oracle (per-element key and draw forms on several
    threads) gives the serial run's records and final state."""
    x1, y1, w1, r1 = O.wealth_init(n, 4, mode)
    x2, y2, w2, r2 = x1.copy(), y1.copy(), w1.copy(), r1.copy()
    m1, g1 = O.wealth_run(x1, y1, w1, r1, steps, mode)
    m2, g2 = O.wealth_run_threads(x2, y2, w2, r2, steps, 4, mode)
    for a, b in ((m1, m2), (g1, g2), (x1, x2), (y1, y2), (w1, w2), (r1, r2)):
        np.testing.assert_array_equal(a, b)


def test_sir_threaded_run_is_identical():
    """The all-cores CPU baseline computes the same simulation: each agent's
    next state depends only on the pre-step copy."""
    rp, col = O.ba_csr(20000, 4, 3)
    st1, r1 = O.sir_init(20000, 50, 2, O.MODE_PARTITIONABLE)
    st2, r2 = st1.copy(), r1.copy()
    c1 = O.sir_run(rp, col, st1, r1, np.float32(0.07), np.float32(0.1), 60, O.MODE_PARTITIONABLE)
    c2 = O.sir_run_threads(rp, col, st2, r2, np.float32(0.07), np.float32(0.1), 60, 4,
                           O.MODE_PARTITIONABLE)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(st1, st2)
    np.testing.assert_array_equal(r1, r2)


def test_deffuant_grad_threads_same_bits_as_one_thread():
    """The all-cores Deffuant gradient (instances over OpenMP threads, each on
    its own column arrays) gives the single-threaded oracle's loss and
    gradients bit for bit."""
    n, batch, steps = 3000, 7, 6
    x0, rng = O.deffuant_init(n, batch, 4, O.MODE_PARTITIONABLE)
    eps = np.linspace(0.05, 0.5, batch)
    mu = np.linspace(0.1, 0.5, batch)[::-1].copy()
    target = np.linspace(0.0, 0.3, O.STATS_LEN)
    a = O.deffuant_grad(x0, eps, mu, rng, steps, target)
    for threads in (1, 3, 8):
        b = O.deffuant_grad_threads(x0, eps, mu, rng, steps, target, threads)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u.view(np.uint64), v.view(np.uint64))


@pytest.mark.parametrize("mode", [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE])
def test_schelling_run_threads_same_results(mode):
    """The all-cores Schelling oracle (agent phase on OpenMP threads, keys per
    agent) gives the single-threaded run's records, grid and positions."""
    H, W, steps = 96, 80, 12
    a = O.schelling_init(H, W, 3, mode)
    b = tuple(np.array(v, copy=True) for v in a)
    hf1, mv1 = O.schelling_run(H, W, a[0], a[1], a[2], a[3], steps, mode)
    hf2, mv2 = O.schelling_run_threads(H, W, b[0], b[1], b[2], b[3], steps, 5, mode)
    np.testing.assert_array_equal(hf1, hf2)
    np.testing.assert_array_equal(mv1, mv2)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[3], b[3])
    assert mv1.sum() > 0
