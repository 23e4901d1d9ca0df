"""GPU parity: differentiable Deffuant model (forward, statistics, adjoint)
against the oracle.

Bars: pairings bit-exact; float32 opinions within 1e-5 relative (the gate
calls the platform expf, which may differ from the oracle's by an ulp);
statistics and losses of a given state to float64 rounding; gradients
against the oracle's float64 restatement of jax.grad through the whole run
(tolerance below, with its reason) and against central differences."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MODES = [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE]
PART = O.MODE_PARTITIONABLE


@pytest.fixture(scope="module")
def dm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd.programs import deffuant

    return deffuant


def oracle_pairings(rng0, steps, n, mode):
    r = rng0.copy()
    out = []
    for _ in range(steps):
        r, sk = O.split(r, 2, mode)
        sk, _ck = O.split(sk, 2, mode)
        sk, uk = O.split(sk, 2, mode)
        out.append(O.permutation(uk, n, mode))
    return np.array(out)


def oracle_target(n, steps, seed, mode=PART):
    x0, rng0 = O.deffuant_init(n, 1, seed, mode)
    xt = x0.copy()
    O.deffuant_forward(xt, np.float32([0.2]), np.float32([0.3]), rng0.copy(), steps, mode)
    return O.deffuant_stats(xt)[0]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,steps", [(1000, 5), (1001, 3), (65536, 2), (7, 4)])
def test_pairings_match_oracle(dm, mode, n, steps):
    m = dm.DeffuantModel(n, 1, steps, seed=3, mode=mode)
    want = oracle_pairings(m.rng0, steps, n, mode)
    np.testing.assert_array_equal(m.perms[:steps].cpu().numpy(), want)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,B,steps,k", [(1000, 8, 30, 7), (1001, 3, 25, 25), (4096, 256, 12, 5),
                                         (2, 4, 3, 1), (999, 1, 10, 3)])
def test_forward_matches_oracle(dm, mode, n, B, steps, k):
    rs = np.random.default_rng(n + B)
    eps = rs.uniform(0.05, 0.5, B).astype(np.float32)
    mu = rs.uniform(0.05, 0.5, B).astype(np.float32)
    target = rs.uniform(0.0, 0.1, 34)
    m = dm.DeffuantModel(n, B, steps, seed=1, ckpt_every=k, mode=mode)
    x0, rng = O.deffuant_init(n, B, 1, mode)
    np.testing.assert_array_equal(m.x0.cpu().numpy(), x0)
    loss = m.forward(eps, mu, target).cpu().numpy()
    x = x0.copy()
    O.deffuant_forward(x, eps, mu, rng, steps, mode)
    xf = m.final_state().cpu().numpy()
    # the gate's exp is the shared abm_math.h function on both sides, and
    # every other operation is one IEEE-rounded float32 operation: same bits
    np.testing.assert_array_equal(xf, x)
    # statistics and loss of the GPU's own final state: float64 arithmetic
    st = O.deffuant_stats(xf)
    np.testing.assert_allclose(m.stats.cpu().numpy(), st, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(loss, [O.deffuant_loss(s, target) for s in st], rtol=1e-10,
                               atol=1e-15)
    # pairwise updates conserve each instance's opinion sum up to rounding
    np.testing.assert_allclose(xf.astype(np.float64).sum(0), x0.astype(np.float64).sum(0),
                               rtol=1e-5)


def grads_close(g, want, rtol):
    """|g - want| <= rtol (|want| + max|want| / 100): relative where the
    gradient is sizeable, scaled by the batch's largest gradient where it is
    near zero (flat loss in the consensus regime)."""
    scale = np.abs(want).max()
    return np.all(np.abs(g - want) <= rtol * (np.abs(want) + scale / 100))


def test_adjoint_matches_restated_jax_grad(dm):
    """Float32 adjoint on the GPU against the float64 reverse sweep of the
    same program.  The float32 trajectory differs from the float64 one by
    rounding, which the sensitivity of 60 gated steps amplifies; 1e-3 holds
    with margin (the float32-vs-float64 gap measured on the CPU is ~1e-6)."""
    n, B, steps = 2048, 8, 60
    target = oracle_target(n, steps, 0)
    eps = np.linspace(0.05, 0.45, B).astype(np.float32)
    mu = np.linspace(0.45, 0.1, B).astype(np.float32)
    m = dm.DeffuantModel(n, B, steps, seed=0, ckpt_every=20)
    loss = m.forward(eps, mu, target).cpu().numpy()
    ge, gm = (g.cpu().numpy() for g in m.backward())
    x0, rng0 = O.deffuant_init(n, B, 0, PART)
    ol, oge, ogm = O.deffuant_grad(x0, eps.astype(np.float64), mu.astype(np.float64), rng0,
                                   steps, target, PART)
    print("loss", loss, ol)
    print("grad eps", ge, oge)
    print("grad mu", gm, ogm)
    np.testing.assert_allclose(loss, ol, rtol=1e-3, atol=1e-9)
    assert grads_close(ge, oge, 1e-3)
    assert grads_close(gm, ogm, 1e-3)


def test_checkpoint_interval_does_not_change_gradients(dm):
    """Recomputation replays the forward kernels bit for bit, so every
    checkpoint interval gives identical gradients; so does the forward's tape
    of the last segment (first backward) against recomputing it (a second
    backward of the same forward).  Last segments of 23 % k steps: 3, 2, 1
    (k = 2, 11), whole (k = 1, 23) and the whole run (k = 50)."""
    n, B, steps = 1500, 4, 23
    target = oracle_target(n, steps, 4)
    eps = np.float32([0.1, 0.2, 0.3, 0.15])
    mu = np.float32([0.3, 0.2, 0.1, 0.45])
    ref = None
    for k in (1, 2, 4, 7, 11, 23, 50):
        m = dm.DeffuantModel(n, B, steps, seed=4, ckpt_every=k)
        loss = m.forward(eps, mu, target).cpu().numpy()
        g = np.stack([x.cpu().numpy() for x in m.backward()])
        g2 = np.stack([x.cpu().numpy() for x in m.backward()])
        np.testing.assert_array_equal(g2, g)
        if ref is None:
            ref = (loss, g)
        else:
            np.testing.assert_array_equal(loss, ref[0])
            np.testing.assert_array_equal(g, ref[1])


def test_central_differences(dm):
    """The batch evaluates the stencil (eps +- h, mu +- h) in one launch;
    h = 1e-3.  The float32 program's differences match the oracle's float64
    program's at the same (float32-rounded) parameters, and the adjoint lies
    between the h = 1e-3 secant and its h -> 0 limit the oracle's own tests
    establish."""
    n, steps, h = 2048, 40, 1e-3
    target = oracle_target(n, steps, 0)
    e0, m0 = 0.12, 0.25
    eps = np.float32([e0 + h, e0 - h, e0, e0, e0])
    mu = np.float32([m0, m0, m0 + h, m0 - h, m0])
    m = dm.DeffuantModel(n, 5, steps, seed=0, ckpt_every=10)
    L = m.forward(eps, mu, target).cpu().numpy()
    ge, gm = (g.cpu().numpy()[4] for g in m.backward())
    fd_e = (L[0] - L[1]) / (float(eps[0]) - float(eps[1]))
    fd_m = (L[2] - L[3]) / (float(mu[2]) - float(mu[3]))
    x0, rng0 = O.deffuant_init(n, 5, 0, PART)
    OL, oge, ogm = O.deffuant_grad(x0, eps.astype(np.float64), mu.astype(np.float64), rng0, steps,
                                   target, PART)
    ofd_e = (OL[0] - OL[1]) / (float(eps[0]) - float(eps[1]))
    ofd_m = (OL[2] - OL[3]) / (float(mu[2]) - float(mu[3]))
    print("fd", fd_e, fd_m, "oracle fd", ofd_e, ofd_m, "adjoint", ge, gm, "oracle", oge[4], ogm[4])
    np.testing.assert_allclose([fd_e, fd_m], [ofd_e, ofd_m], rtol=2e-3)
    np.testing.assert_allclose([ge, gm], [oge[4], ogm[4]], rtol=1e-3)
    # secant and tangent agree to the curvature term (same sign, within 25%)
    assert np.sign(fd_e) == np.sign(ge) and abs(fd_e - ge) <= 0.25 * abs(ge)
    assert np.sign(fd_m) == np.sign(gm) and abs(fd_m - gm) <= 0.25 * abs(gm)


def test_autograd_function(dm):
    import torch

    n, B, steps = 1024, 4, 20
    target = oracle_target(n, steps, 2)
    m = dm.DeffuantModel(n, B, steps, seed=2, ckpt_every=6)
    eps = torch.tensor([0.1, 0.2, 0.3, 0.4], device="cuda", requires_grad=True)
    mu = torch.tensor([0.2, 0.3, 0.4, 0.1], device="cuda", requires_grad=True)
    loss = dm.deffuant_loss(eps, mu, m, target)
    w = torch.tensor([1.0, 2.0, 0.5, -1.0], dtype=torch.float64, device="cuda")
    (loss * w).sum().backward()
    ge, gm = m.backward()
    assert torch.equal(eps.grad, (w * ge).float())
    assert torch.equal(mu.grad, (w * gm).float())
    # a backward after another forward on the same model is refused
    l1 = dm.deffuant_loss(eps, mu, m, target)
    dm.deffuant_loss(eps, mu, m, target)
    with pytest.raises(RuntimeError, match="another forward"):
        l1.sum().backward()


def test_full_size_batch(dm):
    """65,536 agents, 256 (eps, mu) instances, 200 steps, checkpoints every 20
    (the measured workload): opinion sums conserved, the instance at the
    target parameters reproduces the target exactly, four instances' final
    opinions and two instances' gradients match the oracle."""
    n, B, steps = 65536, 256, 200
    target = dm.target_stats(n, steps, seed=0)
    ge_grid = np.linspace(0.05, 0.5, 16)
    mu_grid = np.linspace(0.05, 0.5, 16)
    eps = np.repeat(ge_grid, 16).astype(np.float32)
    mu = np.tile(mu_grid, 16).astype(np.float32)
    eps[37], mu[37] = 0.2, 0.3
    m = dm.DeffuantModel(n, B, steps, seed=0, ckpt_every=20)
    loss = m.forward(eps, mu, target).cpu().numpy()
    ge, gm = (g.cpu().numpy() for g in m.backward())
    assert loss[37] == 0.0
    x0 = m.x0.cpu().numpy().astype(np.float64)
    xf = m.final_state().cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(xf.sum(0), x0.sum(0), rtol=1e-5)
    assert np.isfinite(ge).all() and np.isfinite(gm).all()
    # the measured batch's final opinions: four instances (the target one,
    # the grid's corners) through the whole 200 steps against the oracle's
    # forward (the same opinions start every instance)
    fpick = [0, 15, 37, 255]
    x0f, rngf = O.deffuant_init(n, len(fpick), 0, PART)
    np.testing.assert_array_equal(m.x0.cpu().numpy()[:, fpick], x0f)
    O.deffuant_forward(x0f, eps[fpick], mu[fpick], rngf, steps, PART)
    # every opinion after the 200 steps is the oracle's, bit for bit (the
    # gate's exp is the shared abm_math.h function on both sides)
    xf4 = m.final_state().cpu().numpy()[:, fpick]
    np.testing.assert_array_equal(xf4, x0f)
    pick = [20, 200]
    x0o, rng0 = O.deffuant_init(n, len(pick), 0, PART)
    ol, oge, ogm = O.deffuant_grad(x0o, eps[pick].astype(np.float64), mu[pick].astype(np.float64),
                                   rng0, steps, target, PART)
    print("picked", eps[pick], mu[pick], loss[pick], ol, ge[pick], oge, gm[pick], ogm)
    np.testing.assert_allclose(loss[pick], ol, rtol=1e-3, atol=1e-9)
    assert grads_close(ge[pick], oge, 2e-3) and grads_close(gm[pick], ogm, 2e-3)


def test_batch_size_does_not_change_an_instance(dm):
    """An instance's loss and gradients are the same bits in a batch of 41
    (step and adjoint kernels on single columns) and of 96 (on 4-column
    vectors): instances are independent columns of every kernel."""
    n, steps, k = 4096, 30, 7
    rs = np.random.default_rng(9)
    eps = rs.uniform(0.05, 0.5, 96).astype(np.float32)
    mu = rs.uniform(0.05, 0.5, 96).astype(np.float32)
    target = rs.uniform(0.0, 0.1, 34)
    out = []
    for B in (41, 96):
        m = dm.DeffuantModel(n, B, steps, seed=2, ckpt_every=k)
        loss = m.forward(eps[:B], mu[:B], target).cpu().numpy()
        ge, gm = (g.cpu().numpy() for g in m.backward())
        out.append((loss[:41], ge[:41], gm[:41]))
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
