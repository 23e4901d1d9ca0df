"""GPU: record-time reductions (jaxabm_amd.reductions) against the oracle --
Gini and the int32 mean bit for bit (the reference's int32 / float32
formulas), counts exactly, the float32 mean to its correctly rounded value
within one ulp, the soft histogram at 1e-10 -- and inside a step replayed as
a graph (no memset nodes)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd import reductions

    return reductions


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 100_003])
def test_gini_and_int_mean_match_oracle(R, n):
    rs = np.random.default_rng(n)
    cases = [rs.integers(0, 50, n), rs.integers(-20, 20, n), np.zeros(n),
             rs.integers(0, 2**30, n),        # sums wrap around int32
             np.ones(n)]
    for v in cases:
        v = v.astype(np.int32)
        assert R.gini(_dev(v)).item() == O.gini_i32(v), v[:8]
        assert R.mean(_dev(v)).item() == O.mean_i32(v)


def test_float_mean_correctly_rounded(R):
    rs = np.random.default_rng(3)
    for n in (1, 33, 65_536, 1_000_003):
        x = (rs.standard_normal(n) * 10.0 ** rs.integers(-3, 4, n)).astype(np.float32)
        got = np.float32(R.mean(_dev(x)).item())
        exact = np.float32(x.astype(np.float64).sum() / n)
        assert got == exact or abs(int(got.view(np.int32)) - int(exact.view(np.int32))) <= 1


@pytest.mark.parametrize("n,k", [(0, 3), (1, 1), (1000, 3), (1_000_000, 3), (4097, 256)])
def test_state_counts(R, n, k):
    rs = np.random.default_rng(n + k)
    s = rs.integers(0, min(k + 2, 256), n).astype(np.uint8)   # values >= k are not counted
    got = R.state_counts(_dev(s), k).cpu().numpy()
    want = np.bincount(s, minlength=256)[:k]
    np.testing.assert_array_equal(got, want)


def test_soft_histogram_matches_deffuant_statistics(R):
    rs = np.random.default_rng(5)
    for n in (1, 1000, 65_536):
        x = rs.random(n).astype(np.float32)
        got = R.soft_histogram(_dev(x), 32, 128.0).cpu().numpy()
        want = O.deffuant_stats(x.reshape(n, 1))[0, 2:34]
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-14)
        assert abs(got.sum() - want.sum()) < 1e-12


def test_reductions_in_a_replayed_step(R):
    """A generic model whose metrics use these reductions replays as a graph
    with no memset node and the eager loop's results."""
    import torch

    import jaxabm_amd as J
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    class Giver:
        def init_state(self, cfg, key):
            return {"w": jr.randint(key, (), 0, 5), "s": jr.randint(key, (), 0, 3)}

        def update(self, state, ms, cfg, key):
            k1, k2 = jr.split(key)
            w = jnp.where(jr.uniform(k1) < 0.5, state["w"] + 1, jnp.maximum(state["w"] - 1, 0))
            return {"w": w, "s": jr.randint(k2, (), 0, 3)}

    def metrics(env, agents, params):
        a = agents["g"]
        return {"gini": R.gini(a["w"]), "mean": R.mean(a["w"]),
                "counts": R.state_counts(a["s"].to(torch.uint8), 3)}

    def make():
        m = J.JaxModel(config=J.ModelConfig(seed=2), metrics_fn=metrics)
        m.add_agent_collection("g", J.AgentCollection(Giver(), 50_000))
        return m

    a, b = make(), make()
    ra, rb = a.run(12), b.run(12, graph=True)
    info = b.last_run_info
    assert info["mode"] == "graph" and info["memsets_replaced"] == 0, info
    for k in ("gini", "mean", "counts"):
        for i, (x, y) in enumerate(zip(ra[k], rb[k])):
            np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy(), err_msg=f"{k} {i}")
    w = a.agent_collections["g"].states["w"].cpu().numpy()
    assert ra["gini"][-1].item() == O.gini_i32(w)


@pytest.mark.parametrize("n", [1, 7, 10_000, 65_536, 65_537, 1_000_003])
def test_whole_array_sums_and_means(R, n):
    """jnp.sum / jnp.mean / jnp.count_nonzero of a plain GPU tensor run the
    HIP kernels (one launch up to 65,536 values, block partials above):
    int32 sums wrap exactly as JAX's x64-off int32, boolean sums count,
    float32 sums and means are the float64 sum rounded once (within an ulp of
    numpy's float64 pairwise sum, rounded)."""
    import torch

    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import reductions

    rs = np.random.default_rng(n)
    xi = rs.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    ti = torch.from_numpy(xi).cuda()
    assert jnp.sum(ti).dtype == torch.int32
    wrapped = (int(xi.astype(np.int64).sum()) + 2**31) % 2**32 - 2**31
    assert int(jnp.sum(ti)) == wrapped
    xb = rs.random(n) < 0.3
    tb = torch.from_numpy(xb).cuda()
    assert int(jnp.sum(tb)) == int(xb.sum()) == int(jnp.count_nonzero(tb))
    assert int(jnp.count_nonzero(ti)) == int((xi != 0).sum())
    xf = (rs.standard_normal(n) * 100).astype(np.float32)
    tf = torch.from_numpy(xf).cuda()
    s = float(jnp.sum(tf))
    want = np.float32(xf.astype(np.float64).sum())
    assert abs(np.float32(s) - want) <= np.spacing(np.abs(want)) + 1e-30
    m = float(jnp.mean(tf))
    wm = np.float32(xf.astype(np.float64).sum() / n)
    assert abs(np.float32(m) - wm) <= np.spacing(np.abs(wm)) + 1e-30
    small = rs.integers(0, 1000, n).astype(np.int32)
    np.testing.assert_equal(float(jnp.mean(torch.from_numpy(small).cuda())),
                            float(np.float32(small.astype(np.int64).sum()) / np.float32(n)))
    # the one-launch and multi-block paths are the module's entry points too
    assert reductions.sum_all(ti).item() == jnp.sum(ti).item()


def test_reduce_batch_equals_single_launches(R):
    """abm_reduce_batch through the C ABI: every op at 1, 1,000 and 65,536
    values, 15 items in one launch, each result bit-identical to its own
    single-reduction call; bad items are refused."""
    import ctypes as C

    import torch

    from jaxabm_amd._lib import (REDUCE_MEAN_F32, REDUCE_MEAN_I32, REDUCE_SUM_F32,
                                 REDUCE_SUM_I32, REDUCE_SUM_U8, AbmError, ReduceItem, check, lib)

    rs = np.random.default_rng(3)
    single = {REDUCE_SUM_F32: ("abm_sum_f32", torch.float32),
              REDUCE_MEAN_F32: ("abm_mean_f32", torch.float32),
              REDUCE_SUM_I32: ("abm_sum_i32", torch.int32),
              REDUCE_MEAN_I32: ("abm_mean_i32", torch.float32),
              REDUCE_SUM_U8: ("abm_sum_u8", torch.int32)}
    items, keep, want = [], [], []
    for op, (name, odt) in single.items():
        for n in (1, 1000, 65_536):
            if op in (REDUCE_SUM_F32, REDUCE_MEAN_F32):
                x = torch.from_numpy((rs.standard_normal(n) * 2.0 ** rs.integers(-20, 20, n))
                                     .astype(np.float32)).cuda()
            elif op == REDUCE_SUM_U8:
                x = torch.from_numpy(rs.integers(0, 256, n).astype(np.uint8)).cuda()
            else:
                x = torch.from_numpy(rs.integers(-2**31, 2**31 - 1, n).astype(np.int32)).cuda()
            out = torch.full((), -7, dtype=odt, device="cuda")
            ref = torch.full((), -9, dtype=odt, device="cuda")
            check(getattr(lib(), name)(x.data_ptr(), n, ref.data_ptr(), None, 0, None), name)
            items.append(ReduceItem(x.data_ptr(), n, out.data_ptr(), op, 0))
            keep.append((x, out))
            want.append(ref)
    arr = (ReduceItem * len(items))(*items)
    check(lib().abm_reduce_batch(arr, len(items), None), "reduce_batch")
    torch.cuda.synchronize()
    for (x, out), ref in zip(keep, want):
        assert out.view(torch.int32).item() == ref.view(torch.int32).item(), (x.dtype, x.numel())
    bad = (ReduceItem * 1)(ReduceItem(keep[0][0].data_ptr(), 65_537, keep[0][1].data_ptr(),
                                      REDUCE_SUM_F32, 0))
    with pytest.raises(AbmError, match="bad item"):
        check(lib().abm_reduce_batch(bad, 1, None), "reduce_batch")
    with pytest.raises(AbmError, match="count"):
        check(lib().abm_reduce_batch(arr, 17, None), "reduce_batch")


def test_deferred_callback_reductions(R):
    """Inside reductions.deferring() (a step's env and metrics callbacks)
    whole-array sums and means wait and go out as one launch: reading a
    result launches the waiting batch first; values equal the immediate
    launches bit for bit; more than 16 waiting reductions take several
    launches; materialize() hands back plain tensors."""
    import torch

    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import reductions

    g = torch.Generator(device="cuda").manual_seed(5)
    xs = [torch.randn(1000 + 37 * i, device="cuda", generator=g) * 1e3 for i in range(20)]
    ws = torch.randint(-1000, 1000, (4096,), device="cuda", generator=g, dtype=torch.int32)
    now = [jnp.mean(x) for x in xs] + [jnp.sum(ws), jnp.sum(ws > 3), jnp.mean(ws)]
    with reductions.deferring():
        later = [jnp.mean(x) for x in xs] + [jnp.sum(ws), jnp.sum(ws > 3), jnp.mean(ws)]
        assert all(isinstance(v, reductions._Deferred) for v in later)
        assert later[0].dtype == torch.float32 and later[-3].dtype == torch.int32
        assert len(reductions._pending()) == 23        # metadata reads launch nothing
        used = later[1] * 2.0 + 1.0                     # a read launches the batch
        assert not reductions._pending() and type(used) is torch.Tensor
        tail = jnp.sum(xs[0])
        assert len(reductions._pending()) == 1
        d = reductions.materialize({"a": later[0], "b": 3, "c": tail})
        assert type(d["a"]) is torch.Tensor and d["b"] == 3 and not reductions._pending()
    assert not reductions._pending()
    for a, b in zip(now, later):
        assert a.view(torch.int32).item() == b._abm_plain.view(torch.int32).item()
    assert used.item() == (now[1] * 2.0 + 1.0).item()
    assert d["c"].item() == jnp.sum(xs[0]).item()
    # an input changed in place before the batch goes out is an error, not a
    # silently different sum
    y = xs[3].clone()
    with pytest.raises(RuntimeError, match="modified in place"):
        with reductions.deferring():
            jnp.sum(y)
            y.add_(1.0)
    assert not reductions._pending()


def test_reduction_routing_rules(R):
    """Axis reductions, keepdims, float64 and tensors that record gradients
    stay torch reductions (with JAX's x64-off integer dtypes)."""
    import torch

    from jaxabm_amd import numpy as jnp

    x = torch.arange(12, dtype=torch.int32, device="cuda").reshape(3, 4)
    assert jnp.sum(x, axis=0).tolist() == [12, 15, 18, 21]
    assert jnp.sum(x, axis=(0, 1)).item() == 66 and jnp.sum(x, axis=(0, 1)).dim() == 0
    assert jnp.sum(x, keepdims=True).shape == (1, 1)
    assert jnp.sum(x.double()).dtype == torch.float64
    g = torch.ones(5, device="cuda", requires_grad=True)
    jnp.mean(g * 2).backward()
    assert torch.allclose(g.grad, torch.full((5,), 0.4, device="cuda"))
