"""The learned rank predictor on the GPU: its hand-written forward, backward
and Adam kernels against the CPU FP64 model (oracle/mlp.py, itself pinned by
finite differences), and assembly with learned ACA capacities against the
plain assembly (identical ranks and results, fallbacks counted)."""
import numpy as np
import pytest

from conftest import mixture_points, uniform_points
from oracle import mlp

pytestmark = pytest.mark.gpu


def print_case(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def random_batch(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 16)).astype(np.float32)
    y = rng.uniform(0, 6, size=n).astype(np.float32)
    return X, y


def test_forward_matches_cpu(engine):
    m = engine.RankModel(seed=1)
    p = m.params()
    assert p.shape == (mlp.N_PARAMS,) and np.all(p[1024:1088] == 0)
    X, _ = random_batch(2, 1000)
    out = m.predict(X)
    ref, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.abs(out - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    m.close()


@pytest.mark.parametrize("n", [1, 300, 777])
def test_gradient_matches_cpu_fp64(engine, n):
    """FP32 gradient kernels against the FP64 analytic gradient at the same
    parameters, and on sampled parameters against FP64 central differences
    of the CPU loss."""
    print_case("gradient", n)
    m = engine.RankModel(seed=3)
    p = m.params()
    rng = np.random.default_rng(4)
    p[1024:1088] = rng.uniform(-0.1, 0.1, 64)     # nonzero biases
    p[5184:5248] = rng.uniform(-0.1, 0.1, 64)
    p[9472] = 0.3
    m.set_params(p)
    X, y = random_batch(5 + n, n)
    loss, g = m.gradient(X, y)
    p64 = p.astype(np.float64)
    loss_ref, g_ref = mlp.loss_and_grad(p64, X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 1e-4 * scale, np.abs(g - g_ref).max() / scale
    assert rel(g.astype(np.float64), g_ref) <= 1e-5
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(6, b - a), replace=False)
                          for a, b in ((0, 1024), (1024, 1088), (1088, 5184), (5248, 9344),
                                       (9408, 9472), (9472, 9473))])
    fd = mlp.finite_difference_grad(p64, X.astype(np.float64), y.astype(np.float64), idx)
    assert np.abs(g[idx] - fd).max() <= 1e-4 * scale
    m.close()


def test_adam_step_matches_cpu(engine):
    m = engine.RankModel(seed=6)
    p0 = m.params()
    X, y = random_batch(7, 512)
    _, g = m.gradient(X, y)
    lr = 1e-3
    m.train(X, y, epochs=1, batch=512, lr=lr, seed=1)  # one full-batch step
    p1 = m.params()
    z = np.zeros(mlp.N_PARAMS)
    q1, _, _ = mlp.adam_step(p0.astype(np.float64), g.astype(np.float64), z, z, 1, lr)
    d_gpu = p1.astype(np.float64) - p0
    d_cpu = q1 - p0
    # the first step moves every parameter by about lr sign(g); where g is
    # tiny against eps the step depends on the gradient's last bits
    big = np.abs(g) > 1e-4 * np.abs(g).max()
    assert np.abs(d_gpu[big] - d_cpu[big]).max() <= 1e-3 * lr
    assert np.abs(d_gpu).max() <= lr * 1.001
    m.close()


def tree_samples(engine, specs):
    Xs, rs = [], []
    for n, kernel, tol_pow, pts, seed in specs:
        xy = uniform_points(n, seed)[0] if pts == "uniform" else mixture_points(n, seed)[0]
        kp = [0.01, 0, 0, 0] if kernel == "gaussian" else None
        A = engine.AFMM(n, 64, tol_pow, xy, kernel=kernel, kparams=kp)
        A.assemble()
        X, r = A.rank_samples()
        Xs.append(X)
        rs.append(r)
        A.close()
    return np.concatenate(Xs), np.concatenate(rs)


@pytest.fixture(scope="module")
def trained(engine):
    specs = [(16384, k, t, pts, 100 + i)
             for i, (k, t, pts) in enumerate([(k, t, pts) for k in ("log", "inv_r", "gaussian")
                                              for t in (8, 10) for pts in ("uniform", "mixture")])]
    X, r = tree_samples(engine, specs)
    m = engine.RankModel(seed=11)
    y = engine.rank_target(r)
    l0 = float(np.mean((m.predict(X) - y) ** 2))
    loss = m.train(X, y, epochs=60, batch=256, lr=2e-3, seed=12)
    print_case("rank model: %d samples, loss %.4f -> %.4f" % (len(y), l0, loss))
    yield m, l0, loss
    m.close()


def test_training_learns_ranks(engine, trained):
    m, l0, loss = trained
    assert loss < 0.1 * l0
    # a tree it has not seen
    X, r = tree_samples(engine, [(20000, "log", 8, "uniform", 999)])
    pred = m.predict(X)
    err = np.abs(pred - engine.rank_target(r))
    print_case("held-out |log2 error| mean %.3f p95 %.3f" % (err.mean(), np.quantile(err, 0.95)))
    assert err.mean() < 0.5


@pytest.mark.parametrize("margin", [0.5, -3.0])
def test_learned_capacities_give_the_plain_assembly(engine, trained, margin):
    """Ranks, skeleton sizes and matvec results equal the plain assembly's,
    bit for bit; a negative margin forces fallbacks, which are counted."""
    m, _, _ = trained
    n = 1 << 17
    xy, q = uniform_points(n, 77)
    A = engine.AFMM(n, 64, 8, xy)
    A.assemble()
    B = engine.AFMM(n, 64, 8, xy)
    B.set_rank_model(m, margin)
    B.assemble()
    for j in range(2, A.nLevels + 1):
        np.testing.assert_array_equal(A.ranks(j), B.ranks(j))
    sa, sb = A.stats(), B.stats()
    print_case("margin %.1f: fallbacks %d of %d boxes, scratch %d -> %d bytes, ACA %.3f -> %.3f s"
               % (margin, sb["rank_fallbacks"], sb["rank_boxes"], sa["aca_scratch_peak"],
                  sb["aca_scratch_peak"], sa["aca_seconds"], sb["aca_seconds"]))
    assert sa["rank_fallbacks"] == 0 and sa["rank_boxes"] == sb["rank_boxes"] > 0
    if margin > 0:
        assert sb["rank_fallbacks"] <= 0.2 * sb["rank_boxes"]
        assert sb["aca_scratch_peak"] < sa["aca_scratch_peak"]
    else:
        assert sb["rank_fallbacks"] > 0
    np.testing.assert_array_equal(A.computeMatVecProduct(q), B.computeMatVecProduct(q))
    Q = np.random.default_rng(78).uniform(-1, 1, size=(n, 4))
    np.testing.assert_array_equal(A.matvec(Q, "fp64", blocked=True),
                                  B.matvec(Q, "fp64", blocked=True))
    A.close()
    B.close()


def test_untrained_model_falls_back(engine):
    m = engine.RankModel(seed=21)
    xy, q = uniform_points(20000, 22)
    A = engine.AFMM(20000, 64, 10, xy, kernel="inv_r")
    A.assemble()
    B = engine.AFMM(20000, 64, 10, xy, kernel="inv_r")
    B.set_rank_model(m, 0.0)
    B.assemble()
    st = B.stats()
    assert st["rank_fallbacks"] > 0
    np.testing.assert_array_equal(A.computeMatVecProduct(q), B.computeMatVecProduct(q))
    A.close()
    B.close()
    m.close()


def test_small_aca_budget(engine, trained, monkeypatch):
    """A 16 MiB ACA scratch budget (AFMM2D_ACA_BUDGET_MB) splits every level
    into many launches; plain and learned capacities still give the default
    assembly's ranks and matvec bit for bit, and the learned capacities need
    fewer launches."""
    m, _, _ = trained
    n = 1 << 16
    xy, q = uniform_points(n, 79)
    A = engine.AFMM(n, 64, 8, xy)
    A.assemble()
    ya = A.computeMatVecProduct(q)
    launches = {}
    monkeypatch.setenv("AFMM2D_ACA_BUDGET_MB", "16")
    for arm in ("plain", "learned"):
        B = engine.AFMM(n, 64, 8, xy)
        if arm == "learned":
            B.set_rank_model(m, 0.5)
        B.assemble()
        st = B.stats()
        launches[arm] = st["aca_launches"]
        assert st["aca_scratch_peak"] <= 16 << 20
        for j in range(2, A.nLevels + 1):
            np.testing.assert_array_equal(A.ranks(j), B.ranks(j))
        np.testing.assert_array_equal(ya, B.computeMatVecProduct(q))
        B.close()
    print_case("16 MiB ACA budget: launches plain %d, learned %d (default budget %d)"
               % (launches["plain"], launches["learned"], A.stats()["aca_launches"]))
    assert launches["learned"] < launches["plain"]
    assert launches["plain"] > A.stats()["aca_launches"]
    A.close()
