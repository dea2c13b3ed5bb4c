"""Batches of problems (afmm2d_batch_*): every problem's output against its
own single-problem blocked matvec on the same handle, charges and precision.

A batch runs the same per-task arithmetic as the single-problem passes (one
task list per pass over all problems; a task's rows and their summation order
do not depend on the other tasks in its launch, nor on the tuned launch
parameters), so the bar is bit-for-bit equality.  The single-RHS FP64 passes
are a second reference within 1e-12 (the same factors summed in another
order)."""
import numpy as np
import pytest

from conftest import uniform_points

pytestmark = pytest.mark.gpu

KERNEL_SET = (("log", None), ("inv_r", None), ("gaussian", [0.01, 0, 0, 0]))


def print_case(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_problems(engine, count, seed, nmin=2048, nmax=65536, tol_pow=8):
    """count problems with N log-uniform in [nmin, nmax], kernels drawn
    uniformly from log r, 1/r, exp(-r^2/0.01), points U[-1,1]^2."""
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        name, kp = KERNEL_SET[int(rng.integers(0, 3))]
        xy, _ = uniform_points(n, seed * 1000 + i)
        probs.append(engine.AFMM(n, 64, tol_pow, xy, kernel=name, kparams=kp))
    return probs, rng


def device_io(hip, probs, rng, nrhs):
    Q = [np.asfortranarray(rng.standard_normal((p.N, nrhs))) for p in probs]
    dB = [hip.DeviceArray.from_numpy(q) for q in Q]
    dY = [hip.DeviceArray(q.nbytes) for q in Q]
    return Q, dB, dY


def download(dY, probs, nrhs):
    out = []
    for d, p in zip(dY, probs):
        y = np.zeros((p.N, nrhs), order="F")
        out.append(d.download(y))
    return out


def test_batch_256_problems_graph_and_tuner(engine):
    """The batching workload: 256 problems, N log-uniform in [2048, 65536],
    mixed kernels, 8 RHS ~ N(0,1), FP64 factors; eager, captured graph
    (replayed with other work between replays) and tuned parameters."""
    from afmm2d_amd import hip
    nrhs = 8
    print_case("batch 256 problems: create + assemble")
    probs, rng = make_problems(engine, 256, seed=3)
    engine.batch_assemble(probs)
    Q, dB, dY = device_io(hip, probs, rng, nrhs)
    print_case("  total points", sum(p.N for p in probs))
    B = engine.Batch(probs, [d.ptr for d in dB], [d.ptr for d in dY], nrhs, "fp64")
    inf = B.info()
    assert inf["nproblems"] == 256 and inf["captured"] == 0
    B.matvec()
    hip.synchronize()
    Ye = download(dY, probs, nrhs)
    # each problem against its own blocked matvec (bit for bit) and, for a
    # sample, the single-RHS passes
    for i, p in enumerate(probs):
        Ys = p.matvec(Q[i], "fp64", blocked=True)
        np.testing.assert_array_equal(Ye[i], Ys, err_msg="problem %d (N=%d)" % (i, p.N))
        if i % 32 == 0:
            Y1 = p.matvec(Q[i], "fp64")
            assert max(rel(Ye[i][:, c], Y1[:, c]) for c in range(nrhs)) <= 1e-12
    # the captured graph, replayed with eager work of the problems between
    print_case("  graph capture + replays")
    B.capture(True)
    assert B.info()["captured"] == 1
    for rep in range(3):
        for d in dY:
            d.fill_zero()
        B.matvec()
        hip.synchronize()
        Yg = download(dY, probs, nrhs)
        for i in range(len(probs)):
            np.testing.assert_array_equal(Yg[i], Ye[i], err_msg="replay %d problem %d" % (rep, i))
        probs[rep].matvec(Q[rep][:, 0], "fp64", blocked=True)  # other work between replays
    # the tuner: parameters change, results do not; the graph is captured again
    print_case("  tune")
    ms = B.tune()
    inf = B.info()
    print_case("  tuned %.3f ms/run: %s" % (ms, inf))
    assert ms > 0 and inf["tuned"] == 1 and inf["captured"] == 1
    for d in dY:
        d.fill_zero()
    B.matvec()
    hip.synchronize()
    Yt = download(dY, probs, nrhs)
    for i in range(len(probs)):
        np.testing.assert_array_equal(Yt[i], Ye[i], err_msg="tuned problem %d" % i)
    # a second batch of the same shape class starts from the cached tuning
    B2 = engine.Batch(probs, [d.ptr for d in dB], [d.ptr for d in dY], nrhs, "fp64")
    inf2 = B2.info()
    assert inf2["tuned"] == 1
    assert all(inf2[k] == inf[k] for k in ("depth", "xcd", "near_waves", "rows_m2l"))
    B2.close()
    B.close()
    for d in dB + dY:
        d.free()
    for p in probs:
        p.close()


@pytest.mark.parametrize("precision,nrhs", [("fp32", 16), ("bf16", 20), ("fp64", 32),
                                             ("fp64", 5), ("fp64", 1)])
def test_batch_precisions_and_column_counts(engine, precision, nrhs):
    # FP64 with 5 and 1 columns: the near field's pair partials packed at 6
    # and 2 columns per row (each problem's own matvec keeps 16)
    from afmm2d_amd import hip
    print_case("batch", precision, nrhs)
    probs, rng = make_problems(engine, 12, seed=7, nmin=3000, nmax=40000, tol_pow=10)
    engine.batch_assemble(probs)
    Q, dB, dY = device_io(hip, probs, rng, nrhs)
    B = engine.Batch(probs, [d.ptr for d in dB], [d.ptr for d in dY], nrhs, precision)
    B.matvec()
    hip.synchronize()
    Yb = download(dY, probs, nrhs)
    for i, p in enumerate(probs):
        np.testing.assert_array_equal(Yb[i], p.matvec(Q[i], precision, blocked=True),
                                      err_msg="problem %d" % i)
    B.close()
    for d in dB + dY:
        d.free()
    for p in probs:
        p.close()


def test_batch_leading_dimensions_and_refusals(engine):
    from afmm2d_amd import hip
    probs, rng = make_problems(engine, 3, seed=11, nmin=2048, nmax=8192)
    nrhs = 3
    lds = [p.N + 5 for p in probs]
    Q = [np.asfortranarray(rng.uniform(-1, 1, size=(p.N, nrhs))) for p in probs]
    # unassembled handles are refused
    dB = [hip.DeviceArray(8 * ld * nrhs) for ld in lds]
    dY = [hip.DeviceArray(8 * ld * nrhs) for ld in lds]
    with pytest.raises(engine.AFMM2DError) as ex:
        engine.Batch(probs, [d.ptr for d in dB], [d.ptr for d in dY], nrhs, "fp64", lds, lds)
    assert ex.value.status == engine.ERR_STATE
    engine.batch_assemble(probs)
    for d, q, ld in zip(dB, Q, lds):
        pad = np.zeros((ld, nrhs), order="F")
        pad[: q.shape[0]] = q
        d.upload(pad)
    for d, ld in zip(dY, lds):
        d.upload(np.full((ld, nrhs), 7.0, order="F"))
    B = engine.Batch(probs, [d.ptr for d in dB], [d.ptr for d in dY], nrhs, "fp64", lds, lds)
    B.matvec()
    hip.synchronize()
    for i, (p, d, ld) in enumerate(zip(probs, dY, lds)):
        Y = d.download(np.zeros((ld, nrhs), order="F"))
        np.testing.assert_array_equal(Y[: p.N], p.matvec(Q[i], "fp64", blocked=True))
        assert (Y[p.N:] == 7.0).all()
    B.close()
    # bad arguments
    for nr, ldv, prec in ((33, None, "fp64"), (0, None, "fp64"), (2, [16] * 3, "fp64")):
        with pytest.raises(engine.AFMM2DError) as ex:
            engine.Batch(probs, [d.ptr for d in dB], [d.ptr for d in dY], nr, prec, ldv, ldv)
        assert ex.value.status == engine.ERR_INVALID_ARGUMENT
    # a non-symmetric user kernel (two operators per handle) is refused
    xy, _ = uniform_points(4096, 5)
    U = engine.AFMM(4096, 64, 10, xy, kernel="user", kparams=[0.0, 0.25, 0, 0])
    U.assemble()
    dU = hip.DeviceArray(8 * 4096 * nrhs)
    with pytest.raises(engine.AFMM2DError) as ex:
        engine.Batch([probs[0], U], [dB[0].ptr, dU.ptr], [dY[0].ptr, dU.ptr], nrhs, "fp64",
                     [lds[0], 4096], [lds[0], 4096])
    assert ex.value.status == engine.ERR_UNSUPPORTED
    U.close()
    for d in dB + dY + [dU]:
        d.free()
    for p in probs:
        p.close()


def test_batch_on_48_bit_storage(engine):
    """Problems whose M2L blocks are in 48-bit storage: the batch reads both
    orientations of every block like each problem's own blocked matvec (bit
    for bit) and the single-RHS passes on the same stored values (1e-12);
    mixing storages in one batch is refused."""
    from afmm2d_amd import hip
    nrhs = 8
    rng0 = np.random.default_rng(41)
    probs = []
    for i in range(12):
        n = int(round(np.exp(rng0.uniform(np.log(2048), np.log(40000)))))
        name, kp = KERNEL_SET[i % 3]
        xy, _ = uniform_points(n, 4100 + i)
        p = engine.AFMM(n, 64, 8, xy, kernel=name, kparams=kp)
        p.set_m2l_storage("bfp48")
        probs.append(p)
    engine.batch_assemble(probs)
    Q, dB, dY = device_io(hip, probs, rng0, nrhs)
    B = engine.Batch(probs, [d.ptr for d in dB], [d.ptr for d in dY], nrhs, "fp64")
    B.capture(True)
    for _ in range(2):
        B.matvec()
        hip.synchronize()
        Yb = download(dY, probs, nrhs)
        for i, p in enumerate(probs):
            np.testing.assert_array_equal(Yb[i], p.matvec(Q[i], "fp64", blocked=True))
    for i in (0, 5, 11):
        Y1 = probs[i].matvec(Q[i], "fp64")
        assert max(rel(Yb[i][:, c], Y1[:, c]) for c in range(nrhs)) <= 1e-12
    B.close()
    # one problem with FP64 blocks among 48-bit ones
    xy, _ = uniform_points(3000, 4200)
    f = engine.AFMM(3000, 64, 8, xy)
    f.assemble()
    qf = np.asfortranarray(rng0.standard_normal((3000, nrhs)))
    db, dy = hip.DeviceArray.from_numpy(qf), hip.DeviceArray(qf.nbytes)
    with pytest.raises(engine.AFMM2DError):
        engine.Batch(probs[:2] + [f], [dB[0].ptr, dB[1].ptr, db.ptr],
                     [dY[0].ptr, dY[1].ptr, dy.ptr], nrhs, "fp64")
    for a in dB + dY + [db, dy]:
        a.free()
    for p in probs + [f]:
        p.close()
