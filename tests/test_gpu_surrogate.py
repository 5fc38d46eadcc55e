"""Learned surrogate (gdt_surrogate_*): training lowers the loss, the GPU
forward agrees with a numpy forward of the exported weights (bf16 operand
rounding emulated), the projection conserves mass, momentum and energy
exactly, and fallback cells carry the plain solver's Q bit for bit."""
import numpy as np
import pytest

import gpu_determ_amd as G

from helpers import mixture, moments, rel_l2

pytestmark = pytest.mark.gpu

VMAX = 4.8


def say(msg):
    print(msg, flush=True)


def bf16(x):
    """Round to bf16 (nearest even) and back, as the kernels do."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def setup():
    ctx = G.Context((16,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
    ctx.korobov_build(1 << 16, seed=3, s0=0.25)
    rng = np.random.default_rng(21)
    ns = 4096
    f = np.stack([mixture(ctx.xi, rng, comps=1 + (i % 3), Trange=(0.3, 1.5), umax=1.0) for i in range(ns)])
    fd = G.DeviceArray.from_host(f)
    _, _, Qd = ctx.evaluate(G.METHOD_KOROBOV, fd, ns, G.FLAG_CORRECT)
    ctx.sync()
    return ctx, f, fd, Qd


def test_training_lowers_loss(setup):
    say("surrogate training")
    ctx, f, fd, Qd = setup
    s = G.Surrogate(ctx, hidden=128, seed=5)
    assert s.np == 2176 and s.params > 500000
    l1 = s.train(fd, Qd, epochs=1, batch=256, lr=3e-3)
    l2 = s.train(fd, Qd, epochs=20, batch=256, lr=3e-3)
    say("loss after 1 epoch %.4g, after 21 %.4g" % (l1, l2))
    assert np.isfinite(l2) and l2 < 0.7 * l1
    s.close()


def test_forward_matches_numpy(setup):
    say("surrogate forward vs numpy")
    ctx, f, fd, Qd = setup
    s = G.Surrogate(ctx, hidden=128, seed=6)
    s.train(fd, Qd, epochs=2, batch=256, lr=1e-3)
    W1, b1, W2, b2, W3, b3, norm = s.export()
    n = ctx.n_pnt
    cells = 64
    x = f[:cells]
    dens = x.sum(axis=1) * ctx.d3v
    xs = np.zeros((cells, s.np), np.float32)
    xs[:, :n] = ((x / dens[:, None] - norm[0, :n]) / norm[1, :n]).astype(np.float32)
    xb = bf16(xs).astype(np.float64)
    u = xb @ bf16(W1).T.astype(np.float64) + b1
    w = xb @ bf16(W2).T.astype(np.float64) + b2
    y = bf16(u * w).astype(np.float64) @ bf16(W3).T.astype(np.float64) + b3
    Qref = dens[:, None] ** 2 * (y[:, :n] * norm[3, :n] + norm[2, :n])
    Qgpu = s.predict(G.DeviceArray.from_host(x)).get()
    assert rel_l2(Qgpu, Qref) < 1e-2, rel_l2(Qgpu, Qref)
    s.close()


def test_infer_projection_and_fallback(setup):
    say("surrogate inference, projection, fallback")
    ctx, f, fd, Qd = setup
    s = G.Surrogate(ctx, hidden=128, seed=7)
    s.train(fd, Qd, epochs=6, batch=256, lr=3e-3)
    rng = np.random.default_rng(22)
    cells = 96
    x = np.stack([mixture(ctx.xi, rng, comps=1 + (i % 3), Trange=(0.3, 1.5), umax=1.0) for i in range(cells)])
    xd = G.DeviceArray.from_host(x)
    _, _, Qp = ctx.evaluate(G.METHOD_KOROBOV, xd, cells, G.FLAG_CORRECT | G.FLAG_CELL_TILES)  # the fallback's solver
    Qp = Qp.get()
    # generous threshold: the projected network output is kept for most cells
    Q, fb, res = s.infer(xd, threshold=0.5)
    Q = Q.get()
    for c in range(cells):
        m = moments(ctx.xi, Q[c], ctx.d3v)
        scale = np.abs(moments(ctx.xi, np.abs(Q[c]), ctx.d3v)) + 1e-300
        assert np.all(np.abs(m) / scale < 1e-12), (c, m / scale)
    err = rel_l2(Q[~fb], Qp[~fb]) if (~fb).any() else 0.0
    say("kept %d of %d cells, rel L2 vs plain %.3g, residuals %.3g..%.3g" % ((~fb).sum(), cells, err, res.min(),
                                                                           res.max()))
    assert np.array_equal(Q[fb], Qp[fb])
    # the 1e-4 threshold: every flagged cell is the plain solver's, bitwise
    Q4, fb4, res4 = s.infer(xd, threshold=1e-4)
    Q4 = Q4.get()
    assert np.array_equal(fb4, res4 > 1e-4)
    assert np.array_equal(Q4[fb4], Qp[fb4])
    s.close()


def test_surrogate_pair_24cube():
    """The learned-surrogate configuration's grid (24^3, 7,208 nodes; a small
    pair count here): two networks as a pair.  The kept output is the exact
    projection of the two predictions' mean (it conserves mass, momentum and
    energy, and differs from the mean by f times a combination of 1, xi,
    |xi|^2); every residual is at least the pair's disagreement |Qa - Qb| /
    |Q|; flagged cells carry the plain solver's Q bit for bit; the forward
    pass of one network matches numpy."""
    say("surrogate pair at 24^3")
    ctx = G.Context((24,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
    ctx.korobov_build(1 << 18, seed=3, s0=0.25)
    n = ctx.n_pnt
    assert n == 7208
    rng = np.random.default_rng(24)
    ns = 2048
    f = np.stack([mixture(ctx.xi, rng, comps=1 + (i % 3), Trange=(0.3, 1.5), umax=1.0) for i in range(ns)])
    fd = G.DeviceArray.from_host(f)
    _, _, Qd = ctx.evaluate(G.METHOD_KOROBOV, fd, ns, G.FLAG_CORRECT | G.FLAG_CELL_TILES)
    sa, sb = G.Surrogate(ctx, hidden=128, seed=1), G.Surrogate(ctx, hidden=128, seed=2)
    sa.train(fd, Qd, epochs=2, batch=256, lr=1e-3)
    sb.train(fd, Qd, epochs=2, batch=256, lr=1e-3)
    sa.set_partner(sb)
    cells = 48
    x = np.stack([mixture(ctx.xi, rng, comps=1 + (i % 3), Trange=(0.3, 1.5), umax=1.0) for i in range(cells)])
    xd = G.DeviceArray.from_host(x)
    qa, qb = sa.predict(xd).get(), sb.predict(xd).get()
    # a loose threshold keeps every finite cell: the projection of the mean
    Q1, fb1, res1 = sa.infer(xd, threshold=1e3)
    Q1 = Q1.get()
    assert not fb1.any()
    e = np.sum(ctx.xi ** 2, axis=1)
    A = np.stack([np.ones(n), ctx.xi[:, 0], ctx.xi[:, 1], ctx.xi[:, 2], e], axis=1)
    for c in range(cells):
        m = moments(ctx.xi, Q1[c], ctx.d3v)
        scale = np.abs(moments(ctx.xi, np.abs(Q1[c]), ctx.d3v)) + 1e-300
        assert np.all(np.abs(m) / scale < 1e-12)
        mean = 0.5 * (qa[c] + qb[c])
        y = mean - Q1[c]  # = f (X . psi): fit X, weights f
        B = x[c][:, None] * A
        coef, *_ = np.linalg.lstsq(B, y, rcond=None)
        assert np.max(np.abs(B @ coef - y)) <= 1e-12 * max(np.max(np.abs(mean)), 1e-300)
    dis = np.linalg.norm(qa - qb, axis=1) / np.linalg.norm(Q1, axis=1)
    assert np.all(res1 >= dis * (1 - 1e-9))
    # the median residual as threshold: kept cells as before, the others the
    # plain solver's bit for bit
    thr = float(np.median(res1))
    Q2, fb2, res2 = sa.infer(xd, threshold=thr)
    Q2 = Q2.get()
    assert fb2.any() and (~fb2).any()
    assert np.array_equal(fb2, res2 > thr)
    _, _, Qp = ctx.evaluate(G.METHOD_KOROBOV, xd, cells, G.FLAG_CORRECT | G.FLAG_CELL_TILES)
    Qp = Qp.get()
    assert np.array_equal(Q2[fb2], Qp[fb2])
    assert np.array_equal(Q2[~fb2], Q1[~fb2])
    say("kept %d of %d at threshold %.3g; residual / disagreement min %.3g" % ((~fb2).sum(), cells, thr,
                                                                              (res1 / dis).min()))
    sa.set_partner(None)
    sa.close()
    sb.close()
