"""Spatial transport (first-order upwind, gdt_transport): the oracle's
restatement pinned on the CPU by closed forms, and the GPU kernel against
it (bit for bit: the same roundings in the same order)."""
import numpy as np
import pytest

import gpu_determ_amd as G
import oracle as O
from helpers import mixture


def test_oracle_hand_case():
    """3 cells, f = (1, 2, 4), xi = +1, dt / h = 0.5, periodic:
    f' = f - 0.5 (f - f_left) = (2.5, 1.5, 3.0)."""
    xi = np.array([[1.0, 0.0, 0.0]])
    out = O.transport_upwind(np.array([[1.0], [2.0], [4.0]]), xi, (3, 1, 1), (1.0, 1.0, 1.0), 0.5)
    assert np.array_equal(out.ravel(), [2.5, 1.5, 3.0])
    # zero gradient: the first cell has no left neighbour, so it keeps its value
    out = O.transport_upwind(np.array([[1.0], [2.0], [4.0]]), xi, (3, 1, 1), (1.0, 1.0, 1.0), 0.5, (1, 0, 0))
    assert np.array_equal(out.ravel(), [1.0, 1.5, 3.0])


def test_oracle_courant_one_is_a_shift():
    """At Courant number 1 along one axis the upwind step moves f by exactly
    one cell (periodic), in the direction of the velocity."""
    rng = np.random.default_rng(1)
    xi = np.array([[2.0, 0, 0], [-2.0, 0, 0], [0, 2.0, 0], [0, -2.0, 0], [0, 0, 2.0], [0, 0, -2.0], [0, 0, 0]])
    shape = (5, 4, 3)
    f = rng.random((60, 7))
    out = O.transport_upwind(f, xi, shape, (1.0, 1.0, 1.0), 0.5).reshape(*shape, 7)
    f4 = f.reshape(*shape, 7)
    for v, (axis, sh) in enumerate([(0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)]):
        assert np.array_equal(out[..., v], np.roll(f4[..., v], sh, axis=axis))
    assert np.array_equal(out[..., 6], f4[..., 6])


def test_oracle_conservation_and_max_principle():
    """Periodic: every velocity's cell sum is conserved (the flux differences
    telescope).  Zero gradient with sum |k| <= 1: the step is a convex
    combination, so values stay within the input's range, and a constant
    stays constant."""
    rng = np.random.default_rng(2)
    xi = rng.uniform(-3.0, 3.0, (40, 3))
    shape, h = (6, 5, 4), (1.0, 0.8, 1.2)
    dt = 0.95 / np.max(np.abs(xi) @ (1.0 / np.array(h)))
    f = rng.random((120, 40))
    out = O.transport_upwind(f, xi, shape, h, dt)
    assert np.max(np.abs(out.sum(axis=0) - f.sum(axis=0)) / f.sum(axis=0)) < 1e-14
    out = O.transport_upwind(f, xi, shape, h, dt, (1, 1, 1))
    assert out.min() >= f.min() - 1e-15 and out.max() <= f.max() + 1e-15
    c = np.full((120, 40), 0.37)
    assert np.array_equal(O.transport_upwind(c, xi, shape, h, dt, (1, 0, 1)), c)


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(0, 0, 0), (1, 1, 1), (0, 1, 0)])
@pytest.mark.parametrize("gn", [12, 11])
def test_gpu_transport_matches_oracle(bc, gn):
    """12^3 has 912 velocity nodes (even: two velocities per thread, 16-B
    accesses), 11^3 has 739 (odd: one velocity per thread)."""
    print("transport", gn, bc, flush=True)
    ctx = G.Context((gn,) * 3, (-4.8,) * 3, (4.8,) * 3, 4.8)
    assert ctx.n_pnt % 2 == (0 if gn == 12 else 1)
    shape, h = (5, 4, 3), (0.5, 0.4, 0.6)
    rng = np.random.default_rng(3)
    f = np.stack([mixture(ctx.xi, rng) for _ in range(60)])
    dt = 0.9 / np.max(np.abs(ctx.xi) @ (1.0 / np.array(h)))
    fd = G.DeviceArray.from_host(f)
    out = ctx.transport(fd, shape, h, dt, bc)
    ctx.sync()
    ref = O.transport_upwind(f, ctx.xi, shape, h, dt, bc)
    assert np.array_equal(out.get(), ref)
    if bc == (0, 0, 0):
        s0, s1 = f.sum(axis=0), out.get().sum(axis=0)
        assert np.max(np.abs(s1 - s0) / s0) < 1e-14


@pytest.mark.gpu
def test_gpu_transport_errors():
    ctx = G.Context((8, 8, 8), (-4.0,) * 3, (4.0,) * 3, -1.0)
    f = G.DeviceArray((6, ctx.n_pnt))
    f.zero()
    h = (1.0, 1.0, 1.0)
    dt_max = 1.0 / np.max(np.abs(ctx.xi).sum(axis=1))
    ctx.transport(f, (3, 2, 1), h, dt_max * 0.99)
    with pytest.raises(G.GdtError):
        ctx.transport(f, (3, 2, 1), h, dt_max * 1.01)  # Courant number above 1
    with pytest.raises(G.GdtError):
        ctx.transport(f, (3, 2, 1), h, 0.1, out=f)  # aliasing

    class Shifted:  # f_out one cell into f: a partial overlap
        ptr = f.ptr + 8 * ctx.n_pnt

    with pytest.raises(G.GdtError):
        ctx.transport(f, (3, 2, 1), h, 0.1, out=Shifted)
    with pytest.raises(G.GdtError):
        ctx.transport(f, (3, 2, 1), h, 0.1, (0, 2, 0))  # unknown boundary


@pytest.mark.gpu
def test_gpu_transport_and_collisions_conserve():
    """Alternating transport (periodic) and corrected Korobov collision steps
    over 27 cells keeps the total mass, momentum and energy of the box."""
    print("transport + collisions", flush=True)
    ctx = G.Context((12, 12, 12), (-4.8,) * 3, (4.8,) * 3, 4.8)
    ctx.korobov_build(1 << 15, seed=3, s0=0.25)
    shape, h = (3, 3, 3), (1.0, 1.0, 1.0)
    rng = np.random.default_rng(5)
    f = np.stack([mixture(ctx.xi, rng) for _ in range(27)])
    dt = 0.5 / np.max(np.abs(ctx.xi).sum(axis=1))
    a, b = G.DeviceArray.from_host(f), G.DeviceArray((27, ctx.n_pnt))

    def totals(x):
        w = np.concatenate([np.ones((ctx.n_pnt, 1)), ctx.xi, (ctx.xi ** 2).sum(axis=1, keepdims=True)], axis=1)
        return (x @ w).sum(axis=0)

    t0 = totals(f)
    for _ in range(10):
        ctx.transport(a, shape, h, dt, out=b)
        ctx.run_steps(G.METHOD_KOROBOV, b, 27, dt, 1, G.FLAG_CORRECT)
        a, b = b, a
    ctx.sync()
    t1 = totals(a.get())
    scale = np.abs(t0).max()
    assert np.max(np.abs(t1 - t0)) < 1e-12 * scale
