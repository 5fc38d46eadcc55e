"""GPU checks of the smaller rows of the hot path and of cached state.

* the 14 moment-basis planes (fill_correction_array, velocity.cu:17-50)
  against their closed form;
* relax_f with the Courant time step (get_delta_t / evolve_f,
  collision_integral.cu:776-833) against the oracle, time += dt included;
* the layout transpose entry point (transpose_float, :528-547) against
  numpy;
* cached state that must follow its inputs: the step graph after the scratch
  it captured is reallocated, the matrix-core operands after the tables
  change, the surrogate's residual gate on cells whose sums are not finite,
  and the matrix gain with more than 64 KiB of LDS;
* destroyed contexts give back all their device memory; calls beyond the
  launch limit of 65,535 cells are refused with an error; bad arguments
  return the ABI's error codes.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import gpu_determ_amd as G
import oracle as O
from helpers import maxwellian, mixture, rel_l2

pytestmark = pytest.mark.gpu
VM = 4.8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_basis_planes_closed_form():
    """Plane q of gdt_basis at box node (i, j, k) is the q-th of (x, y, z, e,
    xx, xy, xz, xe, yy, yz, ye, zz, ze, ee) with x = vmin + d (2i + 1) / 2 and
    e = x^2 + y^2 + z^2: the XI_* order of velocity.h:22-35.  A non-cubic,
    offset box, so a swapped axis or a wrong plane order shows."""
    say("basis planes")
    n, vmin, vmax = (6, 8, 10), (-3.0, -4.0, -2.5), (3.0, 2.0, 5.0)
    c = G.Context(n, vmin, vmax, -1.0)
    got = c.basis()
    ax = [vmin[a] + (vmax[a] - vmin[a]) / n[a] * (2 * np.arange(n[a]) + 1) / 2 for a in range(3)]
    x, y, z = (v.ravel() for v in np.meshgrid(*ax, indexing="ij"))
    e = x * x + y * y + z * z
    ref = np.stack([x, y, z, e, x * x, x * y, x * z, x * e, y * y, y * z, y * e, z * z, z * e, e * e])
    assert got.shape == ref.shape
    for q in range(14):
        scale = np.max(np.abs(ref[q]))
        assert np.max(np.abs(got[q] - ref[q])) <= 1e-15 * scale, q


def test_relax_courant_matches_oracle():
    """gdt_relax with GDT_FLAG_COURANT: dt = min courant / nu over every
    value (the oracle's get_delta_t), f <- f (1 - nu dt) + gain dt, and the
    device time advanced by exactly that dt, over two calls in a row."""
    say("relax courant")
    g = O.Grid((16,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    c = G.Context(g.n, g.vmin, g.vmax, g.R)
    c.korobov_build(1 << 16, seed=4, s0=0.25)
    rng = np.random.default_rng(31)
    cells = 3
    f = np.stack([mixture(g.xi, rng) for _ in range(cells)])
    fd = G.DeviceArray.from_host(f)
    t = G.DeviceArray.from_host(np.array([0.125]))
    t_ref = 0.125
    f_ref = f.copy()
    for courant in (0.4, 0.25):
        nu, gain, _ = c.evaluate(G.METHOD_KOROBOV, fd, cells, G.FLAG_CORRECT)
        nu_h, gain_h = nu.get(), gain.get()
        assert np.array_equal(fd.get(), f_ref)
        c.relax(fd, nu, gain, cells, courant=courant, time=t, flags=G.FLAG_COURANT)
        c.sync()
        dt = O.courant_dt(nu_h, courant)
        assert 0.0 < dt < 1.0e10
        t_ref = t_ref + dt
        f_ref = f_ref * (1.0 - nu_h * dt) + gain_h * dt
        got = fd.get()
        assert t.get()[0] == t_ref  # the min of correctly rounded quotients: exact
        assert np.max(np.abs(got - f_ref) / np.max(np.abs(f_ref))) < 1e-15
        f_ref = got


@pytest.mark.parametrize("rows,cols", [(37, 1000), (1000, 37), (64, 64), (1, 513)])
def test_transpose_entry(rows, cols):
    say("transpose", rows, cols)
    c = G.Context((4, 4, 4), (-1.0,) * 3, (1.0,) * 3, -1.0)
    a = np.random.default_rng(rows * 7 + cols).standard_normal((rows, cols))
    src = G.DeviceArray.from_host(a)
    dst = G.DeviceArray((cols, rows))
    c.transpose(src, dst, rows, cols)
    c.sync()
    assert np.array_equal(dst.get(), a.T)


def test_step_graph_follows_reallocated_scratch():
    """A graph of steps on one cell, then an evaluation of 64 cells on the
    same context (its scratch grows: the captured buffers are freed), then
    the graph steps again: the result equals eager steps bit for bit."""
    say("graph after scratch growth")
    g = O.Grid((16,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    c = G.Context(g.n, g.vmin, g.vmax, g.R)
    c.korobov_build(1 << 16, seed=3, s0=0.25)
    rng = np.random.default_rng(5)
    f0 = mixture(g.xi, rng)[None]
    fa = G.DeviceArray.from_host(f0)
    fb = G.DeviceArray.from_host(f0)
    c.run_steps(G.METHOD_KOROBOV, fa, 1, 0.02, 5, G.FLAG_CORRECT | G.FLAG_GRAPH)
    big = G.DeviceArray.from_host(np.stack([mixture(g.xi, rng) for _ in range(64)]))
    _, _, Qbig = c.evaluate(G.METHOD_KOROBOV, big, 64, G.FLAG_CORRECT)
    c.run_steps(G.METHOD_KOROBOV, fa, 1, 0.02, 5, G.FLAG_CORRECT | G.FLAG_GRAPH)
    c.run_steps(G.METHOD_KOROBOV, fb, 1, 0.02, 10, G.FLAG_CORRECT)
    c.sync()
    assert np.all(np.isfinite(Qbig.get()))
    assert np.array_equal(fa.get(), fb.get())


def test_matrix_mfma_follows_new_tables():
    """gdt_matrix_mfma after gdt_set_tables: the matrix-core operands are
    rebuilt from the new tables (doubled b and a double the gain), and the
    gain agrees with the FP64 matrix path on those tables."""
    say("matrix cores after new tables")
    c = G.Context((16,) * 3, (-VM,) * 3, (VM,) * 3, -1.0)
    c.build_tables(G.TABLES_FP64)
    rng = np.random.default_rng(7)
    cells = 4
    f = []
    for _ in range(cells):
        m = maxwellian(c.xi, rng.uniform(0.5, 1.0), rng.uniform(0.6, 1.4), rng.uniform(-0.8, 0.8, 3))
        p = np.clip(m * (1 + 0.05 * rng.standard_normal(m.shape)), 1e-300, None)
        f.append(p * (m.sum() / p.sum()))
    fd = G.DeviceArray.from_host(np.stack(f))
    _, g1, _ = c.matrix_mfma(G.GEMM_F32, fd, cells, 0)
    g1 = g1.get()
    b, a, _ = c.tables()
    c.set_tables(G.DeviceArray.from_host(2.0 * b), G.DeviceArray.from_host(2.0 * a))
    _, g2, _ = c.matrix_mfma(G.GEMM_F32, fd, cells, 0)
    _, g64 = c.collide_matrix(fd, cells)
    g2 = g2.get()
    assert rel_l2(g2, 2.0 * g1) < 1e-12
    assert rel_l2(g2, g64.get()) < 2e-5


def test_surrogate_nonfinite_cells_fall_back():
    """A zero-density cell (0/0 in the network's input scaling) and a cell
    scaled by 1e100 (its squared Q overflows in the residual sums) both reach
    the plain solver: flagged, finite, and equal to the plain solver's Q bit
    for bit; ordinary cells are unaffected."""
    say("surrogate non-finite gate")
    c = G.Context((16,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    c.korobov_build(1 << 16, seed=3, s0=0.25)
    rng = np.random.default_rng(41)
    ns = 512
    f = np.stack([mixture(c.xi, rng, comps=1 + (i % 3), umax=1.0) for i in range(ns)])
    fd = G.DeviceArray.from_host(f)
    _, _, Qd = c.evaluate(G.METHOD_KOROBOV, fd, ns, G.FLAG_CORRECT)
    s = G.Surrogate(c, hidden=128, seed=5)
    s.train(fd, Qd, epochs=1, batch=256, lr=1e-3)
    cells = np.stack([f[0], f[1], np.zeros(c.n_pnt), 1e100 * f[2], f[3]])
    cd = G.DeviceArray.from_host(cells)
    Q, fb, res = s.infer(cd, threshold=1e6)  # a loose gate: only non-finite cells fail it
    Qh = Q.get()
    say("fallback", fb.astype(int).tolist(), "residuals", res.tolist())
    assert fb[2] and fb[3]
    assert not fb[0] and not fb[1] and not fb[4]
    assert np.all(np.isfinite(Qh))
    _, _, Qp = c.evaluate(G.METHOD_KOROBOV, cd, 5, G.FLAG_CORRECT | G.FLAG_CELL_TILES)  # the fallback's solver
    Qp = Qp.get()
    assert np.array_equal(Qh[2], Qp[2]) and np.array_equal(Qh[3], Qp[3])
    assert np.all(Qh[2] == 0.0)
    s.close()


def test_matrix_gain_22cube_large_lds():
    """The matrix gain on a 22^3 box holds f in 108 KiB of LDS (rows padded to
    28 doubles): above the default 64 KiB, so the kernel runs only if it was
    enabled for its dynamic LDS.  Loss everywhere and gain on every 37th
    destination against the oracle."""
    say("matrix 22^3")
    g = O.Grid((22,) * 3, (-VM,) * 3, (VM,) * 3, -1.0)
    c = G.Context(g.n, g.vmin, g.vmax, g.R)
    c.build_tables(G.TABLES_FP64)
    ob, oa = O.build_tables(g, 0)
    rng = np.random.default_rng(17)
    f = mixture(g.xi, rng)[None]
    nu, gain = c.collide_matrix(G.DeviceArray.from_host(f), 1)
    fb = g.to_box(f)
    onu = g.to_cmp(O.loss(g, ob, fb))
    dest = g.box_index()[::37].astype(np.int32)
    og = O.gain(g, oa, fb, dest)[0]
    assert rel_l2(nu.get(), onu) < 1e-13
    assert rel_l2(gain.get()[0, ::37], og) < 1e-13


def test_destroyed_contexts_free_their_device_memory():
    """A destroyed context frees every device buffer it grew, including the
    cell-tile stencil and scratch (node-term rows of a chunk, f tiles, Q0 / L
    rows): after a first round (which may leave the runtime's own one-time
    allocations behind: code objects, kernel-argument pools), creating, using
    and destroying a context again leaves the device's free memory where it
    was (within 16 MiB per round), while one context holds close to a GiB."""
    import ctypes
    import gc
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()

    def free_bytes():
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
        return free.value

    g = O.Grid((24,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    rng = np.random.default_rng(23)
    f = np.stack([mixture(g.xi, rng) for _ in range(130)])
    start, held, after = free_bytes(), [], []
    for it in range(4):
        c = G.Context(g.n, g.vmin, g.vmax, g.R)
        c.korobov_build(1 << 20, seed=1, s0=0.25)
        fd = G.DeviceArray.from_host(f)
        for fl in (G.FLAG_CELL_TILES, G.FLAG_MIXED, 0):
            out = c.evaluate(G.METHOD_KOROBOV, fd, f.shape[0], G.FLAG_CORRECT | fl)
            del out
        c.sync()
        held.append(start - free_bytes())
        del fd
        c.close()
        del c
        gc.collect()
        after.append(start - free_bytes())
        say("context round %d: held %.1f MiB, left %.1f MiB" % (it, held[-1] / 2**20, after[-1] / 2**20))
    assert min(held) > 512 << 20
    assert max(after[1:]) - after[0] < 16 << 20


def test_cells_beyond_the_launch_limit_are_refused():
    """65,536 cells in one call exceed the grid-y limit of the grouped Korobov
    and analytic-matrix launches: the call returns an error naming the limit
    instead of failing inside a launch; 65,535 cells still run."""
    say("grid-y limit")
    g = O.Grid((6,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    c = G.Context(g.n, g.vmin, g.vmax, g.R)
    c.korobov_build(1 << 12, seed=1, s0=0.25)
    rng = np.random.default_rng(29)
    f1 = mixture(g.xi, rng)
    fd = G.DeviceArray.from_host(np.tile(f1, (65536, 1)))
    with pytest.raises(RuntimeError, match="65,535"):
        c.evaluate(G.METHOD_KOROBOV, fd, 65536, G.FLAG_CORRECT)
    _, _, Q = c.evaluate(G.METHOD_KOROBOV, fd, 65535, G.FLAG_CORRECT)
    Qh = Q.get()
    assert np.array_equal(Qh[0], Qh[-1])
    _, _, Q1 = c.evaluate(G.METHOD_KOROBOV, G.DeviceArray.from_host(f1[None]), 1, G.FLAG_CORRECT)
    assert np.array_equal(Qh[0], Q1.get()[0])
    g2 = O.Grid((6,) * 3, (-VM,) * 3, (VM,) * 3, -1.0)
    c2 = G.Context(g2.n, g2.vmin, g2.vmax, g2.R)
    c2.build_tables(G.TABLES_FP64)
    fd2 = G.DeviceArray.from_host(np.tile(np.exp(-np.sum(g2.xi ** 2, axis=1) / 2), (65536, 1)))
    with pytest.raises(RuntimeError, match="65,535"):
        c2.collide_matrix(fd2, 65536)


def test_argument_errors():
    """The C ABI returns its error codes, with a message, instead of
    launching on bad arguments: no cells, null arrays, an unknown method or
    layout, a method whose tables or stencil are not built, negative steps,
    an invalid lattice size or generator; the context stays usable."""
    say("argument errors")
    L = G.lib()
    g = O.Grid((8,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    c = G.Context(g.n, g.vmin, g.vmax, g.R)
    n = g.n_pnt
    f0 = (maxwellian(g.xi, 0.6, 1.0, [0.8, 0, 0]) + maxwellian(g.xi, 0.4, 0.5, [-0.8, 0, 0]))[None]
    f = G.DeviceArray.from_host(f0)
    out = [G.DeviceArray((1, n)) for _ in range(3)]
    ev = lambda method, fp, cells, layout=0, flags=G.FLAG_CORRECT: L.gdt_evaluate(
        c.h, method, fp, out[0].ptr, out[1].ptr, out[2].ptr, cells, layout, flags)
    E_ARG, E_STATE = -1, -3
    assert ev(G.METHOD_KOROBOV, f.ptr, 1) == E_STATE  # stencil not built
    assert ev(G.METHOD_MATRIX, f.ptr, 1) == E_STATE   # tables not built
    assert "not built" in L.gdt_last_error().decode()
    ne, ns = C.c_long(), C.c_long()
    assert L.gdt_korobov_gather_info(c.h, C.byref(ne), C.byref(ns)) == E_STATE  # stencil not built
    assert L.gdt_korobov_build(c.h, 3000, 1, 0, 0.25) == E_ARG      # M not a power of two
    assert L.gdt_korobov_build(c.h, 1 << 12, 1, 6, 0.25) == E_ARG   # even generator
    assert L.gdt_korobov_build(c.h, 1 << 12, 1, 0, -1.0) == E_ARG   # s0 <= 0
    c.korobov_build(1 << 12, seed=1, s0=0.25)
    assert ev(G.METHOD_KOROBOV, f.ptr, 0) == E_ARG
    assert ev(G.METHOD_KOROBOV, None, 1) == E_ARG
    assert ev(7, f.ptr, 1) == E_ARG
    assert "method 7" in L.gdt_last_error().decode()
    assert ev(G.METHOD_KOROBOV, f.ptr, 1, layout=5) == E_ARG
    assert ev(G.METHOD_KOROBOV, f.ptr, -3, flags=G.FLAG_CORRECT | G.FLAG_CELL_TILES) == E_ARG
    assert L.gdt_run_steps(c.h, G.METHOD_KOROBOV, f.ptr, 1, 0.02, -1, G.FLAG_CORRECT) == E_ARG
    # still usable: the same evaluation as the oracle
    assert ev(G.METHOD_KOROBOV, f.ptr, 1) == 0
    st = O.korobov_stencil(g, 1 << 12, c.korobov_info()["a"], 0.25)
    lo, q = O.korobov_collide(g, st, f0)
    qc, _, _ = O.correct_loss(g.xi, lo, q, 1)
    assert np.max(np.abs(out[2].get() - qc)) <= 1e-12 * np.max(np.abs(lo))
