"""Uneven batches (gdt_batch_*): cells of several velocity grids in persistent
work-list launches.  The contract is per-cell equality with non-batched runs,
bitwise, whatever the segment size, node blocking, tuning or graph replay."""
import numpy as np
import pytest

import gpu_determ_amd as G

from helpers import mixture

pytestmark = pytest.mark.gpu

VMAX = 4.8
DT = 0.01


def make_classes(sizes, M=1 << 16, seed=5):
    out = []
    for n in sizes:
        ctx = G.Context((n,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
        ctx.korobov_build(M, seed=seed, s0=0.25)
        out.append(ctx)
    return out


def cell_fs(ctx, cells, rng):
    return np.stack([mixture(ctx.xi, rng, comps=1 + (c % 3)) for c in range(cells)])


def single_runs(ctx, f0, steps, flags):
    """Each cell alone through gdt_run_steps (cells = 1)."""
    out = []
    for c in range(f0.shape[0]):
        d = G.DeviceArray.from_host(f0[c:c + 1])
        ctx.run_steps(G.METHOD_KOROBOV, d, 1, DT, steps, flags)
        ctx.sync()
        out.append(d.get()[0])
    return np.stack(out)


@pytest.fixture(scope="module")
def classes():
    return make_classes((12, 16, 20))


@pytest.mark.parametrize("seg,q", [(64, 2), (2, 4), (1, 8)])
def test_batch_steps_bitwise_equal_single(classes, seg, q):
    print("batch steps: seg %d q %d" % (seg, q), flush=True)
    rng = np.random.default_rng(11)
    cells = [3, 2, 2]
    f0 = [cell_fs(c, k, rng) for c, k in zip(classes, cells)]
    fd = [G.DeviceArray.from_host(x) for x in f0]
    b = G.Batch(classes, fd)
    b.config(seg, q)
    assert b.info()["segments"] == -(-sum(cells) // seg)
    flags = G.FLAG_CORRECT | G.FLAG_GRAPH
    b.run(DT, 2, flags)
    b.sync()
    for ctx, x0, d in zip(classes, f0, fd):
        ref = single_runs(ctx, x0, 2, G.FLAG_CORRECT)
        got = d.get()
        assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    b.close()


def test_batch_tune_keeps_f_and_q_output(classes):
    print("batch tune + Q output", flush=True)
    rng = np.random.default_rng(12)
    cells = [2, 3, 1]
    f0 = [cell_fs(c, k, rng) for c, k in zip(classes, cells)]
    fd = [G.DeviceArray.from_host(x) for x in f0]
    qd = [G.DeviceArray((k, c.n_pnt)) for c, k in zip(classes, cells)]
    b = G.Batch(classes, fd, qd)
    seg, q, ms = b.tune(G.FLAG_CORRECT)
    assert seg >= 1 and q in (2, 8, 32, 128) and ms > 0.0
    for x0, d in zip(f0, fd):  # tuning evaluates only
        assert np.array_equal(d.get(), x0)
    b.run(0.0, 1, G.FLAG_CORRECT)  # evaluation: Q out, f unchanged
    b.sync()
    for ctx, x0, d, qq in zip(classes, f0, fd, qd):
        assert np.array_equal(d.get(), x0)
        _, _, Qref = ctx.evaluate(G.METHOD_KOROBOV, G.DeviceArray.from_host(x0), x0.shape[0], G.FLAG_CORRECT)
        ctx.sync()
        assert np.array_equal(qq.get(), Qref.get())
    b.close()


def test_batch_eager_equals_graph(classes):
    print("batch eager vs graph", flush=True)
    rng = np.random.default_rng(13)
    cells = [2, 2, 2]
    f0 = [cell_fs(c, k, rng) for c, k in zip(classes, cells)]
    outs = []
    for flags in (G.FLAG_CORRECT, G.FLAG_CORRECT | G.FLAG_GRAPH):
        fd = [G.DeviceArray.from_host(x) for x in f0]
        b = G.Batch(classes, fd)
        b.config(4, 2)
        b.run(DT, 3, flags)
        b.sync()
        outs.append([d.get() for d in fd])
        b.close()
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


def test_korobov_set_tiles_and_batch_layout_check():
    """gdt_korobov_set_tiles: per-destination sums are the same bits for any
    tile limit, the corrected Q differs only in the moment sums' last bits,
    and returning to the default restores the default's bits.  A batch built
    before a re-tiling refuses to run until it is configured again, and then
    matches single runs at the new tiling bitwise."""
    print("korobov set tiles", flush=True)
    ctx = make_classes((16,))[0]
    rng = np.random.default_rng(21)
    f0 = cell_fs(ctx, 2, rng)
    fd = G.DeviceArray.from_host(f0)
    l1, q1 = ctx.collide_korobov(fd, 2)
    _, _, Qc1 = ctx.evaluate(G.METHOD_KOROBOV, fd, 2, G.FLAG_CORRECT)
    ctx.korobov_set_tiles(8)
    assert ctx.korobov_tiles()["tiles"] <= 8
    l2, q2 = ctx.collide_korobov(fd, 2)
    _, _, Qc2 = ctx.evaluate(G.METHOD_KOROBOV, fd, 2, G.FLAG_CORRECT)
    assert np.array_equal(q1.get(), q2.get()) and np.array_equal(l1.get(), l2.get())
    a, b_ = Qc1.get(), Qc2.get()
    assert np.max(np.abs(a - b_)) <= 1e-12 * np.max(np.abs(a))
    # a batch configured at 8 tiles, then the class re-tiled: refused until reconfigured
    bd = G.DeviceArray.from_host(f0)
    bt = G.Batch([ctx], [bd])
    bt.config(2, 2)
    ctx.korobov_set_tiles(64)  # the default limit
    with pytest.raises(RuntimeError):
        bt.run(DT, 1, G.FLAG_CORRECT)
    bt.config(2, 2)
    bt.run(DT, 1, G.FLAG_CORRECT)
    bt.sync()
    ref = single_runs(ctx, f0, 1, G.FLAG_CORRECT)
    assert np.array_equal(bd.get(), ref)
    _, _, Qc3 = ctx.evaluate(G.METHOD_KOROBOV, fd, 2, G.FLAG_CORRECT)
    assert np.array_equal(Qc3.get(), a)
    with pytest.raises(RuntimeError):
        ctx.korobov_set_tiles(0)


def test_batch_uneven_grids_vs_oracle():
    """The uneven-batch configuration's grid classes, {16, 20, 24, 32}^3 with
    2^20 lattice nodes each: a batched evaluation of every cell matches the
    FP64 oracle (stencil, collide and correction on the CPU) to 1e-12, and a
    batched explicit step equals single-cell steps bit for bit."""
    import oracle as O
    from helpers import rel_l2
    M = 1 << 20
    sizes = (16, 20, 24, 32)
    cells = [3, 2, 2, 2]
    print("batch {16,20,24,32}^3 at 2^20 vs oracle", flush=True)
    classes = make_classes(sizes, M=M, seed=7)
    rng = np.random.default_rng(17)
    f0 = [cell_fs(c, k, rng) for c, k in zip(classes, cells)]
    fd = [G.DeviceArray.from_host(x) for x in f0]
    qd = [G.DeviceArray((k, c.n_pnt)) for c, k in zip(classes, cells)]
    b = G.Batch(classes, fd, qd)
    b.config(4, 8)
    b.run(0.0, 1, G.FLAG_CORRECT)  # evaluation: Q out, f unchanged
    b.sync()
    for n, ctx, x0, q in zip(sizes, classes, f0, qd):
        g = O.Grid((n,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
        st = O.korobov_stencil(g, M, ctx.korobov_info()["a"], 0.25)
        lo, q0 = O.korobov_collide(g, st, x0)
        oq, _, bad = O.correct_loss(g.xi, lo, q0, x0.shape[0])
        assert bad == 0
        err = rel_l2(q.get(), oq)
        print("  %d^3: %d cells, rel L2 vs oracle %.3g" % (n, x0.shape[0], err), flush=True)
        assert err < 1e-12
    b.run(DT, 1, G.FLAG_CORRECT)
    b.sync()
    for ctx, x0, d in zip(classes, f0, fd):
        assert np.array_equal(d.get(), single_runs(ctx, x0, 1, G.FLAG_CORRECT))
    b.close()


def test_batch_cell_tiles_mode():
    """Batches in cell tiles (GDT_FLAG_CELL_TILES, and the mixed form): the
    uneven-batch grid classes at 2^20 nodes, with a class of 70 cells (two
    tiles, one ragged).  FP64 tiles match the oracle to 1e-12, the mixed form
    to 1e-5; an explicit step equals single-cell steps with the same flags bit
    for bit; graph replay equals eager launches."""
    import oracle as O
    from helpers import rel_l2
    M = 1 << 20
    sizes = (16, 20, 24, 32)
    cells = [70, 3, 2, 3]
    print("batch in cell tiles", flush=True)
    classes = make_classes(sizes, M=M, seed=7)
    rng = np.random.default_rng(19)
    f0 = [cell_fs(c, k, rng) for c, k in zip(classes, cells)]
    for fl, tol in ((G.FLAG_CELL_TILES, 1e-12), (G.FLAG_MIXED, 1e-5)):
        fd = [G.DeviceArray.from_host(x) for x in f0]
        qd = [G.DeviceArray((k, c.n_pnt)) for c, k in zip(classes, cells)]
        b = G.Batch(classes, fd, qd)
        b.run(0.0, 1, G.FLAG_CORRECT | fl)
        b.sync()
        for n, ctx, x0, q in zip(sizes, classes, f0, qd):
            g = O.Grid((n,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
            st = O.korobov_stencil(g, M, ctx.korobov_info()["a"], 0.25)
            k = min(x0.shape[0], 4)
            lo, q0 = O.korobov_collide(g, st, x0[:k])
            oq, _, _ = O.correct_loss(g.xi, lo, q0, k)
            err = rel_l2(q.get()[:k], oq)
            print("  flags %d %d^3: rel L2 vs oracle %.3g" % (fl, n, err), flush=True)
            assert err < tol
        b.run(DT, 1, G.FLAG_CORRECT | fl | G.FLAG_GRAPH)
        b.sync()
        for ctx, x0, d in zip(classes, f0, fd):
            assert np.array_equal(d.get(), single_runs(ctx, x0, 1, G.FLAG_CORRECT | fl))
        b.run(DT, 1, G.FLAG_CORRECT | fl)  # eager after the graph: the same second step
        b.sync()
        e1 = [d.get() for d in fd]
        fd2 = [G.DeviceArray.from_host(x) for x in f0]
        b2 = G.Batch(classes, fd2)
        b2.run(DT, 2, G.FLAG_CORRECT | fl | G.FLAG_GRAPH)
        b2.sync()
        for u, v in zip(e1, fd2):
            assert np.array_equal(u, v.get())
        b.close()
        b2.close()


def test_batch_tune_cell_tiles():
    """The autotuner in cell-tile mode picks a launch shape from its candidates
    (cells per chunk 4, 8 or 16 tiles of 64; gather grid 2048 or 4096), leaves
    f unchanged, and the results after tuning equal those before, bit for
    bit, eager and as graph replay."""
    print("batch tune, cell tiles", flush=True)
    classes = make_classes((16, 20), M=1 << 16, seed=3)
    rng = np.random.default_rng(23)
    cells = [150, 9]
    f0 = [cell_fs(c, k, rng) for c, k in zip(classes, cells)]
    for fl in (G.FLAG_CELL_TILES, G.FLAG_MIXED):
        fd = [G.DeviceArray.from_host(x) for x in f0]
        qd = [G.DeviceArray((k, c.n_pnt)) for c, k in zip(classes, cells)]
        b = G.Batch(classes, fd, qd)
        b.run(0.0, 1, G.FLAG_CORRECT | fl)
        b.sync()
        before = [q.get() for q in qd]
        k, gb, ms = b.tune(G.FLAG_CORRECT | fl)
        print("  flags %d: cells per chunk %d, gather grid %d, %.3f ms" % (fl, k, gb, ms), flush=True)
        assert k in (256, 512, 1024) and gb in (2048, 4096) and ms > 0.0
        for x0, d in zip(f0, fd):
            assert np.array_equal(d.get(), x0)
        for q in qd:
            q.zero()
        b.run(0.0, 1, G.FLAG_CORRECT | fl | G.FLAG_GRAPH)
        b.sync()
        for q, q0 in zip(qd, before):
            assert np.array_equal(q.get(), q0)
        for q in qd:
            q.zero()
        b.run(0.0, 1, G.FLAG_CORRECT | fl)
        b.sync()
        for q, q0 in zip(qd, before):
            assert np.array_equal(q.get(), q0)
        b.close()
