#!/usr/bin/env python3
"""Collision-integral throughput on one MI355X.

Workload (the headline configuration): hard-sphere molecules on a 32^3
velocity grid cut to the sphere |v| < vmax = 4.8 (17,256 nodes), FP64, plain
Korobov/projection method with 2^21 lattice nodes, one spatial cell whose f is
a seeded mixture of 3 Maxwellians (n ~ U(0.2,1), T ~ U(0.3,1.5), |u| <= 1.2),
explicit steps df/dt = Q(f,f) with dt = 0.02 and the conservation correction
every step.  One step = one corrected Q(f,f) evaluation + the Euler update:
three kernels launched onto the stream (--graph: replayed as a hipGraph).
Metric: Q(f,f) evaluations per second.

    python bench.py [--gpus 1] [--steps 200] [--warmup 300] [--graph]

Prints ONE JSON line.  The roofline object reports the dominant kernel's
algorithmic bytes over its hipEvent-timed average duration (a separate
instrumented pass after the timed region); cpu_baseline times the FP64 CPU
oracle (oracle/, a C restatement with OpenMP over lattice nodes and over
destination velocities, compiled for the host with -march=native) on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector), datasheet
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp32": 157.3}  # MI355X dense matrix peaks, datasheet
N_GRID = 32
VMAX = 4.8
M_NODES = 1 << 21
S0 = 0.25  # unit-diameter spheres: differential cross-section 1/4 per steradian
DT = 0.02
SEED = 2026


def mixture(xi, rng, comps=3):
    f = np.zeros(xi.shape[0])
    for _ in range(comps):
        n = rng.uniform(0.2, 1.0)
        T = rng.uniform(0.3, 1.5)
        while True:
            u = rng.uniform(-1.2, 1.2, 3)
            if u @ u <= 1.44:
                break
        f += n * (2 * np.pi * T) ** -1.5 * np.exp(-np.sum((xi - u) ** 2, axis=1) / (2 * T))
    return f


def moments(xi, f, d3v):
    e = np.sum(xi * xi, axis=1)
    return np.array([f.sum(), (f * xi[:, 0]).sum(), (f * xi[:, 1]).sum(), (f * xi[:, 2]).sum(), (f * e).sum()]) * d3v


def entropy(f, d3v):
    fp = np.maximum(f, 1e-300)
    return float(np.sum(fp * np.log(fp)) * d3v)


def batch_fs(xi, cells, rng):
    """Seeded mixtures of 1-3 Maxwellians for `cells` cells (n ~ U(0.2,1),
    T ~ U(0.3,1.5), |u| <= 1.2), as [cells][nodes]."""
    comps = rng.integers(1, 4, cells)
    f = np.zeros((cells, xi.shape[0]))
    for k in range(3):
        n = rng.uniform(0.2, 1.0, cells) * (comps > k)
        T = rng.uniform(0.3, 1.5, cells)
        u = rng.uniform(-1.0, 1.0, (cells, 3))
        u *= (1.2 * rng.uniform(0, 1, cells) ** (1 / 3) / np.maximum(np.linalg.norm(u, axis=1), 1e-12))[:, None]
        d2 = np.sum((xi[None, :, :] - u[:, None, :]) ** 2, axis=2)
        f += (n * (2 * np.pi * T) ** -1.5)[:, None] * np.exp(-d2 / (2 * T[:, None]))
    return f


def uneven_batch(G, cells=10000, steps=3, seed=SEED):
    """Uneven batch: cells on grids drawn from {16,20,24,32}^3 (vmax 4.8,
    sphere cut), one 2^20-node stencil per grid size, every cell corrected
    each step, work-list launches captured as one hipGraph (and, beside it,
    launched straight onto the stream), launch shape from the autotuner.
    Returns the measured figures."""
    rng = np.random.default_rng(seed + 4)
    sizes = (16, 20, 24, 32)
    pick = rng.integers(0, len(sizes), cells)
    ctxs, fds = [], []
    t0 = time.perf_counter()
    for i, n in enumerate(sizes):
        ctx = G.Context((n,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
        ctx.korobov_build(1 << 20, seed=seed, s0=S0)
        ctxs.append(ctx)
        fds.append(G.DeviceArray.from_host(batch_fs(ctx.xi, int(np.sum(pick == i)), rng)))
    setup_s = time.perf_counter() - t0
    b = G.Batch(ctxs, fds)
    t0 = time.perf_counter()
    seg, q, tune_ms = b.tune(G.FLAG_CORRECT)
    tune_s = time.perf_counter() - t0
    flags = G.FLAG_CORRECT | G.FLAG_GRAPH
    b.run(DT, 1, flags)  # capture + warm-up
    b.sync()
    t0 = time.perf_counter()
    b.run(DT, steps, flags)
    b.sync()
    el = time.perf_counter() - t0
    # the same steps launched straight onto the stream (no graph), beside it
    b.run(DT, 1, G.FLAG_CORRECT)
    b.sync()
    t0 = time.perf_counter()
    b.run(DT, steps, G.FLAG_CORRECT)
    b.sync()
    el_eager = time.perf_counter() - t0
    # the same steps with every class's cells in cell tiles (FP64: the same
    # plain method; mixed: FP32 node terms), graph replay
    tiles = {}
    for name, fl in (("fp64_tiles", G.FLAG_CELL_TILES), ("mixed_fp32_tiles", G.FLAG_MIXED)):
        tk, tg, tms = b.tune(G.FLAG_CORRECT | fl)  # launch shape of the classes' tiles
        fl = G.FLAG_CORRECT | G.FLAG_GRAPH | fl
        b.run(DT, 1, fl)  # scratch, capture, warm-up
        b.sync()
        t0 = time.perf_counter()
        b.run(DT, steps, fl)
        b.sync()
        e = time.perf_counter() - t0
        tiles[name] = {"evals_per_s": round(cells * steps / e, 1), "ms_per_step": round(1e3 * e / steps, 3),
                       "tuned": {"cells_per_chunk": tk, "gather_blocks": tg, "best_eval_ms": round(tms, 3)}}
    # roofline per the algorithmic bytes of one evaluation: f in + Q out (16 N)
    # plus the class stencil over the cells sharing one pass of it -- one
    # cell for the work lists (16-byte node records, 8-byte entries of the
    # group CSRs, the first lambda piece of each chunk of 16 nodes), 64 cells
    # for the tiles (node records and FP64 (r, w) 32 B per node, plus
    # the FP32 (r, w) of the mixed form; 16-byte entries of the tile CSR,
    # every entry but each node's alpha one)
    alg_wl = alg_t = alg_m = 0.0
    for i, c in enumerate(ctxs):
        k = int(np.sum(pick == i))
        inf, gi = c.korobov_info(), c.korobov_gather_info()
        tent = inf["entries"] - inf["nodes"]
        alg_wl += k * (16.0 * c.n_pnt + 16.0 * inf["nodes"] + 8.0 * gi["entries"] + 4.0 * ((inf["nodes"] + 15) // 16))
        alg_t += k * (16.0 * c.n_pnt + (32.0 * inf["nodes"] + 16.0 * tent) / 64.0)
        alg_m += k * (16.0 * c.n_pnt + (40.0 * inf["nodes"] + 16.0 * tent) / 64.0)

    def roof(alg, ms):
        gbs = alg / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "alg_bytes_per_step": int(alg), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}

    tiles["fp64_tiles"]["roofline"] = roof(alg_t, tiles["fp64_tiles"]["ms_per_step"])
    tiles["mixed_fp32_tiles"]["roofline"] = roof(alg_m, tiles["mixed_fp32_tiles"]["ms_per_step"])
    out = {"cells": cells, "grids": {"%d^3" % n: int(np.sum(pick == i)) for i, n in enumerate(sizes)},
           "korobov_M": 1 << 20, "steps": steps, "evals_per_s": round(cells * steps / el, 1),
           "roofline": roof(alg_wl, 1e3 * el / steps), "cell_tiles": tiles,
           "ms_per_step": round(1e3 * el / steps, 3), "eager_evals_per_s": round(cells * steps / el_eager, 1),
           "tuned": {"seg_cells": seg, "nodes_per_thread": q, "best_eval_ms": round(tune_ms, 3),
                     "gather_tiles_per_group": {"%d^3" % n: c.korobov_tiles()["tiles"] for n, c in zip(sizes, ctxs)}},
           "setup_s": round(setup_s, 2), "tune_s": round(tune_s, 2)}
    b.close()
    return out


def batched_matrix(G, cells=8192, ref_cells=256, seed=SEED):
    """Batched analytic-matrix form on the matrix cores: 16^3 box (4,096
    nodes), `cells` perturbed Maxwellians (x (1 + 0.05 N(0,1)), clipped
    positive, renormalised), corrected Q per cell, in each operand mode; error
    against the FP64 GPU matrix method (itself pinned to the oracle) on the
    first `ref_cells` cells, and that method's own rate."""
    rng = np.random.default_rng(seed + 3)
    L = 4.8
    ctx = G.Context((16,) * 3, (-L,) * 3, (L,) * 3, -1.0)
    ctx.build_tables(G.TABLES_FP64)
    xi = ctx.xi
    n = rng.uniform(0.5, 1.0, cells)
    T = rng.uniform(0.6, 1.4, cells)
    u = rng.uniform(-0.8, 0.8, (cells, 3))
    f = np.empty((cells, xi.shape[0]))
    for c in range(cells):
        d2 = np.sum((xi - u[c]) ** 2, axis=1)
        m = n[c] * (2 * np.pi * T[c]) ** -1.5 * np.exp(-d2 / (2 * T[c]))
        p_ = np.clip(m * (1 + 0.05 * rng.standard_normal(m.shape)), 1e-300, None)
        f[c] = p_ * (m.sum() / p_.sum())
    fd = G.DeviceArray.from_host(f)
    fr = G.DeviceArray.from_host(f[:ref_cells])
    t0 = time.perf_counter()
    _, _, Qr = ctx.evaluate(G.METHOD_MATRIX, fr, ref_cells, G.FLAG_CORRECT)
    ctx.sync()
    el = time.perf_counter() - t0
    Qr = Qr.get()
    out = {"grid": "16^3 box", "cells": cells, "fp64_gpu_matrix": {"evals_per_s": round(ref_cells / el, 1),
                                                                   "cells_timed": ref_cells}}
    for name, mode in (("bf16_split", G.GEMM_BF16_SPLIT), ("fp32", G.GEMM_F32), ("bf16", G.GEMM_BF16)):
        nu, gain, Q = ctx.matrix_mfma(mode, fd, cells, G.FLAG_CORRECT)  # warm-up (operands built once)
        ctx.sync()
        best = 1e30
        for _ in range(2):
            t0 = time.perf_counter()
            ctx.matrix_mfma(mode, fd, cells, G.FLAG_CORRECT)
            ctx.sync()
            best = min(best, time.perf_counter() - t0)
        Qm = Q.get()[:ref_cells]
        err = float(np.linalg.norm(Qm - Qr) / np.linalg.norm(Qr))
        errs = np.linalg.norm(Qm - Qr, axis=1) / np.linalg.norm(Qr, axis=1)
        # matrix-core work per cell: the gain GEMMs (512 g x 512 n1 per parity
        # class times its s columns: 31^3 over the 8 classes) and the loss GEMM
        # (4,096 x 4,096).  As executed, a tile of 128 columns at s_x runs
        # 64 (kx + 1) rows and n1 each, kx = (min(s_x, 30 - s_x) - p_x) / 2
        # (past that T is zero and no G row is read), over two tiles per s_x;
        # the split mode issues three products per product
        prods = 3 if mode == G.GEMM_BF16_SPLIT else 1
        alg_cell = 2 * 512 * 512 * 31 ** 3 + 2 * 4096 * 4096
        exe_gain = sum(2 * 2 * (64 * ((min(sx, 30 - sx) - px) // 2 + 1)) ** 2 * 128
                       for p in range(8) for px in [(p >> 2) & 1] for sx in range(px, 31, 2))
        exe_cell = prods * (exe_gain + 2 * 4096 * 4096)
        peak = MFMA_PEAK_TFLOPS["fp32" if mode == G.GEMM_F32 else "bf16"]
        exe_tf = exe_cell * cells / best / 1e12
        out[name] = {"evals_per_s": round(cells / best, 1), "ms": round(1e3 * best, 2),
                     "rel_l2_Q_vs_fp64": err, "max_cell_rel_l2": float(errs.max()),
                     "speedup_vs_fp64_gpu": round(cells / best / out["fp64_gpu_matrix"]["evals_per_s"], 1),
                     "mfma": {"alg_tflop": round(alg_cell * cells / 1e12, 1),
                              "executed_tflop": round(exe_cell * cells / 1e12, 1),
                              "executed_tflops": round(exe_tf, 1), "peak_tflops": peak,
                              "frac": round(exe_tf / peak, 4)}}
        del nu, gain, Q
    return out


def transport_step(G, shape=(16, 16, 16), steps=20, seed=SEED):
    """Spatial transport (first-order upwind, periodic) of f on the 24^3
    velocity grid over a box of `shape` cells: HBM rate of its algorithmic
    bytes (f in + f out, 16 B per cell and velocity node)."""
    rng = np.random.default_rng(seed + 6)
    ctx = G.Context((24,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
    cells = int(np.prod(shape))
    f = np.stack([mixture(ctx.xi, rng, comps=1) for _ in range(cells)])
    a, b = G.DeviceArray.from_host(f), G.DeviceArray((cells, ctx.n_pnt))
    h = (1.0, 1.0, 1.0)
    dt = 0.9 / np.max(np.abs(ctx.xi).sum(axis=1))
    for _ in range(3):
        ctx.transport(a, shape, h, dt, out=b)
        a, b = b, a
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.transport(a, shape, h, dt, out=b)
        a, b = b, a
    ctx.sync()
    el = (time.perf_counter() - t0) / steps
    alg = 16.0 * cells * ctx.n_pnt
    return {"grid": "24^3", "nodes": ctx.n_pnt, "cells": "x".join(str(x) for x in shape), "us_per_step": round(1e6 * el, 2),
            "alg_bytes": int(alg), "gbs": round(alg / el / 1e9, 1), "frac_of_hbm": round(alg / el / 1e9 / HBM_PEAK_GBS, 3)}


def plain_tiles(G, ctx, fd, cells, chunk=16384):
    """Corrected Q of every cell by the plain solver in FP64 cell tiles (the
    many-cell plain method, within 1e-12 of the oracle), chunk by chunk."""
    n = ctx.n_pnt
    Q = G.DeviceArray((cells, n))
    ch = min(chunk, cells)
    nu, g = G.DeviceArray((ch, n)), G.DeviceArray((ch, n))
    for c0 in range(0, cells, ch):
        k = min(ch, cells - c0)
        G.check(G.lib().gdt_evaluate(ctx.h, G.METHOD_KOROBOV, fd.ptr + 8 * c0 * n, nu.ptr, g.ptr, Q.ptr + 8 * c0 * n,
                                     k, G.LAYOUT_CELL_MAJOR, G.FLAG_CORRECT | G.FLAG_CELL_TILES))
    ctx.sync()
    return Q


def train_surrogates(G, ctx, rng, train_cells, epochs, hidden, lr, seed=SEED):
    """Two networks (different seeds) trained on `train_cells` (f, Q) pairs
    of seeded 1-3 Maxwellian mixtures (T in [0.3, 1.5], |u| <= 1) labelled by
    the plain solver; the second is the first's partner in the residual
    check."""
    t0 = time.perf_counter()
    ftr = ctx.init_mixture(mixture_params(rng, train_cells, umax=1.0))
    Qtr = plain_tiles(G, ctx, ftr, train_cells)
    gen_s = time.perf_counter() - t0
    nets, losses = [], []
    t0 = time.perf_counter()
    for k in range(2):
        s = G.Surrogate(ctx, hidden=hidden, seed=seed + 101 * k)
        losses.append(s.train(ftr, Qtr, epochs=epochs, batch=256, lr=lr))
        nets.append(s)
        progress("surrogate network %d of 2 trained (%d pairs, %d epochs)" % (k + 1, train_cells, epochs))
    train_s = time.perf_counter() - t0
    nets[0].set_partner(nets[1])
    info = {"train_pairs": train_cells, "epochs": epochs, "hidden": hidden, "lr": lr, "params_per_network": nets[0].params,
            "networks": 2, "final_loss": losses, "gen_s": round(gen_s, 2), "train_s": round(train_s, 2)}
    return nets, info


def surrogate_eval(G, ctx, nets, fin, cells, Qplain, plain_evals, threshold=1e-4):
    """Surrogate with the plain-solver fallback at `threshold` on `cells`
    cells: throughput (network pair + projection + fallback), the network's
    error against the plain solver (which judges every cell), and per
    threshold the fraction of cells the gate keeps and the true error of the
    kept cells -- for the pair's disagreement gate and for the projection
    residual of one network alone."""
    s = nets[0]
    s.infer(fin, threshold)  # warm-up
    t0 = time.perf_counter()
    Q, fb, res = s.infer(fin, threshold)
    infer_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    Qn, _, res_pair = s.infer(fin, 1e30)  # network pair + projection, no fallback
    nn_s = time.perf_counter() - t0
    s.set_partner(None)
    Q1, _, res_one = s.infer(fin, 1e30)
    s.set_partner(nets[1])
    qp, qn, q1 = Qplain.get(), Qn.get(), Q1.get()
    pn = np.linalg.norm(qp, axis=1)
    big = pn > 1e-3 * np.median(pn)  # cells away from equilibrium (Q not ~0)
    err = np.linalg.norm(qn - qp, axis=1)
    rel = err[big] / pn[big]

    def curve(r, qq):
        out = []
        for t in (0.3, 0.1, 0.03, 0.01, 3e-3, 1e-3, 1e-4):
            keep = r <= t
            kb = keep[big]
            e = np.linalg.norm(qq - qp, axis=1)[big] / pn[big]
            out.append({"threshold": t, "kept_frac": round(float(keep.mean()), 5),
                        "kept_max_true_rel_l2": float(e[kb].max()) if kb.any() else None,
                        "kept_median_true_rel_l2": float(np.median(e[kb])) if kb.any() else None})
        return out

    lr = np.log(np.maximum(res_pair[big], 1e-30))
    le = np.log(np.maximum(rel, 1e-30))
    return {"cells": cells, "threshold": threshold, "fallback_cells": int(fb.sum()),
            "evals_per_s": round(cells / infer_s, 1), "network_only_evals_per_s": round(cells / nn_s, 1),
            "plain_evals_per_s": round(plain_evals, 1),
            "speedup_vs_plain": round(cells / infer_s / plain_evals, 3),
            "result_rel_l2_vs_plain": float(np.linalg.norm(Q.get() - qp) / np.linalg.norm(qp)),
            "network_rel_l2_vs_plain": float(np.linalg.norm(qn - qp) / np.linalg.norm(qp)),
            "network_cell_rel_l2": {"median": float(np.median(rel)), "p10": float(np.quantile(rel, 0.1)),
                                    "min": float(rel.min()), "cells": int(big.sum())},
            "gate_log_corr_with_true_error": {"pair": float(np.corrcoef(lr, le)[0, 1]),
                                              "one_network": float(np.corrcoef(
                                                  np.log(np.maximum(res_one[big], 1e-30)), le)[0, 1])},
            "curve_pair_gate": curve(res_pair, qn), "curve_one_network_gate": curve(res_one, q1),
            "conservation_residual": conservation_residual(ctx.xi, Q.get(), ctx.d3v)}


def surrogate_form(G, train_cells=200000, infer_cells=16384, epochs=40, hidden=512, lr=5e-4, threshold=1e-4,
                   seed=SEED):
    """Learned surrogate on 24^3 (vmax 4.8, sphere cut, 7,208 nodes, Korobov
    2^20): two networks trained on `train_cells` plain-solver pairs, inference
    on `infer_cells` new cells with the exact projection and the plain-solver
    fallback above `threshold`."""
    rng = np.random.default_rng(seed + 5)
    ctx = G.Context((24,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
    ctx.korobov_build(1 << 20, seed=seed, s0=S0)
    nets, info = train_surrogates(G, ctx, rng, train_cells, epochs, hidden, lr, seed)
    fin = ctx.init_mixture(mixture_params(rng, infer_cells, umax=1.0))
    Qplain = plain_tiles(G, ctx, fin, infer_cells)
    t0 = time.perf_counter()
    plain_tiles(G, ctx, fin, infer_cells)
    plain_evals = infer_cells / (time.perf_counter() - t0)
    out = {"grid": "24^3", "nodes": ctx.n_pnt}
    out.update(info)
    out.update(surrogate_eval(G, ctx, nets, fin, infer_cells, Qplain, plain_evals, threshold))
    for s_ in nets:
        s_.close()
    return out


def mixture_params(rng, cells, umax=1.2):
    """(n, T, u) of 1-3 Maxwellian components per cell: n ~ U(0.2,1), T ~
    U(0.3,1.5), u uniform in the ball |u| <= umax; absent components have n = 0."""
    comps = rng.integers(1, 4, cells)
    par = np.zeros((cells, 3, 5))
    for j in range(3):
        u = rng.uniform(-1.0, 1.0, (cells, 3))
        u *= (umax * rng.uniform(0, 1, cells) ** (1 / 3) / np.maximum(np.linalg.norm(u, axis=1), 1e-12))[:, None]
        par[:, j, 0] = rng.uniform(0.2, 1.0, cells) * (comps > j)
        par[:, j, 1] = rng.uniform(0.3, 1.5, cells)
        par[:, j, 2:] = u
    return par


def conservation_residual(xi, Q, d3v):
    """max over cells of |sum Q psi| / sum |Q| |psi| (psi = 1, xi, |xi|^2)."""
    e = np.sum(xi * xi, axis=1)
    psi = np.stack([np.ones(len(e)), xi[:, 0], xi[:, 1], xi[:, 2], e])
    m = np.abs(Q @ psi.T) * d3v
    sc = (np.abs(Q) @ np.abs(psi.T)) * d3v
    return float(np.max(m / np.maximum(sc, 1e-300)))


def forms_32cube(G, ctx, single_evals, cells=2048, steps=3, oracle_cells=4, seed=SEED, surrogate=True,
                 surrogate_train=100000, surrogate_epochs=20):
    """The metric's comparison on one set of cells at 32^3 (vmax 4.8, 17,256
    nodes, 2^21 lattice nodes, s0 = 0.25): `cells` seeded mixtures of 1-3
    Maxwellians.  Throughput unit: one corrected Q(f,f) evaluation with the
    Euler update (dt = 0.02) of one cell, as the headline.  Forms:
      plain_fp64_single   the headline (one cell per step, the grouped kernels)
      plain_fp64_grouped  the same kernels on all cells at once
      plain_fp64_tiles    FP64 cell tiles (64 cells per row, same node terms)
      mixed_fp32_tiles    cell tiles with f and the node terms in FP32, sums
                          and the 5x5 correction in FP64
      surrogate_with_fallback  two networks trained on `surrogate_train`
                          plain-solver pairs of 32^3 cells, the exact
                          projection, and the plain solver (FP64 tiles) for
                          every cell whose residual exceeds 1e-4
    Accuracy on the unstepped cells: corrected Q against the FP64 tiles over
    every cell (the FP64 tiles are within 1e-12 of the oracle; checked here on
    `oracle_cells` cells too), the worst cell, and the conservation residual
    of every form."""
    rng = np.random.default_rng(seed + 7)
    n = ctx.n_pnt
    f0 = ctx.init_mixture(mixture_params(rng, cells))
    fh = f0.get()
    forms = {"plain_fp64_grouped": G.FLAG_CORRECT, "plain_fp64_tiles": G.FLAG_CORRECT | G.FLAG_CELL_TILES,
             "mixed_fp32_tiles": G.FLAG_CORRECT | G.FLAG_MIXED}
    out = {"workload": "%d cells, 32^3 sphere cut (17,256 nodes), 2^21 lattice nodes, seeded 1-3 Maxwellian "
                       "mixtures; unit = one corrected evaluation + Euler step of one cell" % cells,
           "cells": cells, "steps_timed": steps,
           "plain_fp64_single": {"evals_per_s": round(single_evals, 1), "note": "the headline value"}}
    Qs = {}
    nu_o, gain_o, gain_ref = G.DeviceArray((cells, n)), G.DeviceArray((cells, n)), G.DeviceArray((cells, n))
    for name, fl in forms.items():
        Q = G.DeviceArray((cells, n))
        G.check(G.lib().gdt_evaluate(ctx.h, G.METHOD_KOROBOV, f0.ptr, nu_o.ptr,
                                     (gain_ref if name == "plain_fp64_tiles" else gain_o).ptr, Q.ptr, cells,
                                     G.LAYOUT_CELL_MAJOR, fl))
        ctx.sync()
        Qs[name] = Q.get()
        fd = G.DeviceArray.from_host(fh)
        ctx.run_steps(G.METHOD_KOROBOV, fd, cells, DT, 1, fl)  # warm-up (scratch, first launches)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.run_steps(G.METHOD_KOROBOV, fd, cells, DT, steps, fl)
        ctx.sync()
        el = time.perf_counter() - t0
        ctx.kernel_stats_reset()
        ctx.run_steps(G.METHOD_KOROBOV, fd, cells, DT, 1, fl | G.FLAG_TIME)
        ctx.sync()
        ks = {}
        for k in ("kor_node", "kor_gather", "kor_apply", "ct_pack", "ct_node", "ct_gather", "ct_moments", "ct_solve",
                  "ct_apply"):
            ms, cnt = ctx.kernel_stats(k)
            if cnt:
                ks[k] = round(1e3 * ms / cells, 4)
        out[name] = {"evals_per_s": round(cells * steps / el, 1), "kernel_us_per_cell": ks}
        del fd
    # per-cell errors: relative to the cell's Q where Q is not negligible
    # against its loss term (single-Maxwellian cells have Q = 0 to rounding),
    # and for every cell relative to the loss term's norm |f nu|
    ref = Qs["plain_fp64_tiles"]
    lref = (gain_ref.get() - ref)  # gain - Q = f nu (the loss term) of the FP64 tiles
    refn, ln = np.linalg.norm(ref, axis=1), np.linalg.norm(lref, axis=1)
    neq = refn > 1e-6 * ln
    for name, Q in Qs.items():
        d = np.linalg.norm(Q - ref, axis=1)
        out[name].update({"rel_l2_Q_vs_fp64": float(np.linalg.norm(Q - ref) / np.linalg.norm(ref)),
                          "max_cell_rel_l2": float((d[neq] / refn[neq]).max()),
                          "max_cell_err_vs_loss": float((d / ln).max()),
                          "conservation_residual": conservation_residual(ctx.xi, Q, ctx.d3v)})
    out["cells_near_equilibrium"] = int((~neq).sum())
    # algorithmic bytes of one evaluation: f in and f out (16 N) plus the
    # stencil over the cells that share one pass of it -- one cell for the
    # grouped kernels (16-byte node records, 8-byte entries of the group
    # CSRs, the first lambda piece of each chunk of 16 nodes), 64 for the
    # tiles (node records and FP64 (r, w), 32 B per node, plus the FP32
    # (r, w) of the mixed form; 16-byte entries of the tile CSR, which holds
    # every entry but the alpha one of each node: the node kernel sums those)
    inf, gi = ctx.korobov_info(), ctx.korobov_gather_info()
    tent = inf["entries"] - inf["nodes"]
    alg = {"plain_fp64_grouped": 16.0 * n + 16.0 * inf["nodes"] + 8.0 * gi["entries"] + 4.0 * ((inf["nodes"] + 15) // 16),
           "plain_fp64_tiles": 16.0 * n + (32.0 * inf["nodes"] + 16.0 * tent) / 64.0,
           "mixed_fp32_tiles": 16.0 * n + (40.0 * inf["nodes"] + 16.0 * tent) / 64.0}
    for name, b_ in alg.items():
        gbs = b_ * out[name]["evals_per_s"] / 1e9
        out[name]["roofline"] = {"bound": "hbm", "alg_bytes_per_eval": int(b_), "achieved": round(gbs, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
    # FP64 work of the FP64 tiles' node kernel per cell (the one-cell node
    # kernel's operation count: 54 per node with r != 0, 8 with r = 0), over
    # its event-timed time per cell
    nz = (inf["entries"] - 4 * inf["nodes"]) // 2
    node_flop = 54 * nz + 8 * (inf["nodes"] - nz)
    us = out["plain_fp64_tiles"]["kernel_us_per_cell"].get("ct_node")
    if us:
        tf = node_flop / (us * 1e-6) / 1e12
        out["plain_fp64_tiles"]["ct_node_fp64"] = {"flop_per_cell": node_flop, "tflops": round(tf, 2),
                                                   "peak_tflops": FP64_PEAK_TFLOPS,
                                                   "frac": round(tf / FP64_PEAK_TFLOPS, 4)}
    for name in ("plain_fp64_grouped", "plain_fp64_tiles", "mixed_fp32_tiles"):
        e = out[name]["evals_per_s"]
        out[name]["speedup_vs_plain_fp64_single"] = round(e / single_evals, 2)
        out[name]["speedup_vs_plain_fp64_grouped"] = round(e / out["plain_fp64_grouped"]["evals_per_s"], 2)
        out[name]["speedup_vs_plain_fp64_tiles"] = round(e / out["plain_fp64_tiles"]["evals_per_s"], 2)
    if surrogate:
        # surrogate with the plain-solver fallback on the same cells (the
        # plain solver of the fallback: the FP64 tiles)
        srng = np.random.default_rng(seed + 8)
        # at 32^3 a rate of 1e-3 diverges; 3e-4 over 100k pairs reaches ~0.26
        # relative error, 256 hidden units as well as 512
        nets, info = train_surrogates(G, ctx, srng, surrogate_train, surrogate_epochs, 256, 3e-4, seed)
        sg = {"train": info}
        sg.update(surrogate_eval(G, ctx, nets, f0, cells, G.DeviceArray.from_host(ref),
                                 out["plain_fp64_tiles"]["evals_per_s"], 1e-4))
        sg["speedup_vs_plain_fp64_single"] = round(sg["evals_per_s"] / single_evals, 2)
        out["surrogate_with_fallback"] = sg
        for s_ in nets:
            s_.close()
    if oracle_cells:
        import oracle as O
        g = O.Grid((N_GRID,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
        st = O.korobov_stencil(g, M_NODES, ctx.korobov_info()["a"], S0)
        lo, q = O.korobov_collide(g, st, fh[:oracle_cells])
        oq, _, _ = O.correct_loss(g.xi, lo, q, oracle_cells)
        out["oracle_check"] = {"cells": oracle_cells,
                               "plain_fp64_tiles_rel_l2": float(np.linalg.norm(ref[:oracle_cells] - oq) / np.linalg.norm(oq)),
                               "mixed_fp32_tiles_rel_l2": float(np.linalg.norm(Qs["mixed_fp32_tiles"][:oracle_cells] - oq)
                                                                / np.linalg.norm(oq))}
    return out


def profiled_traffic():
    """HBM bytes per launch per kernel, measured with rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy-kernel calibration) and
    recorded in profiles/; empty if the summary is absent."""
    path = os.path.join(ROOT, "profiles", "korobov_step_32cube_2e21.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic":
                out[parts[1]] = int(parts[2])
    return out


def cpu_baseline(xi, f0, n_pnt, d3v, budget_s=15.0):
    """FP64 oracle on a bounded number of evaluations of the same workload,
    compiled on this host with -march=native (falls back to the prebuilt
    x86-64-v3 library if the host compiler fails) and run with OpenMP on the
    threads OMP_NUM_THREADS allows.  Stencil and destination lists once
    (untimed), then corrected Q(f,f) + Euler steps in two forms with the same
    bits: the serial scatter after parallel node values, and the gather over
    destinations (parallel over velocities, the one-cell parallel form).
    Each form: one warm-up, then the median of at least 5 runs within half of
    budget_s.  value is the faster form."""
    import tempfile

    import oracle as O
    build = "x86-64-v3 prebuilt"
    try:
        O.use_native_build(tempfile.mkdtemp(prefix="oracle_native_"))
        build = "gcc -O3 -march=native -fopenmp on this host"
    except Exception as e:  # noqa: BLE001 -- the prebuilt library still measures
        print("cpu_baseline: native build failed (%s); timing the prebuilt oracle" % e, file=sys.stderr)
    g = O.Grid((N_GRID,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
    a, _ = O.korobov_pick(M_NODES, SEED, 8)
    st = O.korobov_stencil(g, M_NODES, a, S0)
    lists = O.korobov_dest_lists(g, st)
    forms = {"serial scatter": lambda f: O.korobov_collide(g, st, f),
             "destination gather": lambda f: O.korobov_collide_dest(g, st, lists, f)}
    res = {}
    for name, collide in forms.items():
        f = f0.copy()[None]
        collide(f)  # warm-up
        times = []
        t_all = time.perf_counter()
        while len(times) < 5 or (time.perf_counter() - t_all < budget_s / 2 and len(times) < 50):
            t0 = time.perf_counter()
            lo, q = collide(f)
            qc, _, _ = O.correct_loss(g.xi, lo, q, 1)
            f = f + DT * qc
            times.append(time.perf_counter() - t0)
        res[name] = (1.0 / float(np.median(times)), len(times))
    best = max(res, key=lambda k: res[k][0])
    threads = int(os.environ.get("OMP_NUM_THREADS", 0) or len(os.sched_getaffinity(0)))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": res[best][0], "unit": "Q(f,f) evals/s", "cores": threads, "kind": "port", "cpu_model": model,
            "nproc": len(os.sched_getaffinity(0)), "build": build, "form": best,
            "forms": {k: round(v[0], 2) for k, v in res.items()},
            "sample": "corrected evaluations + Euler steps of the same 32^3 / 2^21-node cell, median of %s runs "
                      "per form after one warm-up; stencil and destination lists built untimed"
                      % "/".join(str(v[1]) for v in res.values())}


def progress(msg):
    """One line on stderr per finished leg, so a long default run shows where it is."""
    print("bench.py: %s" % msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300,
                    help="untimed steps first (300, about 8 ms, bring the GPU to its steady state: "
                         "27.0 against 27.6 us per step after 5)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-roofline", action="store_true", help="skip the instrumented kernel-timing pass")
    ap.add_argument("--fused", action="store_true", help="per-entry recomputation path (GDT_FLAG_FUSED)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the steps as a hipGraph (GDT_FLAG_GRAPH) instead of launching them on the stream")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the extension legs (uneven batch, matrix cores, surrogate)")
    ap.add_argument("--ext", default="uneven_batch,batched_matrix_mfma,surrogate,transport",
                    help="comma-separated extension legs to run (default: all four)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the 32^3 many-cell comparison of the plain, tiled and mixed forms")
    ap.add_argument("--no-form-surrogate", action="store_true",
                    help="skip the surrogate (training and inference) in the 32^3 form comparison")
    ap.add_argument("--form-cells", type=int, default=2048, help="cells of the 32^3 form comparison")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (cells shard with no exchange; "
              "multi-GPU runs are not part of it): --gpus %d refused" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import gpu_determ_amd as G

    ctx = G.Context((N_GRID,) * 3, (-VMAX,) * 3, (VMAX,) * 3, VMAX)
    t0 = time.perf_counter()
    ctx.korobov_build(M_NODES, seed=SEED, s0=S0)
    info = ctx.korobov_info()
    build_s = time.perf_counter() - t0
    rng = np.random.default_rng(SEED)
    f0 = mixture(ctx.xi, rng)
    fd = G.DeviceArray.from_host(f0[None])

    # steps are launched straight onto the context stream by default: on this
    # stack a hipGraph replay of the same three kernels per step costs about
    # 1.7 us more per kernel (36.5 against 31.5 us per step); --graph measures
    # the replay
    flags = G.FLAG_CORRECT | (G.FLAG_GRAPH if args.graph else 0) | (G.FLAG_FUSED if args.fused else 0)
    ctx.run_steps(G.METHOD_KOROBOV, fd, 1, DT, args.warmup, flags)
    ctx.sync()
    # the timed steps start again from the initial mixture (after hundreds of
    # warm-up steps f sits at its discrete equilibrium and no longer changes)
    fd.upload(f0[None])  # same buffer: a captured step graph stays valid
    ctx.sync()
    f_start = fd.get()[0]
    t0 = time.perf_counter()
    ctx.run_steps(G.METHOD_KOROBOV, fd, 1, DT, args.steps, flags)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    f_end = fd.get()[0]
    progress("headline steps done (%.1f evals/s)" % (args.steps / elapsed))

    m0 = moments(ctx.xi, f_start, ctx.d3v)
    m1 = moments(ctx.xi, f_end, ctx.d3v)
    cons = float(np.max(np.abs(m1 - m0) / np.maximum(np.abs(m0), np.abs(m0).max() * 1e-3)))
    h0, h1 = entropy(f_start, ctx.d3v), entropy(f_end, ctx.d3v)

    n = ctx.n_pnt
    nn, ne = info["nodes"], info["entries"]
    # algorithmic bytes per launch (cells = 1)
    tl = ctx.korobov_tiles()
    G_ = tl["groups"]              # node groups of the gather (one per XCD)
    nparts = G_ * tl["tiles"]      # gather tiles per cell: one moment partial each
    gi = ctx.korobov_gather_info()
    neg, npc = gi["entries"], gi["slots"]  # the gather's entries; lambda value slots
    alg = {
        # node records (one uint4 each) + f + the chunks' first slots in;
        # (D, 2wP) per node and one value per lambda piece out
        "kor_node": nn * 16 + 8 * n + 4 * ((nn + 15) // 16) + nn * 16 + npc * 16,
        # entry records (node|role, num|den) + node values and piece sums
        # (each read at least once) + group row pointers + centres in; group
        # partials of Q and loss + moment sums out
        "kor_gather": neg * 8 + (nn + npc) * 16 + G_ * 4 * (n + 1) + 24 * n + G_ * 16 * n + nparts * 160,
        # moment sums + centres + f + group partials in; f out
        "kor_apply": nparts * 160 + 24 * n + 8 * n + G_ * 16 * n + 8 * n,
        # alternative fused kernel: 20-byte entry records + f + ptr + centres
        "kor_fused": ne * 20 + 4 * (n + 1) + 8 * n + 24 * n + 16 * n + 256 * 160,
    }
    # FP64 operations per node of k_kor_node, counted from node_terms and
    # kmath.h (an fma counts 2, a division or square root 1; frexp, rint,
    # ldexp and conversions 0): with r != 0 (the log / exp branch) 5 products
    # and quotients before the logarithm, kmt_log 24, r * log 1, kmt_exp 17,
    # Pl * exp 1, w 2, D 2, 2wP 2 = 54; with r = 0, P, Pl, w, D, 2wP = 8.
    # Every node with r != 0 contributes two r-role entries, the others none.
    nz = (ne - 4 * nn) // 2
    node_flop = 54 * nz + 8 * (nn - nz)
    traffic = profiled_traffic()
    roof = None
    kstats = {}
    if not args.no_roofline:
        ctx.kernel_stats_reset()
        reps = max(10, min(50, args.steps))
        ctx.run_steps(G.METHOD_KOROBOV, fd, 1, DT, reps, (flags & ~G.FLAG_GRAPH) | G.FLAG_TIME)
        ctx.sync()
        for k in alg:
            ms, cnt = ctx.kernel_stats(k)
            if cnt:
                kstats[k] = {"avg_us": 1e3 * ms / cnt, "launches": cnt,
                             "alg_bytes": alg[k], "gbs": alg[k] / (ms / cnt * 1e-3) / 1e9}
        if "kor_node" in kstats:
            tf = node_flop / (kstats["kor_node"]["avg_us"] * 1e-6) / 1e12
            kstats["kor_node"].update({"fp64_flop": node_flop, "fp64_tflops": tf,
                                       "fp64_frac": tf / FP64_PEAK_TFLOPS})
        dom = max(kstats, key=lambda k: kstats[k]["avg_us"])
        ach = kstats[dom]["gbs"]
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic.get(dom),
                "traffic_source": "profiles/korobov_step_32cube_2e21.md (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
                if dom in traffic else None,
                "avg_us": round(kstats[dom]["avg_us"], 3), "alg_bytes": alg[dom],
                "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                            for k, v in kstats.items()}}

    if args.calib_copy_mib:
        nb = args.calib_copy_mib << 20
        a_ = G.DeviceArray(nb // 8)
        b_ = G.DeviceArray(nb // 8)
        G.check(G.lib().gdt_memcpy_dtod(b_.ptr, a_.ptr, nb))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx.xi, f0, n, ctx.d3v)
        progress("cpu baseline done")

    value = args.steps / elapsed
    forms = None
    if not args.no_forms:
        forms = forms_32cube(G, ctx, value, cells=args.form_cells, surrogate=not args.no_form_surrogate)
        progress("32^3 form comparison done")

    ext = {}
    legs = {"uneven_batch": uneven_batch, "batched_matrix_mfma": batched_matrix, "surrogate": surrogate_form,
            "transport": transport_step}
    if not args.no_batch:
        for name in [x for x in args.ext.split(",") if x]:
            if name not in legs:
                print("bench.py: unknown extension leg %r (known: %s)" % (name, ", ".join(legs)), file=sys.stderr)
                sys.exit(2)
            ext[name] = legs[name](G)
            progress("extension leg %s done" % name)

    out = {
        "metric": "Q(f,f) evals/s at 32^3 grid: plain FP64 vs mixed/surrogate @ rel-err 1e-4",
        "value": round(value, 3),
        "unit": "Q(f,f) evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 3-Maxwellian mixture)",
        "config": {"workload": "plain FP64 Korobov/projection Q(f,f), full relaxation step", "grid": "32^3",
                   "vmax": VMAX, "nodes": n, "korobov_M": M_NODES, "korobov_a": info["a"],
                   "stencil_nodes": nn, "csr_entries": ne, "cells": 1, "dt": DT, "s0": S0,
                   "correction": True, "graph": bool(args.graph), "stencil_build_s": round(build_s, 3)},
        "checks": {"moment_drift_rel": cons, "entropy_start": h0, "entropy_end": h1,
                   "min_f": float(f_end.min())},
        "roofline": roof,
        "cpu_baseline": cpu,
        "metric_32cube": forms,
        "extensions": ext,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
