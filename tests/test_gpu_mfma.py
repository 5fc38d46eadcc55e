"""Matrix-core paths: the NT GEMM in its three operand modes, and the batched
analytic-matrix operator on a 16^3 box against the FP64 GPU path and the FP64
oracle.  Tolerances are the stated accuracy of each operand mode (bf16 has 8
mantissa bits, the hi + lo split about 16, fp32 24); conservation after the
FP64 correction is exact to FP64 rounding whatever the mode."""
import numpy as np
import pytest

import gpu_determ_amd as G
import oracle as O

from helpers import maxwellian, moments, rel_l2

pytestmark = pytest.mark.gpu

L = 4.8


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("shape", [(256, 384, 512), (256, 128, 4096)])
@pytest.mark.parametrize("mode,tol", [(G.GEMM_BF16, 1e-2), (G.GEMM_BF16_SPLIT, 2e-5), (G.GEMM_F32, 2e-6)])
def test_gemm_nt_modes(mode, tol, shape):
    """Both a square-ish product and a skinny one (2 tiles, K 4096: split K
    over partial tiles added in a fixed order)."""
    say("gemm mode %d shape %s" % (mode, shape))
    rng = np.random.default_rng(3)
    M, N, K = shape
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    C = G.gemm_nt(mode, G.DeviceArray.from_host(A), G.DeviceArray.from_host(B)).get()
    ref = A @ B.T
    assert rel_l2(C, ref) < tol, rel_l2(C, ref)
    # accumulate flag adds to C
    C2 = G.DeviceArray.from_host(C.astype(np.float32), np.float32)
    G.gemm_nt(mode, G.DeviceArray.from_host(A), G.DeviceArray.from_host(B), accumulate=C2)
    assert rel_l2(C2.get(), 2 * ref) < tol


def perturbed(xi, cells, rng):
    """Maxwellians x (1 + 0.05 N(0,1)), clipped positive, renormalised to the
    Maxwellian's density."""
    out = []
    for _ in range(cells):
        n, T = rng.uniform(0.5, 1.0), rng.uniform(0.6, 1.4)
        u = rng.uniform(-0.8, 0.8, 3)
        m = maxwellian(xi, n, T, u)
        f = np.clip(m * (1 + 0.05 * rng.standard_normal(m.shape)), 1e-300, None)
        out.append(f * (m.sum() / f.sum()))
    return np.stack(out)


@pytest.fixture(scope="module")
def box16():
    ctx = G.Context((16, 16, 16), (-L,) * 3, (L,) * 3, -1.0)
    ctx.build_tables(G.TABLES_FP64)
    return ctx


@pytest.mark.parametrize("mode,tol", [(G.GEMM_BF16_SPLIT, 1e-4), (G.GEMM_F32, 2e-5), (G.GEMM_BF16, 2e-2)])
def test_matrix_mfma_vs_fp64(box16, mode, tol):
    say("matrix mfma mode %d vs FP64" % mode)
    ctx = box16
    rng = np.random.default_rng(4)
    cells = 5
    f = perturbed(ctx.xi, cells, rng)
    fd = G.DeviceArray.from_host(f)
    nu_r, gain_r = ctx.collide_matrix(fd, cells)
    _, _, Q_r = ctx.evaluate(G.METHOD_MATRIX, fd, cells, G.FLAG_CORRECT)
    nu_r, gain_r, Q_r = nu_r.get(), gain_r.get(), Q_r.get()
    nu, gain, Q = ctx.matrix_mfma(mode, fd, cells, G.FLAG_CORRECT)
    nu_u, gain_u, Q_u = ctx.matrix_mfma(mode, fd, cells, 0)
    gain, Q, nu_u = gain.get(), Q.get(), nu_u.get()
    assert rel_l2(nu_u, nu_r) < tol, rel_l2(nu_u, nu_r)
    assert rel_l2(gain, gain_r) < tol, rel_l2(gain, gain_r)
    assert rel_l2(Q, Q_r) < 10 * tol, rel_l2(Q, Q_r)
    # conservation of the corrected Q: exact to FP64 rounding
    for c in range(cells):
        m = moments(ctx.xi, Q[c], ctx.d3v)
        scale = np.abs(moments(ctx.xi, np.abs(Q[c]), ctx.d3v)) + 1e-300
        assert np.all(np.abs(m) / scale < 1e-12), m / scale


def test_matrix_mfma_vs_oracle(box16):
    """One cell straight against the FP64 oracle (loss and gain of the box)."""
    say("matrix mfma split vs oracle")
    ctx = box16
    g = O.Grid((16, 16, 16), (-L,) * 3, (L,) * 3, -1.0)
    rng = np.random.default_rng(5)
    f = perturbed(g.xi, 1, rng)
    ob, oa = O.build_tables(g, 0)
    onu = g.to_cmp(O.loss(g, ob, g.to_box(f)))
    og = g.to_cmp(O.gain(g, oa, g.to_box(f)))
    nu, gain, _ = ctx.matrix_mfma(G.GEMM_BF16_SPLIT, G.DeviceArray.from_host(f), 1, 0)
    assert rel_l2(nu.get(), onu) < 1e-4
    assert rel_l2(gain.get(), og) < 1e-4


def test_matrix_mfma_bitwise_reproducible(box16):
    """The gain's read-back is a per-output register sum in a fixed order (no
    atomics), so two runs give the same bits in every operand mode."""
    say("matrix mfma reproducibility")
    ctx = box16
    rng = np.random.default_rng(6)
    fd = G.DeviceArray.from_host(perturbed(ctx.xi, 3, rng))
    for mode in (G.GEMM_BF16, G.GEMM_BF16_SPLIT, G.GEMM_F32):
        a = [x.get() for x in ctx.matrix_mfma(mode, fd, 3, G.FLAG_CORRECT)]
        b = [x.get() for x in ctx.matrix_mfma(mode, fd, 3, G.FLAG_CORRECT)]
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
