"""gdt_evaluate_noalloc (the allocation-free Korobov evaluation, through the
C ABI) against gdt_evaluate: the same kernels, so the same bits, on two calls
in a row (the second with fewer cells), corrected and uncorrected; on a fresh
context whose scratch buffers were never allocated; and the size limit of its
static scratch is an error, after which calls still work."""
import numpy as np
import pytest

import gpu_determ_amd as G
from helpers import mixture

pytestmark = pytest.mark.gpu
VM = 4.8


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,M,calls", [(16, 1 << 16, (8, 3)), (32, 1 << 21, (5, 2))])
def test_noalloc_equals_evaluate(n, M, calls):
    say("noalloc", n, M, calls)
    c = G.Context((n,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    c.korobov_build(M, seed=13, s0=0.25)
    rng = np.random.default_rng(n)
    for cells in calls:
        f = G.DeviceArray.from_host(np.stack([mixture(c.xi, rng) for _ in range(cells)]))
        for flags in (G.FLAG_CORRECT, 0):
            a = [x.get() for x in c.evaluate_noalloc(f, cells, flags)]  # first: before any evaluate
            b = [x.get() for x in c.evaluate(G.METHOD_KOROBOV, f, cells, flags)]
            for u, v in zip(a, b):
                assert np.array_equal(u, v)


def test_noalloc_limits():
    say("noalloc limits")
    c = G.Context((32,) * 3, (-VM,) * 3, (VM,) * 3, VM)
    c.korobov_build(1 << 21, seed=13, s0=0.25)
    rng = np.random.default_rng(3)
    f = G.DeviceArray.from_host(np.stack([mixture(c.xi, rng) for _ in range(16)]))
    with pytest.raises(RuntimeError, match="static scratch"):
        c.evaluate_noalloc(f, 16)
    with pytest.raises(RuntimeError):
        c.evaluate_noalloc(f, 2, G.FLAG_CORRECT | G.FLAG_GRAPH)
    _, _, q = c.evaluate_noalloc(f, 2)
    _, _, r = c.evaluate(G.METHOD_KOROBOV, f, 2, G.FLAG_CORRECT)
    assert np.array_equal(q.get(), r.get())
    # Q alone (nu and gain not requested)
    Q = G.DeviceArray((2, c.n_pnt))
    G.check(G.lib().gdt_evaluate_noalloc(c.h, f.ptr, None, None, Q.ptr, 2, G.FLAG_CORRECT))
    c.sync()
    assert np.array_equal(Q.get(), r.get())
    with pytest.raises(RuntimeError):  # gain without nu
        G.check(G.lib().gdt_evaluate_noalloc(c.h, f.ptr, None, Q.ptr, Q.ptr, 2, G.FLAG_CORRECT))
