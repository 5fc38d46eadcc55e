"""ctypes/numpy wrapper of the FP64 oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package gpu_determ_amd.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _ptr_or_null(t):
    """ndpointer that also accepts None (NULL)."""

    class _P(t):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return t.from_param(obj)

    return _P


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
    _LIB = _bind(C.CDLL(path))
    return _LIB


def use_native_build(outdir):
    """Compile the oracle for this host's CPU (gcc -march=native, the Makefile's
    `native` target) into outdir and make it the library every function here
    calls.  For the CPU baseline, which is timed on the machine that runs it.
    Returns the library path."""
    global _LIB
    import subprocess
    out = os.path.join(os.path.abspath(outdir), "liboracle_native.so")
    subprocess.run(["make", "-s", "-C", _HERE, "native", "OUT=" + out], check=True)
    _LIB = _bind(C.CDLL(out))
    return out


def _bind(L):
    i3 = C.POINTER(C.c_int)
    d3 = C.POINTER(C.c_double)
    L.orc_grid_nodes.argtypes = [i3, d3, d3, C.c_double, _i32p, _ptr_or_null(_i32p)]
    L.orc_grid_nodes.restype = C.c_int
    L.orc_grid_centres.argtypes = [i3, d3, d3, C.c_int, _f64p, _f64p, _f64p]
    L.orc_init_f_reference.argtypes = [i3, d3, d3, _f64p]
    L.orc_fterm.argtypes = [C.c_int, C.c_double, C.c_double]
    L.orc_fterm.restype = C.c_double
    L.orc_box_area.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [C.c_double]
    L.orc_box_area.restype = C.c_double
    L.orc_build_tables.argtypes = [i3, d3, d3, C.c_int, _f64p, _f64p]
    L.orc_glist_build.argtypes = [i3, _f64p, _ptr_or_null(_i32p), _ptr_or_null(_i32p),
                                  _ptr_or_null(_f64p), C.c_long]
    L.orc_glist_build.restype = C.c_long
    L.orc_loss.argtypes = [i3, _f64p, _f64p, _f64p, C.c_int]
    L.orc_gain.argtypes = [i3, _f64p, _f64p, _f64p, C.c_int, _ptr_or_null(_i32p), C.c_int]
    L.orc_correct.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, _ptr_or_null(_f64p),
                              _ptr_or_null(_f64p), _ptr_or_null(_f64p), C.c_int]
    L.orc_correct.restype = C.c_int
    L.orc_correct_loss.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, _ptr_or_null(_f64p), C.c_int]
    L.orc_correct_loss.restype = C.c_int
    L.orc_solve5.argtypes = [_f64p, _f64p, _f64p]
    L.orc_solve5.restype = C.c_int
    L.orc_courant_dt.argtypes = [_f64p, C.c_long, C.c_double]
    L.orc_courant_dt.restype = C.c_double
    L.orc_euler.argtypes = [_f64p, _f64p, C.c_long, C.c_double]
    L.orc_korobov_pick.argtypes = [C.c_uint32, C.c_uint64, C.c_int, C.POINTER(C.c_double)]
    L.orc_korobov_pick.restype = C.c_uint32
    L.orc_korobov_p2.argtypes = [C.c_uint32, C.c_uint32, C.c_int]
    L.orc_korobov_p2.restype = C.c_double
    L.orc_sincos_2pi.argtypes = [C.c_double, d3, d3]
    L.orc_korobov_stencil.argtypes = [i3, _i32p, C.c_int, C.c_uint32, C.c_uint32, C.c_double,
                                      _ptr_or_null(_i32p), _ptr_or_null(_f64p), _ptr_or_null(_f64p),
                                      _ptr_or_null(_u32p), C.c_long]
    L.orc_korobov_stencil.restype = C.c_long
    L.orc_korobov_collide.argtypes = [C.c_int, C.c_long, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int]
    L.orc_korobov_dest_lists.argtypes = [C.c_int, C.c_long, _i32p, _f64p, _i64p, _i32p, _f64p, _i64p, _i32p]
    L.orc_korobov_collide_dest.argtypes = [C.c_int, C.c_long, _i32p, _f64p, _f64p, _i64p, _i32p, _f64p, _i64p,
                                           _i32p, _f64p, _f64p, _f64p, C.c_int]
    return L


def _i3(n):
    return (C.c_int * 3)(*[int(x) for x in n])


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


class Grid:
    """Host description of a velocity grid (velocity_grid.cu:73-126)."""

    def __init__(self, n, vmin, vmax, R):
        self.n = tuple(int(x) for x in n)
        self.vmin = tuple(float(x) for x in vmin)
        self.vmax = tuple(float(x) for x in vmax)
        self.R = float(R)
        nb = self.n[0] * self.n[1] * self.n[2]
        self.n_box = nb
        self.box2c = np.empty(nb, np.int32)
        cidx = np.empty(3 * nb, np.int32)
        self.n_pnt = lib().orc_grid_nodes(_i3(self.n), _d3(self.vmin), _d3(self.vmax), self.R, self.box2c, cidx)
        self.cidx = cidx[: 3 * self.n_pnt].reshape(-1, 3).copy()
        self.d = tuple((self.vmax[a] - self.vmin[a]) / self.n[a] for a in range(3))
        self.d3v = self.d[0] * self.d[1] * self.d[2]
        u, v, w = self.centres(False)
        self.xi = np.stack([u[self.cidx[:, 0]], v[self.cidx[:, 1]], w[self.cidx[:, 2]]], axis=1).copy()

    def centres(self, ref_float=False):
        u = np.empty(self.n[0])
        v = np.empty(self.n[1])
        w = np.empty(self.n[2])
        lib().orc_grid_centres(_i3(self.n), _d3(self.vmin), _d3(self.vmax), int(ref_float), u, v, w)
        return u, v, w

    def box_index(self):
        c = self.cidx
        return ((c[:, 0] * self.n[1] + c[:, 1]) * self.n[2] + c[:, 2]).astype(np.int64)

    def to_box(self, f_cmp):
        f_cmp = np.asarray(f_cmp, dtype=np.float64)
        cells = f_cmp.size // self.n_pnt
        out = np.zeros((cells, self.n_box))
        out[:, self.box_index()] = f_cmp.reshape(cells, self.n_pnt)
        return out

    def to_cmp(self, f_box):
        f_box = np.asarray(f_box, dtype=np.float64)
        cells = f_box.size // self.n_box
        return np.ascontiguousarray(f_box.reshape(cells, self.n_box)[:, self.box_index()])


def init_f_reference(g):
    out = np.empty(g.n_box)
    lib().orc_init_f_reference(_i3(g.n), _d3(g.vmin), _d3(g.vmax), out)
    return out


def build_tables(g, mode=0):
    nb = g.n_box
    b = np.empty(nb)
    a = np.empty(nb * nb // 8)
    lib().orc_build_tables(_i3(g.n), _d3(g.vmin), _d3(g.vmax), int(mode), b, a)
    return b, a


def glist(g, a):
    L = lib()
    ne = L.orc_glist_build(_i3(g.n), a, None, None, None, 0)
    gptr = np.empty(g.n_box + 1, np.int32)
    gbox = np.empty(max(ne, 1), np.int32)
    gval = np.empty(max(ne, 1))
    L.orc_glist_build(_i3(g.n), a, gptr, gbox, gval, ne)
    return gptr, gbox[:ne], gval[:ne]


def loss(g, b, f_box):
    f_box = np.ascontiguousarray(f_box, dtype=np.float64)
    cells = f_box.size // g.n_box
    out = np.empty(cells * g.n_box)
    lib().orc_loss(_i3(g.n), b, f_box.ravel(), out, cells)
    return out.reshape(cells, g.n_box)


def gain(g, a, f_box, dest=None):
    f_box = np.ascontiguousarray(f_box, dtype=np.float64)
    cells = f_box.size // g.n_box
    if dest is None:
        out = np.empty(cells * g.n_box)
        lib().orc_gain(_i3(g.n), a, f_box.ravel(), out, cells, None, 0)
        return out.reshape(cells, g.n_box)
    dest = np.ascontiguousarray(dest, dtype=np.int32)
    out = np.empty(cells * dest.size)
    lib().orc_gain(_i3(g.n), a, f_box.ravel(), out, cells, dest, dest.size)
    return out.reshape(cells, dest.size)


def correct(xi, f, nu, gain_, cells):
    npts = xi.shape[0]
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    f, nu, gain_ = (np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (f, nu, gain_))
    nuc = np.empty(cells * npts)
    Q = np.empty(cells * npts)
    X = np.empty(cells * 5)
    bad = lib().orc_correct(npts, xi.ravel(), f, nu, gain_, nuc, Q, X, cells)
    return nuc.reshape(cells, npts), Q.reshape(cells, npts), X.reshape(cells, 5), bad


def correct_loss(xi, loss_, Q, cells):
    npts = xi.shape[0]
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    loss_, Q = (np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (loss_, Q))
    out = np.empty(cells * npts)
    X = np.empty(cells * 5)
    bad = lib().orc_correct_loss(npts, xi.ravel(), loss_, Q, out, X, cells)
    return out.reshape(cells, npts), X.reshape(cells, 5), bad


def solve5(A, B):
    X = np.empty(5)
    rc = lib().orc_solve5(np.ascontiguousarray(A, dtype=np.float64).ravel(),
                          np.ascontiguousarray(B, dtype=np.float64), X)
    return rc, X


def korobov_pick(M, seed, ncand=8):
    p2 = C.c_double()
    a = lib().orc_korobov_pick(int(M), int(seed), int(ncand), C.byref(p2))
    return int(a), p2.value


def korobov_p2(M, a, dims=8):
    return lib().orc_korobov_p2(int(M), int(a), int(dims))


def sincos_2pi(t):
    s = C.c_double()
    c = C.c_double()
    lib().orc_sincos_2pi(float(t), C.byref(s), C.byref(c))
    return s.value, c.value


def korobov_K(g, M, s0):
    """Node-weight scale pi s0 n_box^2 d^4 / M (equal spacing on all axes)."""
    d = g.d[0]
    return np.pi * s0 * float(g.n_box) ** 2 * d ** 4 / float(M)


def korobov_stencil(g, M, a, s0):
    L = lib()
    K = korobov_K(g, M, s0)
    cnt = L.orc_korobov_stencil(_i3(g.n), g.box2c, g.n_pnt, int(M), int(a), K, None, None, None, None, 0)
    idx6 = np.empty(6 * max(cnt, 1), np.int32)
    r = np.empty(max(cnt, 1))
    w = np.empty(max(cnt, 1))
    node = np.empty(max(cnt, 1), np.uint32)
    L.orc_korobov_stencil(_i3(g.n), g.box2c, g.n_pnt, int(M), int(a), K, idx6, r, w, node, cnt)
    return idx6[: 6 * cnt].reshape(cnt, 6), r[:cnt], w[:cnt], node[:cnt]


def korobov_collide(g, st, f_cmp):
    idx6, r, w, _ = st
    f_cmp = np.ascontiguousarray(f_cmp, dtype=np.float64)
    cells = f_cmp.size // g.n_pnt
    lo = np.empty(cells * g.n_pnt)
    Q = np.empty(cells * g.n_pnt)
    lib().orc_korobov_collide(g.n_pnt, idx6.shape[0], np.ascontiguousarray(idx6).ravel(), r, w,
                              f_cmp.ravel(), lo, Q, cells)
    return lo.reshape(cells, g.n_pnt), Q.reshape(cells, g.n_pnt)


def korobov_dest_lists(g, st):
    """Per-destination write lists of a stencil, in the serial scatter's
    order: (qrow, qnode, qcoef) for Q and (lrow, lnode) for the loss; input
    of korobov_collide_dest."""
    idx6 = np.ascontiguousarray(st[0], dtype=np.int32).ravel()
    r = np.ascontiguousarray(st[1], dtype=np.float64)
    nn = idx6.size // 6
    qrow = np.empty(g.n_pnt + 1, np.int64)
    lrow = np.empty(g.n_pnt + 1, np.int64)
    qnode = np.empty(max(6 * nn, 1), np.int32)
    qcoef = np.empty(max(6 * nn, 1), np.float64)
    lnode = np.empty(max(2 * nn, 1), np.int32)
    lib().orc_korobov_dest_lists(g.n_pnt, nn, idx6, r, qrow, qnode, qcoef, lrow, lnode)
    return qrow, qnode, qcoef, lrow, lnode


def korobov_collide_dest(g, st, lists, f_cmp):
    """korobov_collide with OpenMP over destinations (same bits)."""
    idx6, r, w, _ = st
    f_cmp = np.ascontiguousarray(f_cmp, dtype=np.float64)
    cells = f_cmp.size // g.n_pnt
    lo = np.empty(cells * g.n_pnt)
    Q = np.empty(cells * g.n_pnt)
    lib().orc_korobov_collide_dest(g.n_pnt, idx6.shape[0], np.ascontiguousarray(idx6).ravel(), r, w, *lists,
                                   f_cmp.ravel(), lo, Q, cells)
    return lo.reshape(cells, g.n_pnt), Q.reshape(cells, g.n_pnt)


def courant_dt(nu, courant):
    nu = np.ascontiguousarray(nu, dtype=np.float64).ravel()
    return lib().orc_courant_dt(nu, nu.size, float(courant))


def transport_upwind(f, xi, shape, spacing, dt, bc=(0, 0, 0)):
    """One first-order upwind step of df/dt + xi . grad_x f = 0 (no reference
    counterpart; the spatial step beside the collision operator).

    f [cells][n] over a box of shape (nx, ny, nz) cells, cell (ix ny + iy) nz +
    iz; xi [n][3].  Per axis a in the order x, y, z: k = (dt / h_a) xi_a,
    D = f - f[behind] for k > 0, f[ahead] - f otherwise, f <- f - k D (each a
    separate rounding, as the kernel computes it).  bc[a] = 0 periodic, 1
    zero-gradient (a missing neighbour is the cell itself)."""
    nx, ny, nz = shape
    n = xi.shape[0]
    f4 = np.asarray(f, dtype=np.float64).reshape(nx, ny, nz, n)
    acc = f4.copy()
    for a in range(3):
        k = (dt / spacing[a]) * xi[:, a]
        if bc[a] == 0:
            back = np.roll(f4, 1, axis=a)
            ahead = np.roll(f4, -1, axis=a)
        else:
            idx = np.arange(f4.shape[a])
            back = np.take(f4, np.maximum(idx - 1, 0), axis=a)
            ahead = np.take(f4, np.minimum(idx + 1, f4.shape[a] - 1), axis=a)
        d = np.where(k > 0.0, f4 - back, ahead - f4)
        acc = acc - k * d
    return acc.reshape(nx * ny * nz, n)
