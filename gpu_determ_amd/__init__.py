"""gpu_determ_amd: MI355X-native collision-integral library (host binding).

The native code is libgdt.so (C ABI, include/gdt.h) with hand-written gfx950
HIP kernels.  This module binds it with ctypes for tests, the benchmark and
Python callers; it owns no compute path of its own.  Importing it does not
touch the GPU; the library is loaded on first use and a missing build raises
RuntimeError (there is no fallback).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgdt.so")

LAYOUT_CELL_MAJOR = 0
LAYOUT_NODE_MAJOR = 1
BC_PERIODIC = 0
BC_ZERO_GRADIENT = 1
TABLES_FP64 = 0
TABLES_REFERENCE = 1
METHOD_MATRIX = 0
METHOD_KOROBOV = 1
FLAG_CORRECT = 1
FLAG_GRAPH = 2
FLAG_TIME = 4
FLAG_COURANT = 8
FLAG_FUSED = 16
FLAG_CELL_TILES = 32
FLAG_MIXED = 64

GEMM_BF16 = 0
GEMM_BF16_SPLIT = 1
GEMM_F32 = 2

# every symbol that include/gdt.h declares
ABI_SYMBOLS = (
    "gdt_last_error", "gdt_version", "gdt_device_count", "gdt_malloc", "gdt_free", "gdt_memcpy_htod",
    "gdt_memcpy_dtoh", "gdt_memcpy_dtod", "gdt_memset", "gdt_create", "gdt_destroy", "gdt_synchronize",
    "gdt_grid_info", "gdt_grid_nodes", "gdt_basis", "gdt_build_tables", "gdt_tables", "gdt_set_tables",
    "gdt_korobov_build", "gdt_korobov_info", "gdt_korobov_export", "gdt_korobov_export_csr", "gdt_korobov_pick",
    "gdt_collide_matrix", "gdt_collide_korobov", "gdt_correct", "gdt_evaluate", "gdt_relax", "gdt_run_steps",
    "gdt_kernel_stats", "gdt_kernel_stats_reset", "gdt_transpose", "gdt_device_synchronize",
    "gdt_convert_f32_to_f64", "gdt_convert_f64_to_f32",
    "gdt_batch_create", "gdt_batch_destroy", "gdt_batch_set_output", "gdt_batch_config", "gdt_batch_tune",
    "gdt_batch_run", "gdt_batch_synchronize", "gdt_batch_info",
    "gdt_gemm_nt", "gdt_split_f64", "gdt_matrix_mfma",
    "gdt_surrogate_create", "gdt_surrogate_destroy", "gdt_surrogate_train", "gdt_surrogate_predict",
    "gdt_surrogate_infer", "gdt_surrogate_residuals", "gdt_surrogate_export", "gdt_surrogate_info",
    "gdt_init_mixture", "gdt_korobov_tiles", "gdt_korobov_gather_info", "gdt_korobov_set_tiles", "gdt_kmath_pow",
    "gdt_kmath_tables",
    "gdt_transport", "gdt_korobov_cell_tiles", "gdt_evaluate_noalloc", "gdt_surrogate_set_partner",
)

_lib = None


class GdtError(RuntimeError):
    pass


class GridDesc(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("vmin", C.c_double * 3), ("vmax", C.c_double * 3), ("R", C.c_double)]


def lib():
    """Loads libgdt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("gpu_determ_amd/libgdt.so is not built: run `python -m gpu_determ_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.gdt_last_error.restype = C.c_char_p
    L.gdt_korobov_pick.restype = C.c_uint32
    L.gdt_korobov_pick.argtypes = [C.c_uint32, C.c_uint64, C.c_int, dp]
    sigs = {
        "gdt_version": [],
        "gdt_device_count": [ip],
        "gdt_malloc": [C.POINTER(vp), C.c_size_t],
        "gdt_free": [vp],
        "gdt_memcpy_htod": [vp, vp, C.c_size_t],
        "gdt_memcpy_dtoh": [vp, vp, C.c_size_t],
        "gdt_memcpy_dtod": [vp, vp, C.c_size_t],
        "gdt_memset": [vp, C.c_int, C.c_size_t],
        "gdt_create": [C.POINTER(vp), C.POINTER(GridDesc), C.c_int],
        "gdt_destroy": [vp],
        "gdt_synchronize": [vp],
        "gdt_grid_info": [vp, ip, ip, dp, dp],
        "gdt_grid_nodes": [vp, vp, vp],
        "gdt_basis": [vp, vp],
        "gdt_build_tables": [vp, C.c_int],
        "gdt_tables": [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_long), C.POINTER(C.c_long)],
        "gdt_set_tables": [vp, vp, vp],
        "gdt_korobov_build": [vp, C.c_uint32, C.c_uint64, C.c_uint32, C.c_double],
        "gdt_korobov_info": [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_long),
                             C.POINTER(C.c_long), dp],
        "gdt_korobov_export": [vp, vp, vp, vp, vp],
        "gdt_korobov_export_csr": [vp, vp, vp],
        "gdt_collide_matrix": [vp, vp, vp, vp, C.c_int, C.c_int],
        "gdt_collide_korobov": [vp, vp, vp, vp, C.c_int, C.c_int],
        "gdt_correct": [vp, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int],
        "gdt_evaluate": [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int, C.c_uint],
        "gdt_evaluate_noalloc": [vp, vp, vp, vp, vp, C.c_int, C.c_uint],
        "gdt_relax": [vp, vp, vp, vp, C.c_int, C.c_double, C.c_double, vp, C.c_uint],
        "gdt_device_synchronize": [],
        "gdt_convert_f32_to_f64": [vp, vp, C.c_size_t],
        "gdt_convert_f64_to_f32": [vp, vp, C.c_size_t],
        "gdt_run_steps": [vp, C.c_int, vp, C.c_int, C.c_double, C.c_int, C.c_uint],
        "gdt_kernel_stats": [vp, C.c_char_p, dp, C.POINTER(C.c_long)],
        "gdt_kernel_stats_reset": [vp],
        "gdt_transpose": [vp, vp, vp, C.c_int, C.c_int],
        "gdt_batch_create": [C.POINTER(vp), C.c_int, C.POINTER(vp), C.POINTER(vp), ip],
        "gdt_batch_destroy": [vp],
        "gdt_batch_set_output": [vp, C.POINTER(vp)],
        "gdt_batch_config": [vp, C.c_int, C.c_int],
        "gdt_batch_tune": [vp, C.c_uint, ip, ip, dp],
        "gdt_batch_run": [vp, C.c_double, C.c_int, C.c_uint],
        "gdt_batch_synchronize": [vp],
        "gdt_batch_info": [vp, ip, ip, ip],
        "gdt_gemm_nt": [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int],
        "gdt_split_f64": [vp, C.c_long, vp, vp, vp],
        "gdt_matrix_mfma": [vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_uint],
        "gdt_surrogate_create": [C.POINTER(vp), vp, C.c_int, C.c_uint64],
        "gdt_surrogate_destroy": [vp],
        "gdt_surrogate_train": [vp, vp, vp, C.c_long, C.c_int, C.c_int, C.c_double, dp],
        "gdt_surrogate_predict": [vp, vp, vp, C.c_int],
        "gdt_surrogate_infer": [vp, vp, vp, C.c_int, C.c_double, vp, ip],
        "gdt_surrogate_residuals": [vp, vp, C.c_int],
        "gdt_surrogate_set_partner": [vp, vp],
        "gdt_surrogate_export": [vp, vp, vp, vp, vp, vp, vp, vp],
        "gdt_surrogate_info": [vp, ip, ip, ip, C.POINTER(C.c_long)],
        "gdt_init_mixture": [vp, vp, C.c_int, C.c_long, vp],
        "gdt_korobov_tiles": [vp, ip, ip, ip],
        "gdt_korobov_gather_info": [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)],
        "gdt_korobov_set_tiles": [vp, C.c_int],
        "gdt_korobov_cell_tiles": [vp, C.c_int, C.c_int, C.c_int, C.c_int],
        "gdt_transport": [vp, vp, vp, C.c_int, C.c_int, C.c_int, dp, C.c_double, ip],
        "gdt_kmath_pow": [vp, vp, vp, vp, C.c_long],
        "gdt_kmath_tables": [vp],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise GdtError(lib().gdt_last_error().decode())
    return rc


def device_count():
    n = C.c_int()
    check(lib().gdt_device_count(C.byref(n)))
    return n.value


class DeviceArray:
    """FP64 (or other dtype) buffer in device memory, owned by this object."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().gdt_malloc(C.byref(p), max(self.nbytes, 8)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(lib().gdt_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes))

    def get(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().gdt_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        check(lib().gdt_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().gdt_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """A velocity grid on one GPU with its tables, stencil and scratch."""

    def __init__(self, n, vmin, vmax, R=-1.0, device=0):
        g = GridDesc()
        for a in range(3):
            g.n[a] = int(n[a])
            g.vmin[a] = float(vmin[a])
            g.vmax[a] = float(vmax[a])
        g.R = float(R)
        h = C.c_void_p()
        check(lib().gdt_create(C.byref(h), C.byref(g), int(device)))
        self.h = h.value
        n_pnt, n_box = C.c_int(), C.c_int()
        d = (C.c_double * 3)()
        d3v = C.c_double()
        check(lib().gdt_grid_info(self.h, C.byref(n_pnt), C.byref(n_box), d, C.byref(d3v)))
        self.n = tuple(int(x) for x in n)
        self.n_pnt, self.n_box = n_pnt.value, n_box.value
        self.d = tuple(d)
        self.d3v = d3v.value
        self.cidx = np.empty((self.n_pnt, 3), np.int32)
        self.xi = np.empty((self.n_pnt, 3))
        check(lib().gdt_grid_nodes(self.h, self.cidx.ctypes.data, self.xi.ctypes.data))

    def close(self):
        if getattr(self, "h", None):
            lib().gdt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().gdt_synchronize(self.h))

    # -- tables ---------------------------------------------------------
    def build_tables(self, mode=TABLES_FP64):
        check(lib().gdt_build_tables(self.h, int(mode)))
        self.sync()

    def set_tables(self, b, a):
        """Replaces the tables by device FP64 arrays b [n_box] and a
        [n_box^2 / 8] in the reference's index order."""
        check(lib().gdt_set_tables(self.h, b.ptr, a.ptr))
        self.sync()

    def basis(self):
        """The 14 moment-basis planes over the box, host FP64 [14][n_box]."""
        out = DeviceArray((14, self.n_box))
        check(lib().gdt_basis(self.h, out.ptr))
        self.sync()
        return out.get()

    def tables(self):
        """Host copies of (b, a) and the nonzero count."""
        b, a = C.c_void_p(), C.c_void_p()
        n_a, nnz = C.c_long(), C.c_long()
        check(lib().gdt_tables(self.h, C.byref(b), C.byref(a), C.byref(n_a), C.byref(nnz)))
        hb = np.empty(self.n_box)
        ha = np.empty(n_a.value)
        check(lib().gdt_memcpy_dtoh(hb.ctypes.data, b.value, hb.nbytes))
        check(lib().gdt_memcpy_dtoh(ha.ctypes.data, a.value, ha.nbytes))
        return hb, ha, nnz.value

    # -- Korobov ----------------------------------------------------------
    def korobov_build(self, M, seed=0, a=0, s0=1.0):
        check(lib().gdt_korobov_build(self.h, int(M), int(seed), int(a), float(s0)))

    def korobov_info(self):
        M, a = C.c_uint32(), C.c_uint32()
        nodes, ents = C.c_long(), C.c_long()
        K = C.c_double()
        check(lib().gdt_korobov_info(self.h, C.byref(M), C.byref(a), C.byref(nodes), C.byref(ents), C.byref(K)))
        return {"M": M.value, "a": a.value, "nodes": nodes.value, "entries": ents.value, "K": K.value}

    def korobov_tiles(self):
        g, t, cap = C.c_int(), C.c_int(), C.c_int()
        check(lib().gdt_korobov_tiles(self.h, C.byref(g), C.byref(t), C.byref(cap)))
        return {"groups": g.value, "tiles": t.value, "cap": cap.value}

    def korobov_gather_info(self):
        """Entries of the grouped gather's CSRs and the lambda value slots
        the node kernel writes in place of the runs' lambda entries."""
        e, p = C.c_long(), C.c_long()
        check(lib().gdt_korobov_gather_info(self.h, C.byref(e), C.byref(p)))
        return {"entries": e.value, "slots": p.value}

    def korobov_set_tiles(self, tiles_max):
        """Limit of gather tiles per node group (default 64); rebuilds the
        tiles of a built stencil."""
        check(lib().gdt_korobov_set_tiles(self.h, int(tiles_max)))

    def korobov_cell_tiles(self, chunk_tiles=16, node_blocks=4096, gather_blocks=4096, wide=1):
        """Launch shape of the cell-tile paths (chunk of tiles, grid sizes,
        tile width 16 / 32 or 64 cells); the defaults are the library's
        own, which a new context starts with."""
        check(lib().gdt_korobov_cell_tiles(self.h, int(chunk_tiles), int(node_blocks), int(gather_blocks),
                                           int(wide)))

    def kmath_pow(self, x, r):
        """x**r through the node terms' table-driven log / exp (FP64 device
        arrays of equal shape)."""
        assert x.dtype == np.float64 and r.dtype == np.float64 and x.shape == r.shape
        out = DeviceArray(x.shape)
        check(lib().gdt_kmath_pow(self.h, x.ptr, r.ptr, out.ptr, int(np.prod(x.shape))))
        return out

    def korobov_export(self):
        info = self.korobov_info()
        nn = info["nodes"]
        idx6 = np.empty((nn, 6), np.int32)
        r = np.empty(nn)
        w = np.empty(nn)
        node = np.empty(nn, np.uint32)
        check(lib().gdt_korobov_export(self.h, idx6.ctypes.data, r.ctypes.data, w.ctypes.data, node.ctypes.data))
        return idx6, r, w, node

    def korobov_export_csr(self):
        info = self.korobov_info()
        ptr = np.empty(self.n_pnt + 1, np.int32)
        ent = np.empty(info["entries"], np.uint32)
        check(lib().gdt_korobov_export_csr(self.h, ptr.ctypes.data, ent.ctypes.data))
        return ptr, ent

    # -- operator ---------------------------------------------------------
    def collide_matrix(self, f, cells, layout=LAYOUT_CELL_MAJOR):
        nu = DeviceArray((cells, self.n_pnt))
        gain = DeviceArray((cells, self.n_pnt))
        check(lib().gdt_collide_matrix(self.h, f.ptr, nu.ptr, gain.ptr, int(cells), int(layout)))
        return nu, gain

    def collide_korobov(self, f, cells, layout=LAYOUT_CELL_MAJOR):
        loss = DeviceArray((cells, self.n_pnt))
        Q = DeviceArray((cells, self.n_pnt))
        check(lib().gdt_collide_korobov(self.h, f.ptr, loss.ptr, Q.ptr, int(cells), int(layout)))
        return loss, Q

    def evaluate(self, method, f, cells, flags=FLAG_CORRECT, layout=LAYOUT_CELL_MAJOR):
        nu = DeviceArray((cells, self.n_pnt))
        gain = DeviceArray((cells, self.n_pnt))
        Q = DeviceArray((cells, self.n_pnt))
        check(lib().gdt_evaluate(self.h, int(method), f.ptr, nu.ptr, gain.ptr, Q.ptr, int(cells), int(layout),
                                 int(flags)))
        return nu, gain, Q

    def evaluate_noalloc(self, f, cells, flags=FLAG_CORRECT, nu=None, gain=None, Q=None):
        """gdt_evaluate_noalloc (Korobov, cell-major): outputs into the given
        DeviceArrays (new ones when None); the library allocates nothing."""
        nu = DeviceArray((cells, self.n_pnt)) if nu is None else nu
        gain = DeviceArray((cells, self.n_pnt)) if gain is None else gain
        Q = DeviceArray((cells, self.n_pnt)) if Q is None else Q
        check(lib().gdt_evaluate_noalloc(self.h, f.ptr, nu.ptr, gain.ptr, Q.ptr, int(cells), int(flags)))
        return nu, gain, Q

    def correct(self, method, f, nu, gain, Q, cells, X=None):
        check(lib().gdt_correct(self.h, int(method), f.ptr, nu.ptr, gain.ptr, Q.ptr, X.ptr if X else None,
                                int(cells), LAYOUT_CELL_MAJOR))

    def relax(self, f, nu, gain, cells, dt=0.0, courant=0.5, time=None, flags=0):
        """relax_f: f <- f (1 - nu dt) + gain dt (dt from courant with FLAG_COURANT)."""
        check(lib().gdt_relax(self.h, f.ptr, nu.ptr, gain.ptr, int(cells), float(dt), float(courant),
                              time.ptr if time else None, int(flags)))

    def run_steps(self, method, f, cells, dt, steps, flags=FLAG_CORRECT):
        check(lib().gdt_run_steps(self.h, int(method), f.ptr, int(cells), float(dt), int(steps), int(flags)))

    def transport(self, f, shape, spacing, dt, bc=(BC_PERIODIC,) * 3, out=None):
        """One first-order upwind transport step of f [cells][n_pnt] over a box
        of shape (nx, ny, nz) cells (cell (ix ny + iy) nz + iz); returns f_out
        (a new DeviceArray unless `out` is given; it must not be f)."""
        nx, ny, nz = (int(x) for x in shape)
        if out is None:
            out = DeviceArray((nx * ny * nz, self.n_pnt))
        sp = (C.c_double * 3)(*[float(x) for x in spacing])
        b = (C.c_int * 3)(*[int(x) for x in bc])
        check(lib().gdt_transport(self.h, f.ptr, out.ptr, nx, ny, nz, sp, float(dt), b))
        return out

    def init_mixture(self, params):
        """f [cells][n_pnt] (DeviceArray) from params [cells][comps][5] = (n, T, ux, uy, uz)."""
        params = np.ascontiguousarray(params, dtype=np.float64)
        cells, comps = params.shape[0], params.shape[1]
        pd = DeviceArray.from_host(params)
        f = DeviceArray((cells, self.n_pnt))
        check(lib().gdt_init_mixture(self.h, pd.ptr, comps, cells, f.ptr))
        self.sync()
        return f

    def matrix_mfma(self, mode, f, cells, flags=FLAG_CORRECT):
        """Batched analytic-matrix operator on the matrix cores (16^3 box):
        returns (nu corrected, gain, Q) DeviceArrays [cells][4096]."""
        nu = DeviceArray((cells, self.n_pnt))
        gain = DeviceArray((cells, self.n_pnt))
        Q = DeviceArray((cells, self.n_pnt))
        check(lib().gdt_matrix_mfma(self.h, int(mode), f.ptr, nu.ptr, gain.ptr, Q.ptr, int(cells), int(flags)))
        return nu, gain, Q

    def kernel_stats(self, name):
        ms, cnt = C.c_double(), C.c_long()
        check(lib().gdt_kernel_stats(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def kernel_stats_reset(self):
        check(lib().gdt_kernel_stats_reset(self.h))

    def transpose(self, src, dst, rows, cols):
        check(lib().gdt_transpose(self.h, src.ptr, dst.ptr, int(rows), int(cols)))


class Batch:
    """Uneven batch of cells over several Korobov contexts (gdt_batch_*).

    classes: Contexts with a stencil; fs: DeviceArrays [cells][n_pnt] per
    class (kept referenced: the batch holds their device pointers)."""

    def __init__(self, classes, fs, Qs=None):
        self.classes, self.fs, self.Qs = list(classes), list(fs), Qs
        n = len(self.classes)
        hs = (C.c_void_p * n)(*[c.h for c in self.classes])
        fp = (C.c_void_p * n)(*[f.ptr for f in self.fs])
        cells = (C.c_int * n)(*[f.shape[0] for f in self.fs])
        h = C.c_void_p()
        check(lib().gdt_batch_create(C.byref(h), n, hs, fp, cells))
        self.h = h
        if Qs is not None:
            qp = (C.c_void_p * n)(*[q.ptr for q in Qs])
            check(lib().gdt_batch_set_output(self.h, qp))

    def config(self, seg_cells, node_q):
        check(lib().gdt_batch_config(self.h, int(seg_cells), int(node_q)))

    def tune(self, flags=FLAG_CORRECT):
        """Autotune on the batch's cells: (cells per segment, nodes per thread,
        best evaluation ms) for the work lists; with FLAG_CELL_TILES or
        FLAG_MIXED the classes' cell-tile shape, returned as (cells per chunk,
        gather grid, best evaluation ms)."""
        k, q, ms = C.c_int(), C.c_int(), C.c_double()
        check(lib().gdt_batch_tune(self.h, flags, C.byref(k), C.byref(q), C.byref(ms)))
        return k.value, q.value, ms.value

    def run(self, dt, steps, flags=FLAG_CORRECT):
        check(lib().gdt_batch_run(self.h, float(dt), int(steps), flags))

    def sync(self):
        check(lib().gdt_batch_synchronize(self.h))

    def info(self):
        k, q, ns = C.c_int(), C.c_int(), C.c_int()
        check(lib().gdt_batch_info(self.h, C.byref(k), C.byref(q), C.byref(ns)))
        return {"seg_cells": k.value, "node_q": q.value, "segments": ns.value}

    def close(self):
        if getattr(self, "h", None):
            lib().gdt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """Learned Q(f,f) surrogate on the grid of a Context (gdt_surrogate_*)."""

    def __init__(self, ctx, hidden=128, seed=1):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().gdt_surrogate_create(C.byref(h), ctx.h, int(hidden), int(seed)))
        self.h = h
        n, np_, H, P = C.c_int(), C.c_int(), C.c_int(), C.c_long()
        check(lib().gdt_surrogate_info(self.h, C.byref(n), C.byref(np_), C.byref(H), C.byref(P)))
        self.n, self.np, self.hidden, self.params = n.value, np_.value, H.value, P.value

    def train(self, f, Q, epochs=10, batch=256, lr=1e-3):
        loss = C.c_double()
        check(lib().gdt_surrogate_train(self.h, f.ptr, Q.ptr, int(f.shape[0]), int(epochs), int(batch), float(lr),
                                        C.byref(loss)))
        return loss.value

    def set_partner(self, other):
        """Second network of the residual check (None clears)."""
        check(lib().gdt_surrogate_set_partner(self.h, other.h if other is not None else None))
        self.partner = other

    def predict(self, f):
        Q = DeviceArray(f.shape)
        check(lib().gdt_surrogate_predict(self.h, f.ptr, Q.ptr, int(f.shape[0])))
        return Q

    def infer(self, f, threshold=1e-4):
        cells = int(f.shape[0])
        Q = DeviceArray(f.shape)
        fb = DeviceArray((cells,), np.int32)
        nfb = C.c_int()
        check(lib().gdt_surrogate_infer(self.h, f.ptr, Q.ptr, cells, float(threshold), fb.ptr, C.byref(nfb)))
        res = np.empty(cells)
        check(lib().gdt_surrogate_residuals(self.h, res.ctypes.data, cells))
        return Q, fb.get().astype(bool), res

    def export(self):
        H, n_ = self.hidden, self.np
        W1, b1 = np.empty((H, n_), np.float32), np.empty(H, np.float32)
        W2, b2 = np.empty((H, n_), np.float32), np.empty(H, np.float32)
        W3, b3 = np.empty((n_, H), np.float32), np.empty(n_, np.float32)
        norm = np.empty((4, n_), np.float32)
        check(lib().gdt_surrogate_export(self.h, W1.ctypes.data, b1.ctypes.data, W2.ctypes.data, b2.ctypes.data,
                                         W3.ctypes.data, b3.ctypes.data, norm.ctypes.data))
        return W1, b1, W2, b2, W3, b3, norm

    def close(self):
        if getattr(self, "h", None):
            lib().gdt_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def split_f64(x, hi=True, lo=True, f32=True):
    """FP64 DeviceArray -> (bf16 hi, bf16 lo, fp32) DeviceArrays (None where not asked)."""
    n = int(np.prod(x.shape))
    h = DeviceArray(x.shape, np.uint16) if hi else None
    l_ = DeviceArray(x.shape, np.uint16) if lo else None
    f = DeviceArray(x.shape, np.float32) if f32 else None
    check(lib().gdt_split_f64(x.ptr, n, h.ptr if h else None, l_.ptr if l_ else None, f.ptr if f else None))
    return h, l_, f


def gemm_nt(mode, A, B, accumulate=None):
    """C = A B^T on the matrix cores (FP32 result).  A: [M][K], B: [N][K] FP64
    DeviceArrays, converted per mode; returns the FP32 DeviceArray C [M][N]."""
    M, K = A.shape
    N = B.shape[0]
    ah, al, af = split_f64(A)
    bh, bl, bf = split_f64(B)
    C_ = accumulate if accumulate is not None else DeviceArray((M, N), np.float32)
    if mode == GEMM_F32:
        check(lib().gdt_gemm_nt(mode, M, N, K, af.ptr, None, bf.ptr, None, C_.ptr, N, int(accumulate is not None)))
    else:
        check(lib().gdt_gemm_nt(mode, M, N, K, ah.ptr, al.ptr, bh.ptr, bl.ptr, C_.ptr, N, int(accumulate is not None)))
    return C_


def kmath_tables():
    """The node terms' log / exp tables (host, no device needed)."""
    T = np.zeros(520)
    check(lib().gdt_kmath_tables(T.ctypes.data))
    return T


def korobov_pick(M, seed, ncand=8):
    p2 = C.c_double()
    a = lib().gdt_korobov_pick(int(M), int(seed), int(ncand), C.byref(p2))
    return int(a), p2.value
