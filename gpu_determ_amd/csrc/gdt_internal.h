// Internal state and kernel launchers of libgdt.so (not installed).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/gdt.h"

void gdt_set_error(int code, const char *fmt, ...);

#define GDT_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            gdt_set_error(GDT_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return GDT_E_HIP;                                                                           \
        }                                                                                               \
    } while (0)

#define GDT_CHECK(cond, code, ...)                \
    do {                                          \
        if (!(cond)) {                            \
            gdt_set_error((code), __VA_ARGS__);   \
            return (code);                        \
        }                                         \
    } while (0)

// Korobov gather layout
#define KOR_NGROUPS 8     // node groups of the gather: one per XCD
#define KOR_TILE_DST 512  // destinations per gather tile: one per thread
#define KOR_TILE_CAP 4864 // entries per gather pass (76 KB of LDS at most)
#define KOR_TILES_MAX 64  // default gather tiles per group: G * 64 blocks, two per CU, one round
#define KOR_TILES_LIMIT 1024
#define KOR_CHUNK 16       // node positions per lambda-piece chunk: one DPP row of the node kernel
#define KOR_NPTS_MAX 19200 // grid nodes the packed node records and the LDS copy of f admit
// Cells (grouped Korobov and analytic-matrix paths) or 64-cell tiles (cell
// tiles) one launch takes: they index grid y, which admits at most 65,535.
#define GDT_GRID_Y_MAX 65535
#define GDT_CHECK_GRID_Y(n, what)                                                                            \
    GDT_CHECK((n) >= 1 && (n) <= GDT_GRID_Y_MAX, GDT_E_ARG,                                                  \
              "%s: %ld per call is outside the launch limit of 1 to 65,535 (grid y); split the cells into " \
              "several calls",                                                                                \
              (what), (long)(n))
#define KOR_APPLY_THREADS 128 // destinations per apply block (128: 26.0 us per step; 64: 27.0, 256: 26.3, 512: 27.9)

// Device buffer: owns one device allocation: freed by release() or on destruction (so an
// early return or a destroyed context frees every buffer); movable, not
// copyable.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int ensure(size_t need);  // grow (contents not kept)
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};
// Process-wide count of DevBuf frees (growth or release).  A captured graph
// holds raw pointers to context buffers, so it is replayed only while this
// count is the one read at its capture.
unsigned long devbuf_generation();

struct KernelStat {
    double ms = 0.0;
    long count = 0;
};

struct gdt_context {
    int device = 0;
    hipStream_t stream = nullptr;

    // grid
    int n[3] = {0, 0, 0};
    double vmin[3], vmax[3], R = 0.0, d[3], d3v = 0.0;
    int n_box = 0, n_pnt = 0;
    std::vector<int> h_box2c, h_cidx;
    std::vector<double> h_xi;
    DevBuf box2c, cidx, xi, cmp2box;

    // analytic matrix tables
    int table_mode = -1;
    bool have_tables = false;
    long n_a = 0, n_gent = 0;
    DevBuf tb, ta, gptr, gbox, gval;

    // Korobov stencil
    bool have_kor = false;
    uint32_t kM = 0, ka = 0;
    double ks0 = 0.0, kK = 0.0;
    long nn = 0, nent = 0;
    long nv = 0;  // node values per cell of the grouped path: nn nodes, then the lambda value slots
    DevBuf kidx, kr, kw, knode, kptr, kent;
    // per-node records (one uint4: six 15-bit node indices, num / 4, den / 4
    // and D2) and, for the fused path, per-entry records (uint4 as the
    // node's, uint32 D2 | role << 12 | destination << 14)
    bool have_stream = false;
    DevBuf kn4, ke4, ke1, kwb;
    DevBuf kgptr, kgent;          // per node-group CSRs of the gather (uint2 entry records)
    DevBuf kcbase;                // first lambda value slot of each chunk of KOR_CHUNK node positions
    DevBuf ktile;                 // gather tiles (d0, d1, e0, e1), [group][kntile]
    DevBuf kmath;                 // log / exp tables of the node terms (KMT_SIZE doubles)
    int kntile = 0, kcap = 0;     // tiles per group, LDS entries per gather pass
    int ktiles_max = KOR_TILES_MAX;  // limit of tiles per group (gdt_korobov_set_tiles)
    std::vector<int32_t> h_gp;    // host copy of the group CSR row pointers [G][n+1]
    unsigned kor_gen = 0;         // bumped whenever the stencil's gather layout changes
    std::vector<int32_t> h_kptr;  // host copy of the CSR row pointers
    // many-cell (cell-tile) path, over its own node order q (sorted by the
    // alpha destination): packed node records, the alpha run of each
    // destination (positions aptr[v] .. aptr[v + 1]) and the destinations by
    // run length, longest first; one destination-sorted CSR of the other
    // entries (uint4 records: q << 2 | role, a grid node, an FP64 value), and
    // the destinations ordered by its entry count, longest first
    DevBuf kt4, kaptr, kaperm;
    DevBuf kcptr, kcent, kdperm;
    DevBuf kcent2, kcmid;         // the tile CSR's entries, beta first per destination; where the others begin
    // the gather's destinations per XCD: the grid's octant x (of the box's
    // index halves) in Morton order, for the blocks on XCD x; xoff[x] ..
    // xoff[x + 1] index kxperm
    DevBuf kxperm;
    int kxoff[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    DevBuf krw, krwf;             // (r, w) per tile node position, FP64 and FP32
    long ntent = 0;               // entries of the tile CSR
    int wb_blocks = 0;            // blocks per cell the wave bounds kwb are built for
    int lds_attr = 0;             // dynamic-LDS bytes the fused kernel is enabled for
    int gain_lds_attr = 0;        // dynamic-LDS bytes the matrix gain kernel is enabled for

    // batched matrix form on the matrix cores (16^3 box): operands built from
    // the tables (B = Toeplitz b, AwT = a by parity class; bf16 hi / lo, fp32)
    DevBuf mx_bh, mx_bl, mx_bf, mx_ah, mx_al, mx_af, s_mx;
    bool mx_ready = false;

    // scratch
    DevBuf s_fbox, s_out1, s_out2, s_vals, s_loss, s_Q, s_nu, s_tmp, s_dt, s_mpart, s_part;
    // cell-tile scratch: f tiles, node values D of a chunk of tiles, per-tile
    // Q0 / L rows, moment partials, correction coefficients
    DevBuf ct_f, ct_D, ct_Q0, ct_L, ct_mp, ct_X;
    int ct_chunk_tiles = 16;     // tiles whose node values are in flight between node and gather
    int ct_node_blocks = 4096;   // grid of the node kernel per tile (grid-stride over destinations)
    int ct_gather_blocks = 4096; // grid of the gather per tile (grid-stride over destinations)
    int ct_wide = 1;             // 1: tiles of 64 cells (a wave per node / destination)

    // step graph cache
    hipGraphExec_t graph_exec = nullptr;
    int g_method = -1, g_cells = 0, g_chunk = 0;
    double g_dt = 0.0;
    unsigned g_flags = 0;
    double *g_f = nullptr;
    unsigned long g_gen = 0;  // devbuf_generation() at capture

    // timing
    bool timing = false;
    std::map<std::string, KernelStat> stats;
    // timed launches: events recorded before and after each, back to back on
    // the stream (no wait between launches), resolved when the stats are read
    struct PendingTime {
        const char *name;
        int a, b;
    };
    std::vector<hipEvent_t> ev_pool;
    int ev_used = 0;
    std::vector<PendingTime> pending;
};

// Records a kernel launch with hipEvents on the context stream when timing is on.
struct KernelTimer {
    gdt_context *h;
    const char *name;
    int a = -1;
    KernelTimer(gdt_context *h_, const char *name_);
    ~KernelTimer();
};

// Tables of kmt_log / kmt_exp (kmath.h), KMT_SIZE doubles, computed on the
// host in extended precision.
void kmath_tables(double *T);
// Dynamic LDS bytes of the kernels that stage f (f + the tables).
int kor_node_lds(const gdt_context *h);
int launch_kmath_pow(gdt_context *h, const double *x, const double *r, double *out, long n);

// Waits for the pending timed launches and adds them to the stats.
int timing_resolve(gdt_context *h);
// Index of a free event of the context's timing pool (-1 if none can be made).
int timing_event(gdt_context *h);

// Start / stop events of a timed launch of the hot-path kernels: passed to
// hipExtLaunchKernelGGL, which records them in the kernel's own dispatch, so
// the elapsed time is the kernel's execution (no extra packets around it).
struct ExtTimer {
    gdt_context *h;
    const char *name;
    int a = -1, b = -1;
    hipEvent_t start = nullptr, stop = nullptr;
    ExtTimer(gdt_context *h_, const char *name_);
    ~ExtTimer();
};

// Launches `kern` on the context stream, through hipExtLaunchKernelGGL with
// the ExtTimer events when timing is on.
#define GDT_LAUNCH_TIMED(h, name, kern, grid, block, lds, ...)                                                  \
    do {                                                                                                       \
        ExtTimer et_((h), (name));                                                                             \
        if (et_.start)                                                                                         \
            hipExtLaunchKernelGGL(kern, grid, block, lds, (h)->stream, et_.start, et_.stop, 0, __VA_ARGS__);  \
        else                                                                                                   \
            hipLaunchKernelGGL(kern, grid, block, lds, (h)->stream, __VA_ARGS__);                              \
    } while (0)

// ---- setup kernels (setup_kernels.hip) -----------------------------------
int launch_build_tables(gdt_context *h, int mode);
int launch_korobov_nodes(gdt_context *h, uint32_t M, uint32_t a, double K, int32_t *d_idx6, double *d_r,
                         double *d_w, int32_t *d_ints, uint8_t *d_valid);
uint32_t host_korobov_pick(uint32_t M, uint64_t seed, int ncand, double *p2);

// ---- collision kernels (collide_kernels.hip) ------------------------------
int launch_loss(gdt_context *h, const double *f_box, double *nu_box, int cells);
int launch_gain(gdt_context *h, const double *f_box, double *gain_box, int cells);
int gain_prepare(gdt_context *h);
int launch_cmp_to_box(gdt_context *h, const double *f_cmp, double *f_box, int cells, int layout);
int launch_box_to_cmp(gdt_context *h, const double *f_box, double *f_cmp, int cells, int layout);
// ---- matrix cores (mfma_gemm.hip, matrix_mfma.hip) ---------------------------
// C (+)= A B^T on the matrix cores; with a workspace of ws_floats FP32
// values, skinny shapes split K over gridDim.z (gemm_k_splits partials).
int launch_gemm_nt(hipStream_t s, int mode, int M, int N, int K, const void *Ah, const void *Al, const void *Bh,
                   const void *Bl, float *C, int ldc, int accumulate, float *ws = nullptr, size_t ws_floats = 0);
int gemm_k_splits(int mode, int M, int N, int K);
int launch_split_f64(hipStream_t s, const double *x, long n, uint16_t *hi, uint16_t *lo, float *f32);
int launch_mx_operands(hipStream_t s, const double *b, const double *a, uint16_t *Bh, uint16_t *Bl, float *Bf,
                       uint16_t *Ah, uint16_t *Al, float *Af);
int launch_mx_gain(hipStream_t s, int mode, const uint16_t *Ah, const uint16_t *Al, const float *Af, const double *f,
                   int cells, float *gpart);
int launch_mx_combine(hipStream_t s, int cells, const float *nuf, int ldn, const float *gpart, double *nu,
                      double *gain);

// ---- uneven batches (batch_abi.hip, korobov_kernels.hip) -------------------
// One velocity grid (class) of a batch: its stencil, the caller's f / Q of
// all its cells, and the scratch of the current segment's cells (slots).
struct KorClassDev {
    int npts, ntile, cap, nparts;  // nparts = KOR_NGROUPS * ntile
    long nn;
    double K;
    const uint4 *nrec;
    const int4 *tiles;
    const int32_t *gptr;
    const uint2 *ent;
    const double *xi;
    const double *kmt;  // log / exp tables
    double *f;     // [cells][npts], updated in place by steps
    double *Qout;  // [cells][npts] or NULL
    double2 *V;    // [slots][nv]
    double *Lp;    // [slots][G][npts]
    double *Qp;    // [slots][G][npts]
    double *mp;    // [slots][20][nparts]
    int node_blocks;  // node work items per cell
    int pad;
    long nv;               // node values per cell (nodes, then lambda value slots)
    const int32_t *cbase;  // first lambda value slot per chunk of node positions
};
// Work item: (class, cell, block within the cell, slot)
int launch_batch_node(hipStream_t s, const KorClassDev *cls, const int4 *items, int nitems, int blocks, int lds);
int launch_batch_gather(hipStream_t s, const KorClassDev *cls, const int4 *items, int nitems, int blocks, int lds);
int launch_batch_apply(hipStream_t s, bool correct, const KorClassDev *cls, const int4 *items, int nitems, int blocks,
                       double dt);
int batch_set_lds();

// ---- Korobov kernels (korobov_kernels.hip) --------------------------------
int kor_blocks(gdt_context *h, int cells);
int kor_setup(gdt_context *h, int cells);
int launch_kor_node(gdt_context *h, const double *f, double *DW, int cells);
int kor_gather_blocks(gdt_context *h, int cells);  // blocks launched per cell (one per tile)
int kor_gather_parts(gdt_context *h);              // moment partials per cell (one per tile)
int launch_kor_gather(gdt_context *h, const double *DW, double *Lpart, double *Qpart, double *mpart, int cells);
// First node position of gather group g: ceil(g nn / G) rounded up to a
// whole chunk of KOR_CHUNK positions, at most nn (group g is
// [kor_group_begin(g), kor_group_begin(g + 1)), whole chunks but the last)
__host__ __device__ constexpr long kor_group_begin(int g, long nn)
{
    return ((((long)g * nn + KOR_NGROUPS - 1) / KOR_NGROUPS + KOR_CHUNK - 1) / KOR_CHUNK * KOR_CHUNK) < nn
               ? (((long)g * nn + KOR_NGROUPS - 1) / KOR_NGROUPS + KOR_CHUNK - 1) / KOR_CHUNK * KOR_CHUNK
               : nn;
}
// LDS copy of f padded to whole 128-double chunks (the node kernels' 16-byte
// staging writes whole wave chunks); the log / exp tables follow the padding
__host__ __device__ constexpr int kor_fpad(int npts) { return (npts + 127) & ~127; }

// Gain-list entry of relative-velocity triple (gi, gj, gk): the box offsets
// in = 2 in1 + (gi & 1), jn, kn (each < its axis size), 10 bits each, so the
// gain kernel splits an entry with shifts and masks.  Axes of the matrix
// tables are at most GLIST_AXIS_MAX.
#define GLIST_AXIS_MAX 1024
__host__ __device__ constexpr int glist_pack(int in, int jn, int kn) { return in | jn << 10 | kn << 20; }
// ---- allocation-free evaluation (noalloc.hip) --------------------------
// Device address of the static scratch array on `device` (current), looked
// up once per device.
int noalloc_setup(int device);

// ---- many-cell path in cell tiles (cell_tiles.hip) ----------------------
int ct_cells_per_tile(const gdt_context *h, bool mixed);
// Scratch for `cells` cells (outside any stream capture).
int ct_prepare(gdt_context *h, int cells, bool mixed);
// Korobov evaluation of `cells` cell-major cells: FP64 node terms, or FP32
// (mixed) with FP64 sums; outputs cell-major (any may be NULL), corrected when
// `correct`; with dt != 0, f_upd <- f + dt Q.
int ct_config(gdt_context *h, int chunk_tiles, int node_blocks, int gather_blocks, int wide);
int ct_run(gdt_context *h, bool mixed, const double *f, int cells, bool correct, double *Qout, double *Lout,
           double *nu_out, double dt, double *f_upd);
int launch_kor_fused(gdt_context *h, const double *f, double *loss, double *Q, double *mpart, int cells);
// G group partials [cell][G][node] of Q and loss (G = 1: plain arrays)
int launch_kor_apply(gdt_context *h, bool correct, int G, int nparts, const double *mpart, const double *f,
                     const double *Lpart, const double *Qpart, double *Qout, double *Lout, double *nu_out, double *X,
                     int cells, double dt, double *f_upd);
int launch_correct(gdt_context *h, int method, const double *f, const double *nu, const double *gain,
                   double *nu_out, double *Q, double *X, int cells, double dt, double *f_upd);
int launch_transpose(gdt_context *h, const double *src, double *dst, int rows, int cols);
int launch_courant_dt(gdt_context *h, const double *nu, long count, double courant, double *dt_out);
int launch_euler(gdt_context *h, double *f, const double *nu, const double *gain, long count, double dt,
                 const double *dt_dev, double *time);
// f32 -> f64 when (s32, d64) given, else f64 -> f32; default stream, synchronous
int launch_convert(const float *s32, double *d64, size_t count, const double *s64, float *d32);
int launch_basis(gdt_context *h, double *dst);
int launch_fill(gdt_context *h, double *dst, long count, double value);
int launch_init_mixture(gdt_context *h, const double *params, int comps, long cells, double *f);
