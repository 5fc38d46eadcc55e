/*
 * gdt.h -- C ABI of the MI355X collision-integral library (libgdt.so).
 *
 * Plain C: opaque handle, plain pointers and sizes, int status codes.  Every
 * velocity-space array argument is DEVICE memory of the handle's GPU, FP64,
 * in the compact sphere-cut node order of the handle's grid (i-j-k order of
 * the kept nodes, velocity_grid.cu:114-126), laid out [cell][node] unless a
 * `layout` argument says GDT_LAYOUT_NODE_MAJOR ([node][cell], the layout of
 * the reference's arrays, determ.cu:116).  Calls are asynchronous on the
 * handle's stream unless stated; gdt_synchronize() waits.
 *
 * Status: 0 = ok, < 0 = error; gdt_last_error() returns the message of the
 * calling thread's last failure.
 *
 * Each entry point names the reference interface it replaces.  The C++
 * drop-in layer (the headers under include/gpu_determ, libgpu_determ.so) keeps the
 * reference's own names and signatures and is built on these calls.
 */
#ifndef GDT_H
#define GDT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDT_OK 0
#define GDT_E_ARG -1      /* invalid argument (null, size, unsupported grid) */
#define GDT_E_HIP -2      /* HIP runtime failure                             */
#define GDT_E_STATE -3    /* required tables / stencil not built             */
#define GDT_E_NOMEM -4    /* device allocation failed                        */
#define GDT_E_NODEVICE -5 /* no HIP device                                   */

#define GDT_LAYOUT_CELL_MAJOR 0 /* [cell][node] */
#define GDT_LAYOUT_NODE_MAJOR 1 /* [node][cell], the reference's layout */

#define GDT_TABLES_FP64 0      /* FP64 spacing and relative speed          */
#define GDT_TABLES_REFERENCE 1 /* FP32 spacing / g / stored values, as the
                                  reference's store_matrix_elements        */

#define GDT_METHOD_MATRIX 0  /* analytic collision matrix (the reference's) */
#define GDT_METHOD_KOROBOV 1 /* Korobov lattice + two-point projection      */

#define GDT_FLAG_CORRECT 1u    /* apply the conservation correction         */
#define GDT_FLAG_GRAPH 2u      /* replay the step sequence as a hipGraph    */
#define GDT_FLAG_TIME 4u       /* time each kernel with hipEvents           */
#define GDT_FLAG_COURANT 8u    /* dt = courant / max nu (relax_f semantics) */
#define GDT_FLAG_FUSED 16u     /* Korobov: one fused kernel recomputing each
                                  node's term per CSR entry, instead of the
                                  node-value + gather pair                  */
#define GDT_FLAG_CELL_TILES 32u /* Korobov, many cells: cells interleaved in
                                  tiles of 64 (16 in the narrow launch shape
                                  of gdt_korobov_cell_tiles) that share every
                                  stencil read (FP64 node terms and sums)   */
#define GDT_FLAG_MIXED 64u     /* Korobov, many cells: the same tiles (32
                                  cells in the narrow shape) with f and the
                                  node terms in FP32, every sum and the
                                  correction in FP64 (relative L2 error of Q
                                  against FP64: about 1e-7 over many cells,
                                  at most 1e-4 per cell)                    */

typedef struct gdt_context *gdt_handle;

/* Velocity grid: n[3] nodes on [vmin, vmax] per axis, sphere cut by R
 * (R <= 0: keep all).  Replaces VelocityGrid::init (velocity_grid.cu:73-126,
 * velocity_grid.hpp:35-54). */
typedef struct gdt_grid_desc {
    int n[3];
    double vmin[3];
    double vmax[3];
    double R;
} gdt_grid_desc;

/* ---- errors, device plumbing ------------------------------------------ */
const char *gdt_last_error(void);
int gdt_version(void);
int gdt_device_count(int *count);
int gdt_malloc(void **dptr, size_t bytes);
int gdt_free(void *dptr);
int gdt_memcpy_htod(void *dst, const void *src, size_t bytes);
int gdt_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int gdt_memcpy_dtod(void *dst, const void *src, size_t bytes);
int gdt_memset(void *dptr, int value, size_t bytes);
/* Waits for all work on the current device (cudaThreadSynchronize in the
 * reference, e.g. collision_integral.cu:600). */
int gdt_device_synchronize(void);
/* Element-wise precision conversion of device arrays (the reference's FP32
 * arrays enter and leave the FP64 core through these). */
int gdt_convert_f32_to_f64(const float *src, double *dst, size_t count);
int gdt_convert_f64_to_f32(const double *src, float *dst, size_t count);

/* ---- context ------------------------------------------------------------ */
/* Creates the grid on `device`: node lists, centres and the moment basis.
 * Replaces VelocityGrid::device() / DeviceVelocityGrid (velocity_grid.cu:
 * 21-48) and init_correction_array (velocity.cu:53-65). */
int gdt_create(gdt_handle *h, const gdt_grid_desc *grid, int device);
int gdt_destroy(gdt_handle h);
int gdt_synchronize(gdt_handle h);
/* Grid facts: n_pnt kept nodes, n_box box nodes, spacing d[3], d3v. */
int gdt_grid_info(gdt_handle h, int *n_pnt, int *n_box, double d[3], double *d3v);
/* Host copies: box coordinates [n_pnt][3] and centres [n_pnt][3]. */
int gdt_grid_nodes(gdt_handle h, int *cidx, double *xi);
/* 14 moment-basis planes over the box (xi_x, xi_y, xi_z, xi^2, xi_x^2, ...,
 * xi^4), FP64 device array of 14*n_box: fill_correction_array
 * (velocity.cu:17-50).  dst is device memory. */
int gdt_basis(gdt_handle h, double *dst);

/* ---- analytic collision matrix (init_matrices, compute_matrix.cu:564-600) */
/* Builds b[n_box] and a[n_box*n_box/8] (reference index order) on the GPU in
 * FP64, and the per-relative-velocity lists of nonzero boxes.  mode is
 * GDT_TABLES_FP64 or GDT_TABLES_REFERENCE.  Requires even n of at most
 * 1,024 on every axis and a box of at most 32,768 nodes.  The gain holds f of
 * a cell in LDS, rows padded to n_w rounded up to 4 mod 8 doubles: the matrix
 * evaluation takes boxes with n_u n_v rw 8 B <= 160 KiB (26^3 at most for a
 * cube); larger ones return GDT_E_ARG there. */
int gdt_build_tables(gdt_handle h, int mode);
/* Device pointers to the FP64 tables owned by the handle. */
int gdt_tables(gdt_handle h, const double **b, const double **a, long *n_a, long *nnz);
/* Replace the handle's tables by caller-provided ones (device FP64, same
 * layout), e.g. tables imported from FP32 reference arrays.  Same axis limits
 * as gdt_build_tables. */
int gdt_set_tables(gdt_handle h, const double *b, const double *a);

/* ---- Korobov lattice stencil (no reference counterpart) ---------------- */
/* Picks a generator from seed (a = 0) or uses `a`, generates the lattice on
 * the GPU, keeps the nodes whose six projection nodes lie in the grid, and
 * builds the destination-sorted CSR.  s0 is the differential cross-section
 * per steradian (1: the reference's 4 pi g total; 0.25: unit diameter).
 * Requires a cubic grid with equal spacing and M a power of two <= 2^26. */
int gdt_korobov_build(gdt_handle h, uint32_t M, uint64_t seed, uint32_t a, double s0);
int gdt_korobov_info(gdt_handle h, uint32_t *M, uint32_t *a, long *nodes, long *entries, double *K);
/* Host copies of the stencil: idx6[nodes][6] (alpha, beta, lambda, mu,
 * lambda+s, mu-s), r[nodes], w[nodes], node[nodes] (lattice number). */
int gdt_korobov_export(gdt_handle h, int32_t *idx6, double *r, double *w, uint32_t *node);
/* Gather layout of the stencil: node groups (one per XCD), destination
 * tiles per group (one moment partial each), LDS entries per gather pass. */
int gdt_korobov_tiles(gdt_handle h, int *groups, int *tiles, int *cap);
/* Sizes of the grouped gather; no reference counterpart (layout figures for
 * byte counts).  The node kernel forms the prefix sums of the lambda terms
 * over each chunk of 16 node positions and writes each lambda piece (a run
 * of one lambda destination within a chunk) as value slots: +prefix at its
 * last position and -prefix before its first, unless it starts the chunk.
 * The gather reads each slot as one entry in place of the piece's lambda
 * entries.  *entries is the entries of the node groups' CSRs, *slots the
 * value slots (node values per cell: nodes + slots). */
int gdt_korobov_gather_info(gdt_handle h, long *entries, long *slots);
/* Limit of gather tiles per node group (default 64, 1..1024); no reference
 * counterpart (a launch-shape setting of the gather).  A tile holds at most
 * 512 destinations, so a group has at least ceil(grid nodes / 512) tiles
 * whatever the limit (34 at 32^3).  Fewer, larger tiles give the apply fewer
 * moment partials and the gather less per-tile work; more tiles put more
 * blocks in flight.  On a built stencil the tiles are
 * rebuilt at once (after the context's stream drains).  Per-destination sums
 * do not depend on the tiling; the moment sums do, in their last bits.  A
 * batch holding this context must be configured again (gdt_batch_config);
 * gdt_batch_tune sets it on the batch's classes. */
int gdt_korobov_set_tiles(gdt_handle h, int tiles_max);
/* Launch shape of the cell-tile paths (GDT_FLAG_CELL_TILES / GDT_FLAG_MIXED),
 * no reference counterpart: tiles whose node values are in flight between
 * the node and gather kernels (1..64), grid sizes of the node and gather
 * kernels per tile, and the tile width (wide = 0: 16 cells in FP64, 32 in
 * FP32; wide = 1: 64 cells, a wave per node and per destination).  A
 * destination's sums follow the CSR order whatever the shape, so results are
 * the same bits for every chunk and grid size; the tile width changes the
 * order of the moment sums, so the corrected Q in its last bits.  Defaults:
 * 16 tiles (fewer where a chunk's node values would pass 4 GiB), node grid
 * 4096, gather grid 4096, wide. */
int gdt_korobov_cell_tiles(gdt_handle h, int chunk_tiles, int node_blocks, int gather_blocks, int wide);
/* Accuracy hook, no reference counterpart: out[i] = x[i]^r[i] as the node
 * terms compute it (table-driven FP64 log and exp); device pointers, x > 0;
 * synchronous. */
int gdt_kmath_pow(gdt_handle h, const double *x, const double *r, double *out, long n);
/* Host copy of those log / exp tables (520 doubles: c_j, -log c_j high and
 * low for j < 128; 2^(j/64) high and low for j < 64; ln2 high and low, 64 /
 * ln2, ln2 / 64 high and low; zero padding).  No device needed. */
int gdt_kmath_tables(double *T);
/* Host copy of the CSR: ptr[n_pnt+1], ent[entries] = node*4 + role. */
int gdt_korobov_export_csr(gdt_handle h, int32_t *ptr, uint32_t *ent);
/* Generator choice alone (host): same rule as gdt_korobov_build. */
uint32_t gdt_korobov_pick(uint32_t M, uint64_t seed, int ncand, double *p2);

/* ---- collision operator ------------------------------------------------- */
/* Cells per call: at most 65,535 on the analytic-matrix path and the grouped
 * Korobov path (one launch holds every cell), and 65,535 tiles of 64 cells
 * with GDT_FLAG_CELL_TILES / GDT_FLAG_MIXED; more return GDT_E_ARG (split
 * the cells into several calls).  The reference runs 256 cells per launch
 * (npx, collision_integral.cu:589). */
/* Matrix method: nu (loss frequency, evolve_direct_colisions,
 * collision_integral.cu:32-72) and gain (evolve_inverse_colisions, :75-188,
 * all 512 boxes, intended output slot) for `cells` cells.  Either output may
 * be NULL. */
int gdt_collide_matrix(gdt_handle h, const double *f, double *nu, double *gain, int cells, int layout);
/* Korobov method: loss = f nu and Q (uncorrected) for `cells` cells.
 * Requires at most 19,200 grid nodes (f of a cell is staged in LDS). */
int gdt_collide_korobov(gdt_handle h, const double *f, double *loss, double *Q, int cells, int layout);
/* Conservation correction (before_correction / linear_solver /
 * correct_frequency, collision_integral.cu:217-526), for every cell:
 * method MATRIX: in f, nu, gain -> nu_corr (may alias nu), Q;
 * method KOROBOV: in loss (passed as nu), Q (in/out, passed as gain == Q).
 * X (5 per cell, device) may be NULL. */
int gdt_correct(gdt_handle h, int method, const double *f, double *nu, const double *gain, double *Q,
                double *X, int cells, int layout);
/* One full evaluation = evolve_colisions (collision_integral.cu:550-773):
 * collide, then correct if flags & GDT_FLAG_CORRECT.  Outputs: nu (corrected
 * loss frequency; KOROBOV: loss/f), gain (KOROBOV: Q + loss), Q. */
int gdt_evaluate(gdt_handle h, int method, const double *f, double *nu, double *gain, double *Q, int cells,
                 int layout, unsigned flags);
/* Allocation-free variant of gdt_evaluate for the Korobov method, with no
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and results as gdt_evaluate(h,
 * GDT_METHOD_KOROBOV, f, nu, gain, Q, cells, GDT_LAYOUT_CELL_MAJOR, flags),
 * bit for bit; flags is 0 or GDT_FLAG_CORRECT; nu and gain may be NULL (gain
 * needs nu).  Its scratch (node values, group partials, moment sums) is one
 * static 64 MiB device array per GPU: calls needing more (more than 6 cells
 * of a 32^3 grid with 2^21 lattice nodes) return GDT_E_ARG.  Every
 * call on a GPU shares that array, so calls that overlap on one GPU race.
 * Calls no hipMalloc or hipFree; asynchronous on the handle's stream. */
int gdt_evaluate_noalloc(gdt_handle h, const double *f, double *nu, double *gain, double *Q, int cells,
                         unsigned flags);
/* relax_f (collision_integral.cu:835-880): f <- f (1 - nu dt) + gain dt for
 * all cells, with dt = min courant / nu over every value (flags &
 * GDT_FLAG_COURANT, get_delta_t :776-807) or the given dt.  time (device
 * double, may be NULL) += dt.  Element-wise, so any layout. */
int gdt_relax(gdt_handle h, double *f, const double *nu, const double *gain, int cells, double dt, double courant,
              double *time, unsigned flags);
/* `steps` explicit steps df/dt = Q(f,f) of `cells` cells (cell-major),
 * each: collide, correct (if flags & GDT_FLAG_CORRECT), f += dt Q.  With
 * GDT_FLAG_GRAPH the step is captured once as a hipGraph and replayed. */
int gdt_run_steps(gdt_handle h, int method, double *f, int cells, double dt, int steps, unsigned flags);

/* ---- matrix cores (no reference counterpart) ----------------------------- */
#define GDT_GEMM_BF16 0       /* bf16 operands                                   */
#define GDT_GEMM_BF16_SPLIT 1 /* operands split into bf16 hi + lo, 3 products    */
#define GDT_GEMM_F32 2        /* fp32 operands (fp32 MFMA)                       */
/* C[M][N] (+)= sum_k A[M][K] B[N][K] on MFMA with FP32 accumulation; device
 * operands (bf16 as uint16 bit patterns, or float), M and N multiples of 128,
 * K a multiple of 32 (bf16) / 16 (fp32).  Default stream, synchronous. */
int gdt_gemm_nt(int mode, int M, int N, int K, const void *A_hi, const void *A_lo, const void *B_hi,
                const void *B_lo, float *C, int ldc, int accumulate);
/* FP64 -> bf16 pair (hi = bf16(x), lo = bf16(x - hi)) and / or FP32; any
 * output may be NULL.  Default stream, synchronous. */
int gdt_split_f64(const double *x, long n, uint16_t *hi, uint16_t *lo, float *f32);
/* The analytic-matrix operator of evolve_colisions (collision_integral.cu:
 * 550-773) for many cells of a 16^3 box grid (R keeping every node, tables
 * built), as dense contractions on the matrix cores: the loss a GEMM against
 * the Toeplitz b matrix, the gain a GEMM per parity class of v + v1 over the
 * a table (mode: GDT_GEMM_*), FP32 accumulation; FP64 per-cell sums, Q and
 * conservation correction (flags & GDT_FLAG_CORRECT).  f, nu, gain, Q:
 * device FP64 [cells][4096]; outputs as gdt_evaluate's (nu corrected). */
int gdt_matrix_mfma(gdt_handle h, int mode, const double *f, double *nu, double *gain, double *Q, int cells,
                    unsigned flags);

/* ---- learned surrogate (no reference counterpart) ------------------------
 * Bilinear MLP on the node values of the grid of h: x = standardised f / n,
 * h = (W1 x + b1) * (W2 x + b2) elementwise, W3 h + b3 -> standardised
 * Q / n^2 (n = density), bf16 weights and activations on MFMA with FP32
 * accumulation, FP32 master weights and Adam; hidden a multiple of 128.
 * The handle h must outlive the surrogate. */
typedef struct gdt_surrogate *gdt_surrogate_handle;
int gdt_surrogate_create(gdt_surrogate_handle *s, gdt_handle h, int hidden, uint64_t seed);
int gdt_surrogate_destroy(gdt_surrogate_handle s);
/* Trains on ns pairs (device FP64 f, Q [ns][n_pnt]); loss_out: mean square
 * error (standardised units) of the last sampled step.  Synchronous. */
int gdt_surrogate_train(gdt_surrogate_handle s, const double *f, const double *Q, long ns, int epochs, int batch,
                        double lr, double *loss_out);
/* Network prediction only (no projection): Q [cells][n_pnt] device FP64. */
int gdt_surrogate_predict(gdt_surrogate_handle s, const double *f, double *Q, int cells);
/* Prediction, exact conservation projection (FP64 5x5 per cell, weights f),
 * residual r = |Q - Q_nn| / |Q| (at least 1 if sum Q log f > 0) and, for r >
 * threshold, the plain solver (corrected Korobov evaluation on the stencil of
 * h, in FP64 cell tiles) in place of the network.  fallback (device int [cells], may be NULL)
 * marks those cells; *n_fallback counts them.  Synchronous. */
int gdt_surrogate_infer(gdt_surrogate_handle s, const double *f, double *Q, int cells, double threshold,
                        int *fallback, int *n_fallback);
/* Second network for the residual check of s (same context; NULL clears):
 * gdt_surrogate_infer then predicts with both, keeps their mean, and raises
 * each cell's residual to |Q_a - Q_b| / |Q| (their disagreement, an estimate
 * of the error against the plain solver).  The partner must outlive s's use
 * of it. */
int gdt_surrogate_set_partner(gdt_surrogate_handle s, gdt_surrogate_handle partner);
/* Host copy of the residuals of the last gdt_surrogate_infer call. */
int gdt_surrogate_residuals(gdt_surrogate_handle s, double *res, int cells);
/* Host copies of the FP32 master weights ([hidden][np], [hidden],
 * [hidden][np], [hidden], [np][hidden], [np]) and the standardisation
 * (mu_x, sd_x, mu_y, sd_y, each [np]); any may be NULL.  np = n_pnt rounded
 * up to 128. */
int gdt_surrogate_export(gdt_surrogate_handle s, float *W1, float *b1, float *W2, float *b2, float *W3, float *b3,
                         float *norm);
int gdt_surrogate_info(gdt_surrogate_handle s, int *n, int *np, int *hidden, long *params);

/* ---- uneven batches (no reference counterpart: cells of several grids) ---
 * Classes are handles with a Korobov stencil built; f[s] is device FP64
 * [cells[s]][n_pnt of class s] (cell-major), updated in place by steps.
 * Each run step evaluates every cell (corrected if GDT_FLAG_CORRECT) and,
 * for dt != 0, applies f += dt Q: per segment of cells, three work-list
 * kernels, or with GDT_FLAG_CELL_TILES / GDT_FLAG_MIXED every class's cells
 * in cell tiles (each class's stencil read once per tile of 64 cells); with
 * GDT_FLAG_GRAPH the whole step is one hipGraph.  Every cell's result is
 * bitwise that of gdt_run_steps on the cell alone with the same flags.
 * The batch keeps the f / Q pointers and the class handles: they must
 * outlive it. */
typedef struct gdt_batch *gdt_batch_handle;
int gdt_batch_create(gdt_batch_handle *b, int nclass, const gdt_handle *classes, double *const *f,
                     const int *cells);
int gdt_batch_destroy(gdt_batch_handle b);
/* Q of the last run step per class ([cells][n_pnt], device), or NULL. */
int gdt_batch_set_output(gdt_batch_handle b, double *const *Q);
/* Explicit launch shape: cells per segment, nodes per thread of the node kernel (>= 2). */
int gdt_batch_config(gdt_batch_handle b, int seg_cells, int node_q);
/* Autotuner: on the batch's cells (evaluation only, f unchanged) times the
 * classes' gather tile limit (32, 64, 128; set on every class with
 * gdt_korobov_set_tiles, so single runs of those contexts keep matching the
 * batch bitwise), then every (segment size, nodes per thread) candidate, and
 * keeps the fastest.  With GDT_FLAG_CELL_TILES or GDT_FLAG_MIXED in flags it
 * tunes the cell tiles instead: chunk of tiles (4, 8, 16) x node grid (2048,
 * 4096) x gather grid (2048, 4096), set on every class as
 * gdt_korobov_cell_tiles does with the classes' own tile width (results do
 * not change); *seg_cells then returns the cells per chunk and *node_q the
 * gather grid. */
int gdt_batch_tune(gdt_batch_handle b, unsigned flags, int *seg_cells, int *node_q, double *best_ms);
int gdt_batch_run(gdt_batch_handle b, double dt, int steps, unsigned flags);
int gdt_batch_synchronize(gdt_batch_handle b);
int gdt_batch_info(gdt_batch_handle b, int *seg_cells, int *node_q, int *segments);

/* ---- spatial transport (no reference counterpart) ------------------------ */
#define GDT_BC_PERIODIC 0
#define GDT_BC_ZERO_GRADIENT 1 /* a missing neighbour takes the cell's own value */
/* One explicit first-order upwind step of df/dt + xi . grad_x f = 0 over a box
 * of nx x ny x nz cells (cell c = (ix ny + iy) nz + iz), f and f_out FP64
 * [cell][node] device arrays that must not overlap; spacing[3] the cell
 * sizes, bc[3] the boundary of each axis.  Per axis a, k = dt xi_a / h_a and
 * f_out = f - sum_a k (f - f[c - e_a]) for k > 0, k (f[c + e_a] - f)
 * otherwise (axes x, y, z in that order).  Rejects dt with
 * max_v sum_a |k_a| > 1.  Periodic boundaries conserve every velocity's cell
 * sum.  Asynchronous on the handle's stream.  This is the step a kinetic
 * solver alternates with the collision operator (gdt_run_steps). */
int gdt_transport(gdt_handle h, const double *f, double *f_out, int nx, int ny, int nz, const double *spacing,
                  double dt, const int *bc);

/* ---- initial conditions ---------------------------------------------------- */
/* Sums of Maxwellians on the handle's grid (the reference's init_f is one
 * component, determ.cu:101-121): f[c][v] = sum_k n (2 pi T)^(-3/2)
 * exp(-|xi_v - u|^2 / 2T) with params[c][k] = (n, T, ux, uy, uz), device FP64;
 * a component with n = 0 is skipped.  Asynchronous on the handle's stream. */
int gdt_init_mixture(gdt_handle h, const double *params, int comps, long cells, double *f);

/* ---- measurement --------------------------------------------------------- */
/* Kernel timing under GDT_FLAG_TIME: total milliseconds and launch count of
 * the kernel `name` since the last reset.  Each timed launch is bracketed by
 * two events on the context stream, back to back with the other launches (no
 * wait in between); reading the stats waits for them. */
int gdt_kernel_stats(gdt_handle h, const char *name, double *total_ms, long *count);
int gdt_kernel_stats_reset(gdt_handle h);
/* Layout transposes (transpose_float, collision_integral.cu:528-547):
 * [rows][cols] -> [cols][rows], device FP64. */
int gdt_transpose(gdt_handle h, const double *src, double *dst, int rows, int cols);

#ifdef __cplusplus
}
#endif
#endif
