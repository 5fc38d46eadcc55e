// Many-cell Korobov operator in cell tiles (gfx950, wave64).
//
// The single-cell path reads every stencil record and entry once per cell and
// gathers 16-byte node values one request per entry.  Here K cells share each
// of those reads: a tile holds K cells interleaved [node][K], so one row
// carries a node's value for all K cells, and one stencil record or entry
// serves the K lanes of a row.  By default K = 64: a wave owns one node (node
// kernel) or one destination (gather), so records, row offsets and the
// branch on an entry's role are wave-uniform (scalar loads and scalar
// address arithmetic); narrow tiles (K = 16 in FP64, 32 in FP32, lane groups
// of K within a wave) remain as a launch-shape option.
//
// Node-term type T: double (the plain method: the same operations as the
// single-cell node kernel, so the node values are the same bits) or float
// (the mixed form: f and the node term D in FP32 with the hardware's log2 /
// exp2; every sum and the correction in FP64).
//
// Per evaluation of C cells in T = ceil(C / K) tiles (cells past C are zero):
//   k_ct_pack     f [cell][node] FP64  ->  ft[tile][node][K] of T
//   k_ct_node     per chunk of tiles: D of every (node, cell) into
//                 D[tile][q][K] (q: the node's position in the tile order,
//                 sorted by the alpha destination); r and w come precomputed
//                 per node.  A destination's alpha nodes are one run of
//                 positions, walked by one lane group, which also sums the
//                 run's alpha terms (-D into Q0, 2 w f_beta into the loss)
//   k_ct_gather   per chunk: the K lanes of a destination start from those
//                 sums and walk its entries of the destination-sorted tile CSR
//                 (beta, lambda, mu, lambda+s, mu-s) in order (one FP64 sum
//                 per cell, no partials over node groups): Q0[tile][d][K] =
//                 sum coef D and the loss L[tile][d][K] = f_d sum 2 w
//                 f_partner over the entries where d is alpha or beta (the
//                 loss is linear in f given f_d, so it reads the f tile
//                 instead of a stored 2 w P); destinations are taken longest
//                 first, so concurrent waves have similar entry counts
//   k_ct_moments  per (destination range, tile): the 20 moment sums of the
//                 range, one partial per cell and range
//   k_ct_solve    per cell: the partials in a fixed order, the 5x5 solve: X
//   k_ct_apply    per (destination range, tile): Q = Q0 - L X.psi, loss,
//                 corrected nu, f += dt Q, written cell-major through an LDS
//                 transpose
// No atomics: every sum has a fixed order, so results are reproducible run to
// run and do not depend on the launch shape or on the other cells of a tile.
// Chunks of tiles bound the node values in flight between the node and
// gather kernels (16 tiles of 64 cells by default: 1.65 GB at 32^3 and 2^21
// nodes in FP32, 3.3 GB in FP64; fewer where a chunk would pass the 4 GiB
// that 32-bit offsets reach; the batch autotuner picks 4, 8 or 16).
//
// What bounds them (counters at 32^3, 2^21 nodes): the gather reads each
// node's D row about 4.7 times (once per non-alpha entry, scattered over the
// grid) and each beta entry's partner row of f; a tile's rows (103 MB in
// FP32) do not stay in L2, and 71% (FP32) to 80% (FP64) of its L2 requests
// miss even with the octant order.  Its FETCH_SIZE + WRITE_SIZE, scaled by a
// copy of known size, come to 7.8 TB/s: the memory side's rate.  The node
// kernel writes each row once and reads the f tile (4.4 MB in FP32, a little
// over one XCD's L2) at 6.3 (FP32) to 7.35 TB/s (FP64), about half its
// instruction issue busy.
#include <algorithm>

#include "gdt_internal.h"
#include "kernel_common.h"
#include "kmath.h"

#define CT_THREADS 256
#define CT_MOM_PARTS 64   // destination ranges (moment partials) per tile
// entries whose loads a lane of the gather issues together: 12 FP64 rows
// (2.216 -> 2.137 us per cell against 8; 16: 2.153), 8 FP32 rows (12: 1.333
// -> 1.357, 16: 1.50)
// entries whose D rows a lane of the FP32 gather loads together past the beta
// run (16: 1.088 us per cell; 12: 1.100; 8: 1.104)
#define CT_UNROLL_D 16
// nodes of an alpha run a wave evaluates together in the node kernel, their
// loads in flight at once (us per cell FP64 / FP32: 4: 1.726 / 0.732;
// 2: 1.725 / 0.797; 3: 1.750 / 0.762; 6: 1.726 / 0.776; 8: 1.92 / 0.92)
#define CT_NODE_UNROLL 4
template <class T> struct CtUnroll {
    static constexpr int value = sizeof(T) == 8 ? 12 : 8;
};

template <class T> struct CtK;
template <> struct CtK<double> {
    static constexpr int K = 16;
};
template <> struct CtK<float> {
    static constexpr int K = 32;
};

// base[idx] with a 32-bit byte offset: a uniform base and a per-lane 32-bit
// offset (global_load ... saddr), no 64-bit address arithmetic per lane.  The
// scratch the offsets index stays below 4 GiB (checked by ct_prepare).
template <class T>
__device__ __forceinline__ T ld_off(const T *__restrict__ base, uint32_t idx)
{
    return *(const T *)((const char *)base + (size_t)(uint32_t)(idx * (uint32_t)sizeof(T)));
}

// ---------------------------------------------------------------------------
// f tiles
// ---------------------------------------------------------------------------
// Block (x, t): nodes [64 x, 64 x + 64) of tile t, through an LDS tile: rows of
// 64 contiguous nodes in, rows of K contiguous cells out.
template <class T, int K>
__global__ __launch_bounds__(CT_THREADS) void k_ct_pack(int npts, int cells, const double *__restrict__ f,
                                                        T *__restrict__ ft)
{
    __shared__ T tile[K][65];
    const int v0 = blockIdx.x * 64, t = blockIdx.y;
    const int lane = threadIdx.x % 64;
    for (int k = threadIdx.x / 64; k < K; k += CT_THREADS / 64) {
        const long c = (long)t * K + k;
        const int v = v0 + lane;
        tile[k][lane] = (c < cells && v < npts) ? (T)f[c * npts + v] : (T)0;
    }
    __syncthreads();
    T *out = ft + ((size_t)t * npts + v0) * K;
    for (int i = threadIdx.x; i < 64 * K; i += CT_THREADS) {
        const int vl = i / K, k = i % K;
        if (v0 + vl < npts) out[i] = tile[k][vl];
    }
}

// ---------------------------------------------------------------------------
// node terms
// ---------------------------------------------------------------------------
// Packed node record x = (a | b << 15 | D2 >> 10 << 30, l | m << 15,
// ls | ms << 15, num / 4 | den / 4 << 11 | (D2 & 1023) << 22): the six node
// indices, and num = 0 where r = 0.
__device__ __forceinline__ int ct_num(uint32_t w) { return (int)(w & 0x7FFu); }

// The six f values of a node for one lane: alpha's (a, the same for a whole
// alpha run, so the caller passes it in), beta's, and the projection pairs'.
template <class T> struct CtF {
    T a, b, l, m, ls, ms;
};
template <class T, class LD>
__device__ __forceinline__ CtF<T> ct_load(const uint4 x, T fa, LD ld)
{
    CtF<T> v;
    v.a = fa;
    v.b = ld((x.x >> 15) & 0x7FFFu);
    v.l = ld(x.y & 0x7FFFu);
    v.m = ld(x.y >> 15);
    v.ls = ld(x.z & 0x7FFFu);
    v.ms = ld(x.z >> 15);
    return v;
}

// FP64: the operations of the single-cell node kernel (table-driven log and
// exp from ts, contraction off), so D is the same bits.
// r = num / den and w = K sqrt(D2) come precomputed per node (rw: the same
// correctly rounded values the single-cell kernel forms).
__device__ __forceinline__ double ct_terms(const uint4 x, const double2 rw, const CtF<double> &v, const double *ts)
{
#pragma clang fp contract(off)
    const int num = ct_num(x.w);
    const double P = v.a * v.b;
    const double Pl = v.l * v.m;
    double R;
    if (num == 0) {
        R = Pl;
    } else {
        const double r = rw.x;
        const double Ps = v.ls * v.ms;
        if (Pl > 0.0 && Ps > 0.0)
            R = Pl * kmt_exp(r * kmt_log(Ps / Pl, ts), ts);
        else
            R = (1.0 - r) * Pl + r * Ps;
    }
    return rw.y * (P - R);
}

// FP32: the same formula in single precision with the hardware's
// transcendentals.  The ratio Ps / Pl is split by frexp into mantissas (whose
// quotient lies in (1/2, 2): v_log_f32 never sees a denormal) and an exact
// integer exponent difference; the quotient is one reciprocal refined by a
// Newton step (within an ulp of the rounded quotient); r and w come per
// node, rounded to FP32.  R = Pl 2^(r log2(Ps / Pl)); an exponent below -126
// gives 0 (a value below 2^-126 Pl, far below the sums' rounding).  Products
// that are not positive take the linear form, as in FP64.
__device__ __forceinline__ float ct_terms(const uint4 x, const float2 rw, const CtF<float> &v, const double *ts)
{
    (void)ts;
    const int num = ct_num(x.w);
    const float P = v.a * v.b;
    const float Pl = v.l * v.m, Ps = v.ls * v.ms;
    const float r = rw.x;
    float R = Pl;
    if (num != 0) {
        if (Pl > 0.0f && Ps > 0.0f) {
            const float ms = __builtin_amdgcn_frexp_mantf(Ps), ml = __builtin_amdgcn_frexp_mantf(Pl);
            const int es = __builtin_amdgcn_frexp_expf(Ps), el = __builtin_amdgcn_frexp_expf(Pl);
            const float rc = __builtin_amdgcn_rcpf(ml);
            float qm = ms * rc;
            qm = fmaf(fmaf(-ml, qm, ms), rc, qm);
            const float lg = __builtin_amdgcn_logf(qm) + (float)(es - el);
            R = Pl * __builtin_amdgcn_exp2f(r * lg);
        } else {
            R = (1.0f - r) * Pl + r * Ps;
        }
    }
    return rw.y * (P - R);
}

// One term of a destination's FP64 sum q: the FP64 tiles keep the
// single-cell path's term (coef D, then the sum), the mixed form one FMA.
template <class T>
__device__ __forceinline__ double ct_acc(double q, double coef, T d)
{
#pragma clang fp contract(off)
    if (sizeof(T) == 8) return q + coef * (double)d;
    return fma(coef, (double)d, q);
}

// Lane group of K lanes: its group index, made uniform (an SGPR) when the
// group is the whole wave, so that records, node rows and their addresses
// are scalar.
template <int K>
__device__ __forceinline__ int ct_group()
{
    if (K == WAVE) return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    return (int)(threadIdx.x / K);
}

// Lane k's value of grid node i in the rows of `fr` (K values per row): a row
// pointer when the group is the wave (uniform i: a scalar base, one lane
// offset), a 32-bit offset load otherwise.
template <int K, class T>
__device__ __forceinline__ T ct_row(const T *__restrict__ fr, uint32_t i, int k)
{
    if (K == WAVE) return (fr + (size_t)i * K)[k];
    return ld_off(fr, i * K + k);
}

// Block (x, j) of a chunk: tile t0 + j; lane group g = tid / K takes the
// destination slots s = x NPB + g, + gridDim.x NPB, ... (NPB = CT_THREADS / K)
// of the longest-first order aperm; lane k of the group is cell k of the
// tile.  For destination a the group walks a's alpha run, positions
// [aptr[a], aptr[a + 1]) of the tile node order, CT_NODE_UNROLL nodes per
// pass (their loads in flight together): it stores each node's D row, and
// sums the run's alpha terms of Q0 (coefficient -1) and of the loss (2 w
// f_beta, FP64 w) in the CSR's term form; the sums start the gather's sums
// of destination a (Q0 and L of the tile, L before its factor f_a).  Record
// loads are one
// address per group; the f loads and the value stores are K contiguous
// elements (a 64- or 128-byte row).
// Dv holds the chunk's D values [j][q][K].
template <class T> struct CtRW;
template <> struct CtRW<double> {
    typedef double2 type;
};
template <> struct CtRW<float> {
    typedef float2 type;
};

// (Slices of K < 64 cells per launch keep a slice of the f tile in L2, but
// lose the wave-uniform records: 0.90 and 0.98 us per cell for halves and
// quarters against 0.84 us.  Slices per XCD in one launch -- block b on cells
// (b % 8) % 2 of every row, so each XCD reads a 2.2 MB half of the f tile --
// measured 0.85 against 0.77 us in FP32 and 1.86 against 1.69 in FP64
// quarters, with the same bits.)
template <class T, int K>
__global__ __launch_bounds__(CT_THREADS) void k_ct_node(int npts, int nn, const int32_t *__restrict__ aptr,
                                                        const int32_t *__restrict__ aperm,
                                                        const uint4 *__restrict__ nrec,
                                                        const typename CtRW<T>::type *__restrict__ nrw,
                                                        const double2 *__restrict__ rw64,
                                                        const T *__restrict__ ft, int t0, T *__restrict__ Dv,
                                                        double *__restrict__ Q0, double *__restrict__ L,
                                                        const double *__restrict__ kmt)
{
#pragma clang fp contract(off)
    __shared__ double ts[KMT_SIZE];
    if (sizeof(T) == 8)
        for (int i = threadIdx.x; i < KMT_SIZE; i += CT_THREADS) ts[i] = kmt[i];
    __syncthreads();
    constexpr int NPB = CT_THREADS / K;
    const int k = threadIdx.x % K, g = ct_group<K>();
    const T *fr = ft + (size_t)(t0 + blockIdx.y) * npts * K;
    T *Dt = Dv + (size_t)blockIdx.y * nn * K;
    const size_t to = (size_t)(t0 + blockIdx.y) * npts * K;
    auto ld = [&](uint32_t i) { return ct_row<K>(fr, i, k); };
    // streaming stores for the wide rows, so the rows do not push the f tile
    // out of L2 (node kernel 0.805 -> 0.783 us per cell in FP32, 1.764 ->
    // 1.655 in FP64; the gather is unchanged, and streaming loads there are
    // slower: 1.62 -> 1.89)
    auto store = [&](int q, T d) {
        if (K == WAVE)
            __builtin_nontemporal_store(d, &(Dt + (size_t)q * K)[k]);
        else
            *(T *)((char *)Dt + (size_t)(uint32_t)(((uint32_t)q * K + k) * (uint32_t)sizeof(T))) = d;
    };
    for (int s = blockIdx.x * NPB + g; s < npts; s += gridDim.x * NPB) {
        const int a = aperm[s];
        const int q0 = aptr[a], q1 = aptr[a + 1];
        const T fa = ld((uint32_t)a);
        double qs = 0.0, ls = 0.0;
        constexpr int NU = CT_NODE_UNROLL;
        for (int q = q0; q < q1; q += NU) {
            uint4 x[NU];
            int qq[NU];
#pragma unroll
            for (int u = 0; u < NU; u++) {
                qq[u] = q + u < q1 ? q + u : q1 - 1;  // clamped: in bounds
                x[u] = nrec[qq[u]];
            }
            CtF<T> v[NU];
#pragma unroll
            for (int u = 0; u < NU; u++) v[u] = ct_load(x[u], fa, ld);
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const T d = ct_terms(x[u], nrw[qq[u]], v[u], ts);
                if (q + u < q1) {
                    store(qq[u], d);
                    qs = ct_acc(qs, -1.0, d);
                    ls += (2.0 * rw64[qq[u]].y) * (double)v[u].b;
                }
            }
        }
        Q0[to + (size_t)a * K + k] = qs;
        L[to + (size_t)a * K + k] = ls;
    }
}

// ---------------------------------------------------------------------------
// gather
// ---------------------------------------------------------------------------
// Lane group g of block (x, j) takes destination slots s = x NPB + g, +
// gridDim.x NPB, ... of the longest-first order dperm; lane k sums cell k of
// the destination: q and l start from the node kernel's sums over d's alpha
// run, then add the tile CSR's entries in order: q += coef D (coef -1 for
// role 0, beta) and l += 2 w f_alpha over the role-0 entries; L = f_d l.
// FP64 products and sums; the FP64 tiles keep the single-cell path's term
// for q (coef D, then the sum), the mixed form one FMA per term.  Each round
// issues the loads of several entries (records, then D rows and f rows)
// before adding any of them; entries past a run's end are clamped to its
// last one and add exact zeros (+-0 leaves a sum unchanged).  The FP64
// gather walks the entries in (q, slot) order and loads a partner's f row on
// the beta entries only (the role is wave-uniform); the FP32 gather reads a
// copy of the tile CSR with each destination's beta entries first and runs
// the beta run (D and f rows) and then the rest (D rows only).
// Destination order: with a grid of a multiple of 8 blocks (workgroups are
// dealt to the XCDs round robin), block b takes its slots from XCD b % 8's
// list, one octant of the box in Morton order (xperm / xo), else from the
// longest-first order dperm.  A node's row is read by lambda and mu and by
// their grid neighbours lambda + s and mu - s, so on one XCD, working
// through a compact part of the grid, the second read often finds the row
// in L2: FP64 gather 2.475 -> 2.224 us per cell against longest first (the
// FP32 gather, then still loading an f row for every entry, was bound by its
// L2 request count and unchanged at 1.33).
struct CtXcd {
    int off[9];
};
template <class T, int K>
__global__ __launch_bounds__(CT_THREADS) void k_ct_gather(int npts, int nn, const int32_t *__restrict__ cptr,
                                                          const int32_t *__restrict__ cmid,
                                                          const uint4 *__restrict__ cent,
                                                          const int32_t *__restrict__ dperm,
                                                          const int32_t *__restrict__ xperm, CtXcd xo,
                                                          const T *__restrict__ Dv, const T *__restrict__ ft, int t0,
                                                          double *__restrict__ Q0, double *__restrict__ L)
{
#pragma clang fp contract(off)
    constexpr int NPB = CT_THREADS / K;
    const int k = threadIdx.x % K, g = ct_group<K>();
    const T *Dt = Dv + (size_t)blockIdx.y * nn * K;
    const T *fr = ft + (size_t)(t0 + blockIdx.y) * npts * K;
    const size_t to = (size_t)(t0 + blockIdx.y) * npts * K;
    // grids of a multiple of 8 blocks: block b walks XCD b % 8's octant list
    const bool xcd = (gridDim.x & 7u) == 0u;
    const int xc = blockIdx.x & 7, lo = xcd ? xo.off[xc] : 0, cnt = xcd ? xo.off[xc + 1] - lo : npts;
    const int sb = xcd ? (int)(blockIdx.x >> 3) : (int)blockIdx.x, nb = xcd ? (int)(gridDim.x >> 3) : (int)gridDim.x;
    const int32_t *perm = xcd ? xperm + lo : dperm;
    for (int s = sb * NPB + g; s < cnt; s += nb * NPB) {
        const int d = perm[s];
        const int e0 = cptr[d], em = cmid[d], e1 = cptr[d + 1];
        // the node kernel's sums of d's alpha run come first
        double q = Q0[to + (size_t)d * K + k], l = L[to + (size_t)d * K + k];
        constexpr int U = CtUnroll<T>::value;
        if (sizeof(T) == 8) {
            // FP64 rows: one pass over [e0, e1); with a wave per destination
            // the role is uniform, so only the beta entries load their
            // partner's f row (no divergence; narrow tiles load every row).
            // The two-run form below is slower here (2.22 against 2.15 us
            // per cell).
            for (int e = e0; e < e1; e += U) {
                uint4 x[U];
#pragma unroll
                for (int u = 0; u < U; u++) {  // clamped to the destination's last entry: in bounds
                    const uint32_t eu = (uint32_t)min(e + u, e1 - 1);
                    x[u] = K == WAVE ? cent[eu] : ld_off(cent, eu);
                }
                T dv[U], fv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    dv[u] = ct_row<K>(Dt, x[u].x >> 2, k);
                    fv[u] = (K != WAVE || (x[u].x & 3u) == 0u) ? ct_row<K>(fr, x[u].y, k) : (T)0;
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool in = e + u < e1, role0 = (x[u].x & 3u) == 0u;
                    const double val = __hiloint2double((int)x[u].w, (int)x[u].z);
                    q = ct_acc(q, in ? (role0 ? -1.0 : val) : 0.0, dv[u]);
                    l += (in && role0) ? val * (double)fv[u] : 0.0;
                }
            }
        } else {
            // FP32 rows: the beta entries [e0, em) load the node's D row and
            // the partner's f row, the others D rows only -- no branch and no
            // f row for the other 79% of the entries (1.34 -> 1.10 us per
            // cell against one pass loading every entry's f row)
            for (int e = e0; e < em; e += U) {
                uint4 x[U];
#pragma unroll
                for (int u = 0; u < U; u++) {  // clamped to the run's last entry: in bounds
                    const uint32_t eu = (uint32_t)min(e + u, em - 1);
                    x[u] = K == WAVE ? cent[eu] : ld_off(cent, eu);
                }
                T dv[U], fv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    dv[u] = ct_row<K>(Dt, x[u].x >> 2, k);
                    fv[u] = ct_row<K>(fr, x[u].y, k);
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool in = e + u < em;
                    const double val = __hiloint2double((int)x[u].w, (int)x[u].z);
                    q = ct_acc(q, in ? -1.0 : 0.0, dv[u]);
                    l += in ? val * (double)fv[u] : 0.0;
                }
            }
            // lambda, mu, lambda+s, mu-s entries [em, e1): D rows only
            constexpr int U2 = CT_UNROLL_D;
            for (int e = em; e < e1; e += U2) {
                uint4 x[U2];
#pragma unroll
                for (int u = 0; u < U2; u++) {  // clamped to the destination's last entry: in bounds
                    const uint32_t eu = (uint32_t)min(e + u, e1 - 1);
                    x[u] = K == WAVE ? cent[eu] : ld_off(cent, eu);
                }
                T dv[U2];
#pragma unroll
                for (int u = 0; u < U2; u++) dv[u] = ct_row<K>(Dt, x[u].x >> 2, k);
#pragma unroll
                for (int u = 0; u < U2; u++) {
                    const double val = __hiloint2double((int)x[u].w, (int)x[u].z);
                    q = ct_acc(q, e + u < e1 ? val : 0.0, dv[u]);
                }
            }
        }
        Q0[to + (size_t)d * K + k] = q;
        L[to + (size_t)d * K + k] = (double)ct_row<K>(fr, (uint32_t)d, k) * l;
    }
}

// ---------------------------------------------------------------------------
// moments, solve, apply
// ---------------------------------------------------------------------------
// Block (b, t): destinations [b npts / P, (b + 1) npts / P) of tile t; lane
// group g of K lanes takes destinations lo + g, + NPB, ...; lane k adds cell
// k's moments.  The lanes of one cell are reduced across the wave (lane bits
// above K) and then across the block's waves in a fixed order; partial b of
// cell (t K + k) goes to mp[((t K + k) 20 + m) P + b].
template <int K>
__global__ __launch_bounds__(CT_THREADS) void k_ct_moments(int npts, const double *__restrict__ xi,
                                                           const double *__restrict__ Q0,
                                                           const double *__restrict__ L, double *__restrict__ mp)
{
    constexpr int NPB = CT_THREADS / K, NW = CT_THREADS / WAVE;
    __shared__ double red[NW][K][21];
    const int k = threadIdx.x % K, g = threadIdx.x / K, t = blockIdx.y, P = gridDim.x;
    const int lo = (int)((long)blockIdx.x * npts / P), hi = (int)((long)(blockIdx.x + 1) * npts / P);
    const size_t to = (size_t)t * npts * K;
    double m[20];
#pragma unroll
    for (int i = 0; i < 20; i++) m[i] = 0.0;
    for (int d = lo + g; d < hi; d += NPB)
        moments_add(m, xi, d, Q0[to + (size_t)d * K + k], L[to + (size_t)d * K + k]);
#pragma unroll
    for (int sh = K; sh < WAVE; sh <<= 1)
#pragma unroll
        for (int i = 0; i < 20; i++) m[i] += __shfl_xor(m[i], sh, WAVE);
    const int wv = threadIdx.x / WAVE;
    if (threadIdx.x % WAVE < K)
#pragma unroll
        for (int i = 0; i < 20; i++) red[wv][k][i] = m[i];
    __syncthreads();
    for (int i = threadIdx.x; i < K * 20; i += CT_THREADS) {
        const int kk = i / 20, mm = i % 20;
        double sm = 0.0;
        for (int w = 0; w < NW; w++) sm += red[w][kk][mm];
        mp[(((size_t)t * K + kk) * 20 + mm) * P + blockIdx.x] = sm;
    }
}

// One wave per cell: lane i adds partials i, i + 64, ... of each moment, the
// wave sums them (fixed pairing), then the 5x5 solve; X[c][5].
__global__ __launch_bounds__(WAVE) void k_ct_solve(int cells, int P, const double *__restrict__ mp,
                                                   double *__restrict__ X)
{
    const int c = blockIdx.x;
    if (c >= cells) return;
    double s[20];
#pragma unroll
    for (int i = 0; i < 20; i++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < P; b += WAVE) a += mp[((size_t)c * 20 + i) * P + b];
        s[i] = a;
    }
    wave_sum_n<20>(s);
    double x[5];
    moments_solve(s, x);
    if (threadIdx.x < 5) {
        double v = x[0];
#pragma unroll
        for (int i = 1; i < 5; i++) v = threadIdx.x == i ? x[i] : v;
        X[(size_t)c * 5 + threadIdx.x] = v;
    }
}

// Destinations per apply block: its LDS holds two [K][DCH + 1] FP64 tiles.
__host__ __device__ constexpr int ct_apply_dests(int K) { return K <= 32 ? 64 : 32; }

// Block (x, t): destinations [DCH x, DCH x + DCH) of tile t.  Q0 / L rows go
// through LDS so that every output row (a cell's 64 destinations) is written
// contiguously (DCH destinations):  Q = Q0 - L s, s = X.psi (X = 0 uncorrected); loss = L;
// nu = (L / f)(1 + s) (0 where f <= 0); f += dt Q.
template <int K>
__global__ __launch_bounds__(CT_THREADS) void k_ct_apply(int npts, int cells, const double *__restrict__ xi,
                                                         const double *__restrict__ Q0,
                                                         const double *__restrict__ L,
                                                         const double *__restrict__ X, const double *f,
                                                         double *__restrict__ Qout, double *__restrict__ Lout,
                                                         double *__restrict__ nu_out, double dt, double *f_upd)
{
    constexpr int DCH = ct_apply_dests(K);
    __shared__ double sq[K][DCH + 1], sl[K][DCH + 1];
    const int d0 = blockIdx.x * DCH, t = blockIdx.y;
    const size_t to = (size_t)t * npts * K;
    for (int i = threadIdx.x; i < DCH * K; i += CT_THREADS) {
        const int dl = i / K, k = i % K, d = d0 + dl;
        if (d < npts) {
            sq[k][dl] = Q0[to + (size_t)d * K + k];
            sl[k][dl] = L[to + (size_t)d * K + k];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < DCH * K; i += CT_THREADS) {
        const int k = i / DCH, dl = i % DCH, d = d0 + dl;
        const long c = (long)t * K + k;
        if (c >= cells || d >= npts) continue;
        const double x = xi[3 * d], y = xi[3 * d + 1], z = xi[3 * d + 2];
        const double e = x * x + y * y + z * z;
        double s = 0.0;
        if (X) {
            const double *Xc = X + 5 * c;
            s = Xc[0] + Xc[1] * x + Xc[2] * y + Xc[3] * z + Xc[4] * e;
        }
        const double lv = sl[k][dl], q = sq[k][dl] - lv * s;
        const size_t o = (size_t)c * npts + d;
        if (Qout) Qout[o] = q;
        if (Lout) Lout[o] = lv;
        if (nu_out || dt != 0.0) {
            const double fv = f[o];
            if (nu_out) nu_out[o] = fv > 0.0 ? (lv / fv) * (1.0 + s) : 0.0;
            if (dt != 0.0) f_upd[o] = fv + dt * q;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int ct_cells_per_tile(const gdt_context *h, bool mixed)
{
    return h->ct_wide ? WAVE : (mixed ? CtK<float>::K : CtK<double>::K);
}

// Tiles of one node / gather launch: the configured chunk, at most the tiles
// there are, and at most what keeps a chunk's node terms within 32-bit byte
// offsets (the narrow tiles index them so).
static long ct_chunk(const gdt_context *h, long tiles, size_t sz, int K)
{
    const long fit = (long)(4294967295.0 / ((double)sz * (double)h->nn * K));
    return std::max(1L, std::min({tiles, (long)h->ct_chunk_tiles, fit}));
}

int ct_prepare(gdt_context *h, int cells, bool mixed)
{
    const int K = ct_cells_per_tile(h, mixed);
    const size_t sz = mixed ? sizeof(float) : sizeof(double);
    const long tiles = (cells + K - 1) / K;
    const long chunk = ct_chunk(h, tiles, sz, K);
    const size_t np = h->n_pnt;
    GDT_CHECK_GRID_Y(tiles, "cell tiles: tiles of cells");
    // one tile's node values are indexed with 32-bit byte offsets
    GDT_CHECK((double)sz * h->nn * K < 4294967296.0 && h->nn < (1L << 30), GDT_E_ARG,
              "cell tiles: %ld node values per tile exceed the 32-bit offsets", h->nn * K);
    GDT_CHECK(h->ct_f.ensure(sz * tiles * np * K) == 0 && h->ct_D.ensure(sz * chunk * h->nn * K) == 0 &&
                  h->ct_Q0.ensure(8 * tiles * np * K) == 0 &&
                  h->ct_L.ensure(8 * tiles * np * K) == 0 &&
                  h->ct_mp.ensure(8 * (size_t)tiles * K * 20 * CT_MOM_PARTS) == 0 &&
                  h->ct_X.ensure(8 * (size_t)tiles * K * 5) == 0,
              GDT_E_NOMEM, "cell tiles: scratch allocation failed (%ld tiles of %d cells)", tiles, K);
    return GDT_OK;
}

template <class T, int K>
static int ct_launch(gdt_context *h, const double *f, int cells, bool correct, double *Qout, double *Lout,
                     double *nu_out, double dt, double *f_upd)
{
    const int np = h->n_pnt, tiles = (cells + K - 1) / K;
    const int chunk = (int)ct_chunk(h, tiles, sizeof(T), K);
    T *ft = h->ct_f.as<T>(), *Dv = h->ct_D.as<T>();
    double *Q0 = h->ct_Q0.as<double>(), *L = h->ct_L.as<double>();
    GDT_LAUNCH_TIMED(h, "ct_pack", (k_ct_pack<T, K>), dim3((np + 63) / 64, tiles), dim3(CT_THREADS), 0u, np, cells,
                     f, ft);
    const int nblk = std::max(1, h->ct_node_blocks), gblk = std::max(1, h->ct_gather_blocks);
    CtXcd xo;
    for (int i = 0; i < 9; i++) xo.off[i] = h->kxoff[i];
    for (int t0 = 0; t0 < tiles; t0 += chunk) {
        const int nt = std::min(chunk, tiles - t0);
        const auto *nrw = (const typename CtRW<T>::type *)(sizeof(T) == 8 ? h->krw.p : h->krwf.p);
        GDT_LAUNCH_TIMED(h, "ct_node", (k_ct_node<T, K>), dim3(nblk, nt), dim3(CT_THREADS), 0u, np, (int)h->nn,
                         h->kaptr.as<int32_t>(), h->kaperm.as<int32_t>(), h->kt4.as<uint4>(), nrw,
                         h->krw.as<double2>(), (const T *)ft, t0, Dv, Q0, L, h->kmath.as<double>());
        GDT_LAUNCH_TIMED(h, "ct_gather", (k_ct_gather<T, K>), dim3(gblk, nt), dim3(CT_THREADS), 0u, np, (int)h->nn,
                         h->kcptr.as<int32_t>(), h->kcmid.as<int32_t>(),
                         (sizeof(T) == 8 ? h->kcent : h->kcent2).as<uint4>(), h->kdperm.as<int32_t>(),
                         h->kxperm.as<int32_t>(), xo, (const T *)Dv, (const T *)ft, t0, Q0, L);
    }
    double *X = nullptr;
    if (correct) {
        GDT_LAUNCH_TIMED(h, "ct_moments", k_ct_moments<K>, dim3(CT_MOM_PARTS, tiles), dim3(CT_THREADS), 0u, np,
                         h->xi.as<double>(), (const double *)Q0, (const double *)L, h->ct_mp.as<double>());
        GDT_LAUNCH_TIMED(h, "ct_solve", k_ct_solve, dim3(cells), dim3(WAVE), 0u, cells, CT_MOM_PARTS,
                         (const double *)h->ct_mp.as<double>(), h->ct_X.as<double>());
        X = h->ct_X.as<double>();
    }
    GDT_LAUNCH_TIMED(h, "ct_apply", k_ct_apply<K>, dim3((np + ct_apply_dests(K) - 1) / ct_apply_dests(K), tiles), dim3(CT_THREADS), 0u, np, cells,
                     h->xi.as<double>(), (const double *)Q0, (const double *)L, (const double *)X, f, Qout, Lout,
                     nu_out, dt, f_upd);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int ct_config(gdt_context *h, int chunk_tiles, int node_blocks, int gather_blocks, int wide)
{
    GDT_CHECK(chunk_tiles >= 1 && chunk_tiles <= 64 && node_blocks >= 1 && node_blocks <= 65535 &&
                  gather_blocks >= 1 && gather_blocks <= 65535 && (wide == 0 || wide == 1),
              GDT_E_ARG, "cell tiles: launch shape (%d, %d, %d, %d) out of range", chunk_tiles, node_blocks,
              gather_blocks, wide);
    h->ct_chunk_tiles = chunk_tiles;
    h->ct_node_blocks = node_blocks;
    h->ct_gather_blocks = gather_blocks;
    h->ct_wide = wide;
    return GDT_OK;
}

int ct_run(gdt_context *h, bool mixed, const double *f, int cells, bool correct, double *Qout, double *Lout,
           double *nu_out, double dt, double *f_upd)
{
    GDT_CHECK(h->kcent.p && h->kcptr.p && h->kdperm.p && h->krw.p && h->krwf.p && h->kt4.p && h->kaptr.p &&
                  h->kaperm.p && h->kxperm.p && h->kcmid.p && h->kcent2.p, GDT_E_STATE,
              "cell tiles: stencil not built");
    if (h->ct_wide)
        return mixed ? ct_launch<float, WAVE>(h, f, cells, correct, Qout, Lout, nu_out, dt, f_upd)
                     : ct_launch<double, WAVE>(h, f, cells, correct, Qout, Lout, nu_out, dt, f_upd);
    return mixed ? ct_launch<float, CtK<float>::K>(h, f, cells, correct, Qout, Lout, nu_out, dt, f_upd)
                 : ct_launch<double, CtK<double>::K>(h, f, cells, correct, Qout, Lout, nu_out, dt, f_upd);
}
