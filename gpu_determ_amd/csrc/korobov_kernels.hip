// Korobov / two-point projection collision operator on gfx950 (wave64, FP64).
//
// Per lattice node t of the stencil (pre-collision pair a, b; projection pairs
// (l, m) and (ls, ms); r = num/den; w = K sqrt(D2)) and cell c:
//   P = f_a f_b,  R = f_l f_m (f_ls f_ms / f_l f_m)^r  (linear if a product
//   is not positive),  D = w (P - R),
//   Q[a], Q[b] -= D;  Q[l], Q[m] += (1-r) D;  Q[ls], Q[ms] += r D;
//   loss[a], loss[b] += 2 w P.
//
// Default path, three kernels per step:
//   k_kor_node    node-parallel, f of the cell staged in LDS (16-byte LDS
//                 DMA): D and 2wP of every node, once, written contiguously
//                 (double2 per node), and the lambda value slots (prefix
//                 sums of the lambda terms over each chunk of 16 node
//                 positions; the records are sorted by lambda, so a run of
//                 one lambda within a chunk reads as two slots at most);
//                 block b evaluates a slice of node group b % KOR_GROUPS, on
//                 the XCD that gathers that group;
//   k_kor_gather  destination-sorted reduction, XCD-aware: the nodes are
//                 split into KOR_GROUPS contiguous ranges, each with its own
//                 destination-sorted CSR of (node | role, coefficient -1 /
//                 1-r / r) entries, where the value slots (coefficient 1)
//                 stand for the lambda entries.
//                 Workgroups are dealt to the XCDs round robin, so block b
//                 serves group b % KOR_GROUPS and every XCD gathers only its
//                 range's (D, 2wP) and slots -- under 1 MB, resident in
//                 that XCD's L2.  A block takes one tile of a group's
//                 destinations: it parks every entry's term in LDS, then a
//                 thread (or, for tiles of long rows, a quad) sums each
//                 destination's terms in a fixed order and writes the group's
//                 partial Q and loss; each block adds its destinations' moment
//                 sums (linear, so partials add up);
//   k_kor_apply   sums the group partials in group order, solves the 5x5
//                 correction system from the moment sums, applies it (and the
//                 explicit step).
// Alternative path (GDT_FLAG_FUSED): k_kor_fused streams one 20-byte record
// per entry of the global CSR and recomputes the node's D per entry from the
// LDS copy of f.  Per-entry terms use the same operations on both paths
// (contraction off), so they agree with each other and with the oracle to
// FP64 rounding.
#include <algorithm>
#include <vector>

#include "gdt_internal.h"
#include "kernel_common.h"
#include "kmath.h"

#define KOR_THREADS 1024
#define KOR_WAVES (KOR_THREADS / WAVE)
#define KOR_GROUPS KOR_NGROUPS

// D and 2 w P of one node from the LDS copy of f (log / exp tables at ts).
// x = (a | b << 15 | D2 >> 10 << 30, l | m << 15, ls | ms << 15,
//      |num| / 4 | |den| / 4 << 11 | (D2 & 1023) << 22): r = num / den is the
// quotient of the quarters (the same rational, so the same rounded double).
__device__ __forceinline__ int kor_num(uint32_t w) { return (int)(w & 0x7FFu); }
__device__ __forceinline__ int kor_den(uint32_t w) { return (int)((w >> 11) & 0x7FFu); }
__device__ __forceinline__ void node_terms(const uint4 x, const double *fs, const double *ts, double K, double &D,
                                           double &W2P, double &r)
{
#pragma clang fp contract(off)
    const int num = kor_num(x.w), den = kor_den(x.w);
    const uint32_t d2 = (x.w >> 22) | ((x.x >> 30) << 10);
    const double P = fs[x.x & 0x7FFFu] * fs[(x.x >> 15) & 0x7FFFu];
    const double Pl = fs[x.y & 0x7FFFu] * fs[x.y >> 15];
    double R;
    r = 0.0;
    if (num == 0) {
        R = Pl;
    } else {
        r = (double)num / (double)den;
        const double Ps = fs[x.z & 0x7FFFu] * fs[x.z >> 15];
        if (Pl > 0.0 && Ps > 0.0)
            R = Pl * kmt_exp(r * kmt_log(Ps / Pl, ts), ts);
        else
            R = (1.0 - r) * Pl + r * Ps;
    }
    const double w = K * sqrt((double)d2);
    D = w * (P - R);
    W2P = 2.0 * (w * P);
}

// Stages f of cell c into LDS fs[0, npts) and the log / exp tables at
// fs[kor_fpad(npts), + KMT_SIZE); the gap is padding that the staging may
// overwrite.  Returns with every staging write done (the caller's barrier
// then publishes them).
//   16-byte path (f 16-byte aligned, npts even): LDS DMA.  Wave w copies the
//   64-pair chunks w, w + KOR_WAVES, ... with global_load_lds_dwordx4, which
//   writes lane i's 16 bytes to the chunk's LDS base + 16 i without passing
//   through registers; lanes past the end re-read the last pair into the
//   padding (kor_fpad rounds up to whole chunks).  About 9 loads per thread
//   instead of 17 register round trips, and nothing waits between them.
//   Otherwise: 8-byte loads, every one of the thread's issued before the
//   first LDS store (npts <= KOR_STAGE * KOR_THREADS = 19,456; the stencil
//   build admits at most 19,200).
#define KOR_STAGE 19
static_assert(KMT_SIZE <= KOR_THREADS, "one table entry per thread");
typedef __attribute__((address_space(1))) void gmem_void;
typedef __attribute__((address_space(3))) void lds_void;
__device__ __forceinline__ void stage_f(double *fs, const double *fc, int npts, const double *kmt)
{
    const double tv = kmt[threadIdx.x < KMT_SIZE ? threadIdx.x : 0];
    if ((npts & 1) == 0 && ((uintptr_t)fc & 15) == 0) {
        const int np2 = npts >> 1, nch = (np2 + WAVE - 1) / WAVE;
        const int lane = threadIdx.x % WAVE;
        for (int c = threadIdx.x / WAVE; c < nch; c += KOR_WAVES) {
            const int t = c * WAVE + lane;
            __builtin_amdgcn_global_load_lds((gmem_void *)(fc + 2 * (t < np2 ? t : np2 - 1)),
                                             (lds_void *)(fs + 2 * WAVE * c), 16, 0, 0);
        }
    } else {
        double v[KOR_STAGE];
#pragma unroll
        for (int j = 0; j < KOR_STAGE; j++) {
            const int t = j * KOR_THREADS + threadIdx.x;
            v[j] = fc[t < npts ? t : npts - 1];  // unconditional (clamped) loads: no branches
        }
#pragma unroll
        for (int j = 0; j < KOR_STAGE; j++) {
            const int t = j * KOR_THREADS + threadIdx.x;
            if (t < npts) fs[t] = v[j];
        }
    }
    if (threadIdx.x < KMT_SIZE) fs[kor_fpad(npts) + threadIdx.x] = tv;
    __builtin_amdgcn_s_waitcnt(0);  // LDS DMA writes complete under vmcnt
}

// ---------------------------------------------------------------------------
// node values
// ---------------------------------------------------------------------------
// Per node t of cell c, the values the destination entries read:
//   V[t] = (D, 2 w P)
// and, after the nn nodes, the lambda value slots (below).
// Block b of a grid of B blocks per cell owns a contiguous node range:
//   B a multiple of KOR_GROUPS (xcd): the range is slice b / G of the S = B / G
//     equal slices of gather group g = b % G.  Workgroups are dealt to the
//     XCDs round robin, so the values of group g are written from the XCD
//     whose L2 the gather of group g then reads them through (one step:
//     29.0 -> 26.7 us against group-blind ranges);
//   otherwise: [b per, (b + 1) per), per = ceil(nn / B).
// Slices are whole waves of positions (per rounded up to 64) and groups
// start on chunk boundaries, so every block has the same share within a
// wave, and a wave's 64 positions start on a chunk boundary.  A thread owns
// nodes t and t + KOR_THREADS of the range and evaluates both together, so
// the two dependency chains (LDS reads, log / exp) overlap; their records
// are loaded before the staging of f, so that latency overlaps too.  Further
// nodes (more than two per thread) are handled one at a time.  f of the cell
// is fc; values go to out (the cell's V).
//
// Lambda pieces.  The records are sorted by the lambda destination, so a
// destination's lambda terms (1 - r) D come from one run of positions.  The
// positions are cut into chunks of KOR_CHUNK = 16, one DPP row of a wave
// (a wave holds 64 consecutive positions from a chunk boundary on), and a
// piece is a run of one lambda within a chunk.  The row forms the prefix sums I of its terms
// (row shifts 1, 2, 4, 8: a fixed pairing order) and a piece [s, e] is
// I_e - I_(s-1): its last position writes +I_e and, unless it starts the
// chunk, its first position writes -I_(s-1), each to its own value slot
// out[nn + k], k = cbase[c] + the slots of the positions before it (a first
// position's slot before a last position's).  Chunks are absolute
// positions, so the values do not depend on the launch shape (the grouped,
// multi-cell and batch grids give the same bits).  The gather reads each
// slot as one entry (role 1 with r = 0: coefficient 1) in place of the
// piece's lambda entries: about 58,000 slots for 403,000 lambda entries at
// 32^3 and 2^21 nodes.  (A segmented scan per piece costs twice the
// instructions of the prefix sums, and the node kernel's arithmetic is
// issue-bound.)
typedef unsigned int u32x4_nt __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 load_nt_u4(const uint4 *p)
{
    const u32x4_nt v = __builtin_nontemporal_load((const u32x4_nt *)p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

#define DPP_ROW_SHL(k) (0x100 + (k))  // lane i of a row reads lane i + k of the row
static_assert(KOR_CHUNK == 16, "a chunk of positions is one DPP row of 16 lanes");
// u: (1 - r) D of the lane's position (0 where not valid); lam its lambda;
// cb: the first value slot of the wave's first chunk (loaded by the caller
// ahead of time: a load here would hold up the end of the wave).
__device__ __forceinline__ void lambda_pieces(double u, uint32_t lam, bool valid, int cb, double2 *__restrict__ P)
{
#pragma clang fp contract(off)
    const int lane = (int)(threadIdx.x % WAVE);
    const int key = valid ? (int)lam : -2 - lane;
    double I = u;
    I += dpp_f64_or<DPP_ROW_SHR(1), 0xF>(I, 0.0);
    I += dpp_f64_or<DPP_ROW_SHR(2), 0xF>(I, 0.0);
    I += dpp_f64_or<DPP_ROW_SHR(4), 0xF>(I, 0.0);
    I += dpp_f64_or<DPP_ROW_SHR(8), 0xF>(I, 0.0);
    const double E = dpp_f64_or<DPP_ROW_SHR(1), 0xF>(I, 0.0);  // I of the previous position
    const int kprev = dpp_i32_or<DPP_ROW_SHR(1), 0xF>(key, -1);
    const int knext = dpp_i32_or<DPP_ROW_SHL(1), 0xF>(key, -1);
    const bool head = valid && kprev != key && (lane & (KOR_CHUNK - 1)) != 0, tail = valid && knext != key;
    const unsigned long long mh = __ballot(head), mt = __ballot(tail);
    if (head || tail) {
        const unsigned long long below = (1ull << lane) - 1ull;
        const int k = cb + __popcll(mh & below) + __popcll(mt & below);
        if (head) {
            __builtin_nontemporal_store(-E, &P[k].x);
            __builtin_nontemporal_store(0.0, &P[k].y);
        }
        if (tail) {
            __builtin_nontemporal_store(I, &P[k + (head ? 1 : 0)].x);
            __builtin_nontemporal_store(0.0, &P[k + (head ? 1 : 0)].y);
        }
    }
}

__device__ __forceinline__ void node_block(double *fs, int npts, long nn, const uint4 *__restrict__ nrec, double K,
                                           const double *__restrict__ fc, double2 *__restrict__ out, int b, int B,
                                           const double *__restrict__ kmt, bool xcd,
                                           const int32_t *__restrict__ cbase)
{
#pragma clang fp contract(off)
    long lo = 0, hi = nn, k = b, S = B;
    if (xcd) {
        const int g = b % KOR_GROUPS;
        lo = kor_group_begin(g, nn);
        hi = kor_group_begin(g + 1, nn);
        k = b / KOR_GROUPS;
        S = B / KOR_GROUPS;
    }
    const long per = ((hi - lo + S - 1) / S + WAVE - 1) / WAVE * WAVE, t0 = lo + k * per;
    const long tend = t0 + per < hi ? t0 + per : hi;
    const long ta = t0 + threadIdx.x, tb = ta + KOR_THREADS;
    const long lane = threadIdx.x % WAVE;
    uint4 xa = make_uint4(0u, 0u, 0u, 0u), xb = xa;
    // streaming (nontemporal) loads: records are read once per step, and
    // should not push the gather's entries out of L2
    if (ta < tend) xa = load_nt_u4(&nrec[ta]);
    if (tb < tend) xb = load_nt_u4(&nrec[tb]);
    // the waves' first value slots, in flight with the records
    const int cba = ta - lane < tend ? cbase[(ta - lane) / KOR_CHUNK] : 0;
    const int cbb = tb - lane < tend ? cbase[(tb - lane) / KOR_CHUNK] : 0;
    stage_f(fs, fc, npts, kmt);
    __syncthreads();
    const double *ts = fs + kor_fpad(npts);
    double Da, Wa, ra, Db, Wb, rb;
    node_terms(xa, fs, ts, K, Da, Wa, ra);
    node_terms(xb, fs, ts, K, Db, Wb, rb);
    // streaming (nontemporal) stores: the values are read by the next kernel
    // only, from other XCDs, so they need not linger in this XCD's L2
    if (ta < tend) {
        __builtin_nontemporal_store(Da, &out[ta].x);
        __builtin_nontemporal_store(Wa, &out[ta].y);
    }
    if (tb < tend) {
        __builtin_nontemporal_store(Db, &out[tb].x);
        __builtin_nontemporal_store(Wb, &out[tb].y);
    }
    double2 *P = out + nn;
    lambda_pieces(ta < tend ? (1.0 - ra) * Da : 0.0, xa.y & 0x7FFFu, ta < tend, cba, P);
    lambda_pieces(tb < tend ? (1.0 - rb) * Db : 0.0, xb.y & 0x7FFFu, tb < tend, cbb, P);
    // further nodes while the wave's first position is before tend (a
    // uniform condition: every lane of the wave takes part in the sums)
    for (long t = tb + KOR_THREADS; t - lane < tend; t += KOR_THREADS) {
        const bool v = t < tend;
        const uint4 x = v ? nrec[t] : make_uint4(0u, 0u, 0u, 0u);
        const int cb = cbase[(t - lane) / KOR_CHUNK];
        double D, W2P, r;
        node_terms(x, fs, ts, K, D, W2P, r);
        if (v) {
            __builtin_nontemporal_store(D, &out[t].x);
            __builtin_nontemporal_store(W2P, &out[t].y);
        }
        lambda_pieces(v ? (1.0 - r) * D : 0.0, x.y & 0x7FFFu, v, cb, P);
    }
}

__global__ __launch_bounds__(KOR_THREADS) void k_kor_node(int npts, long nn, long nv, const uint4 *__restrict__ nrec,
                                                           double K, const double *__restrict__ f,
                                                           double2 *__restrict__ V, const double *__restrict__ kmt,
                                                           const int32_t *__restrict__ cbase)
{
    extern __shared__ double fs[];
    const int c = blockIdx.y;
    node_block(fs, npts, nn, nrec, K, f + (size_t)c * npts, V + (size_t)c * nv, blockIdx.x, gridDim.x, kmt,
               gridDim.x % KOR_GROUPS == 0, cbase);
}

// ---------------------------------------------------------------------------
// block helpers
// ---------------------------------------------------------------------------
// Block sum of the 20 moment sums (DPP row sums, then the rows in a fixed
// order), written to mp[moment * nparts + b] (mp: the cell's moment sums).
template <int NT>
__device__ __forceinline__ void block_moments(double (&m)[20], double (*rpart)[20], double *mp, int b, int nparts)
{
#pragma unroll
    for (int i = 0; i < 20; i++) m[i] += dpp_f64<DPP_QUAD_XOR1>(m[i]);
#pragma unroll
    for (int i = 0; i < 20; i++) m[i] += dpp_f64<DPP_QUAD_XOR2>(m[i]);
#pragma unroll
    for (int i = 0; i < 20; i++) m[i] += dpp_f64<DPP_ROW_HALF_MIRROR>(m[i]);
#pragma unroll
    for (int i = 0; i < 20; i++) m[i] += dpp_f64<DPP_ROW_MIRROR>(m[i]);
    if ((threadIdx.x & 15) == 0)
#pragma unroll
        for (int i = 0; i < 20; i++) rpart[threadIdx.x / 16][i] = m[i];
    __syncthreads();
    if (threadIdx.x < 20) {
        double sm = 0.0;
        for (int r = 0; r < NT / 16; r++) sm += rpart[r][threadIdx.x];
        mp[(size_t)threadIdx.x * nparts + b] = sm;
    }
}

// Destinations of [ga, gb) without entries get exact zeros.
__device__ __forceinline__ void zero_empty(int ga, int gb, const int32_t *ptr, double *Qc, double *Lc, int lane)
{
    for (int g = ga + lane; g < gb; g += WAVE)
        if (ptr[g] == ptr[g + 1]) {
            Qc[g] = 0.0;
            if (Lc) Lc[g] = 0.0;
        }
}

// ---------------------------------------------------------------------------
// destination-sorted gather in tiles, one node group per XCD
// ---------------------------------------------------------------------------
// Entry record: (node << 2 | role, the node record's last word: |num| / 4,
// |den| / 4, D2 bits).  An entry's
// term is, from the node's (D, 2 w P) and r = num / den (the node kernel's
// operations, so the terms are bit-identical to the oracle's):
//   role 0 (a, b):   q = -D,  l = 2 w P;   role 1 (l, m): q = (1 - r) D;
//   role 2 (ls, ms): q = r D.
// Tile (d0, d1, e0, e1) of group g: destinations [d0, d1) of the group's CSR,
// whose entries are [e0, e1).  Block b serves group b % KOR_GROUPS (blocks are
// dealt to the XCDs round robin, so a group's node values stay in one L2).
// Per pass over at most `cap` entries:
//   phase 1  every thread issues the entry loads of its batch, then the node
//            value loads, and parks (q, l) of each entry in LDS by entry
//            position -- no dependent chains beyond entry -> value;
//   phase 2  thread i owns destinations d0 + i + j * TILE_THREADS and adds
//            their entries of the pass in entry order (fixed order:
//            reproducible sums).
// The destinations' row pointers and coordinates are loaded before phase 1.
// By default 64 tiles per group: at 32^3 and 2^21 nodes about 4,600 entries
// per tile (74 KB of LDS, two batches of TILE_BATCH entries per thread; 112
// VGPRs), so the G * 64 blocks of a cell run in one round, two per CU.  (128
// smaller tiles, four per CU, gather a little faster but give the apply twice
// the moment partials: the step is 0.3 us slower.)
#define TILE_THREADS 256
#define TILE_WAVES (TILE_THREADS / WAVE)
#define TILE_DPT (KOR_TILE_DST / TILE_THREADS)  // destinations per thread
#define TILE_BATCH 10

// Block b (group b % G, tile b / G) of a cell: Vc its node values, Lp / Qp its
// group partials [G][npts], mp its moment sums [20][nparts].
template <int BATCH = TILE_BATCH>
__device__ __forceinline__ void gather_block(double2 *ql, double (*rpart)[20], int npts, int ntile, int cap,
                                             const int4 *__restrict__ tiles, const int32_t *__restrict__ gptr,
                                             const uint2 *__restrict__ ent, const double2 *__restrict__ Vc,
                                             const double *__restrict__ xi, double *__restrict__ Lp,
                                             double *__restrict__ Qp, double *__restrict__ mp, int b, int nparts)
{
#pragma clang fp contract(off)
    const int g = b % KOR_GROUPS, k = b / KOR_GROUPS;
    const int4 tl = tiles[(size_t)g * ntile + k];
    const int d0 = tl.x, d1 = tl.y, e0 = tl.z, e1 = tl.w;
    const int32_t *ptr = gptr + (size_t)g * (npts + 1);
    // a tile of at most TILE_THREADS / 4 destinations (long rows) gives each
    // destination a quad: lane r of the quad sums the entries r, r + 4, ...
    // of the destination, the quad adds (s0 + s1) + (s2 + s3) -- the very
    // sums of the one-thread form below, so both give the same bits
    const bool quad = d1 - d0 <= TILE_THREADS / 4;
    const bool owner = !quad || (threadIdx.x & 3) == 0;
    int ea[TILE_DPT], eb[TILE_DPT];
    double px[TILE_DPT], py[TILE_DPT], pz[TILE_DPT], q[TILE_DPT], l[TILE_DPT];
#pragma unroll
    for (int i = 0; i < TILE_DPT; i++) {
        const int d = quad ? (i == 0 ? d0 + (int)(threadIdx.x >> 2) : d1) : d0 + threadIdx.x + i * TILE_THREADS;
        ea[i] = eb[i] = 0;
        px[i] = py[i] = pz[i] = q[i] = l[i] = 0.0;
        if (d < d1) {
            ea[i] = ptr[d];
            eb[i] = ptr[d + 1];
            px[i] = xi[3 * d];
            py[i] = xi[3 * d + 1];
            pz[i] = xi[3 * d + 2];
        }
    }
    for (int w0 = e0; w0 < e1; w0 += cap) {
        const int w1 = min(e1, w0 + cap);
        if (w0 > e0) __syncthreads();  // the previous pass's reads are done
        for (int j0 = w0 + threadIdx.x; j0 < w1; j0 += BATCH * TILE_THREADS) {
            uint2 x[BATCH];
#pragma unroll
            for (int j = 0; j < BATCH; j++) {
                const int e = j0 + j * TILE_THREADS;
                x[j] = e < w1 ? ent[e] : make_uint2(0u, 0u);
            }
            double2 v[BATCH];
#pragma unroll
            for (int j = 0; j < BATCH; j++) {
                const int e = j0 + j * TILE_THREADS;
                v[j] = e < w1 ? Vc[x[j].x >> 2] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int j = 0; j < BATCH; j++) {
                const int e = j0 + j * TILE_THREADS;
                if (e < w1) {
                    const uint32_t role = x[j].x & 3u;
                    const int num = kor_num(x[j].y), den = kor_den(x[j].y);
                    const double r = num == 0 ? 0.0 : (double)num / (double)den;
                    ql[e - w0] = role == 0u ? make_double2(-1.0 * v[j].x, v[j].y)
                                            : make_double2((role == 1u ? (1.0 - r) : r) * v[j].x, 0.0);
                }
            }
        }
        __syncthreads();
        if (quad) {
            const int a = max(ea[0], w0), bb = min(eb[0], w1);
            double qs = 0.0, ls = 0.0;
            for (int e = a + (int)(threadIdx.x & 3); e < bb; e += 4) {
                const double2 t = ql[e - w0];
                qs += t.x;
                ls += t.y;
            }
            qs += dpp_f64<DPP_QUAD_XOR1>(qs);
            ls += dpp_f64<DPP_QUAD_XOR1>(ls);
            qs += dpp_f64<DPP_QUAD_XOR2>(qs);
            ls += dpp_f64<DPP_QUAD_XOR2>(ls);
            q[0] += qs;
            l[0] += ls;
            continue;
        }
        // four interleaved partial sums per destination (loads in flight
        // together), combined in a fixed order: reproducible
#pragma unroll
        for (int i = 0; i < TILE_DPT; i++) {
            const int a = max(ea[i], w0), bb = min(eb[i], w1);
            double q4[4] = {0.0, 0.0, 0.0, 0.0}, l4[4] = {0.0, 0.0, 0.0, 0.0};
            int e = a;
            for (; e + 4 <= bb; e += 4) {
                double2 t[4];
#pragma unroll
                for (int r = 0; r < 4; r++) t[r] = ql[e + r - w0];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    q4[r] += t[r].x;
                    l4[r] += t[r].y;
                }
            }
            for (int r = 0; e < bb; e++, r++) {
                const double2 t = ql[e - w0];
                q4[r] += t.x;
                l4[r] += t.y;
            }
            q[i] += (q4[0] + q4[1]) + (q4[2] + q4[3]);
            l[i] += (l4[0] + l4[1]) + (l4[2] + l4[3]);
        }
    }
    double m[20];
#pragma unroll
    for (int i = 0; i < 20; i++) m[i] = 0.0;
#pragma unroll
    for (int i = 0; i < TILE_DPT; i++) {
        const int d = quad ? (i == 0 ? d0 + (int)(threadIdx.x >> 2) : d1) : d0 + threadIdx.x + i * TILE_THREADS;
        if (d < d1 && owner) {
            const size_t po = (size_t)g * npts + d;
            __builtin_nontemporal_store(q[i], &Qp[po]);
            __builtin_nontemporal_store(l[i], &Lp[po]);
            moments_add_psi(m, px[i], py[i], pz[i], q[i], l[i]);
        }
    }
    block_moments<TILE_THREADS>(m, rpart, mp, b, nparts);
}

__global__ __launch_bounds__(TILE_THREADS) void k_kor_gather(int npts, long nv, int ntile, int cap,
                                                              const int4 *__restrict__ tiles,
                                                              const int32_t *__restrict__ gptr,
                                                              const uint2 *__restrict__ ent,
                                                              const double2 *__restrict__ V,
                                                              const double *__restrict__ xi,
                                                              double *__restrict__ Lpart, double *__restrict__ Qpart,
                                                              double *__restrict__ mpart)
{
    extern __shared__ double2 ql[];
    __shared__ double rpart[TILE_THREADS / 16][20];
    const int c = blockIdx.y;
    const size_t po = (size_t)c * KOR_GROUPS * npts;
    gather_block(ql, rpart, npts, ntile, cap, tiles, gptr, ent, V + (size_t)c * nv, xi, Lpart + po, Qpart + po,
                 mpart + (size_t)c * 20 * gridDim.x, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// fused: per-entry recomputation from the LDS copy of f
// ---------------------------------------------------------------------------
// One 20-byte record per entry of the global destination-sorted CSR: the
// node's record (six indices, num / 4, den / 4, D2) and D2 | role << 12 |
// destination << 14.  Wave w of the grid owns the destinations
// [wbound[w], wbound[w+1]) (whole destinations, balanced entry counts) and
// walks their entries in chunks of 64: every lane evaluates its entry's term
// from f in LDS, a DPP segmented scan sums the chunk's runs of equal
// destination, and the run tails write Q and loss (a run that reaches the
// chunk's end and continues in the next chunk is carried).  Records are
// loaded two chunks ahead; the first two are in flight during the staging.
__global__ __launch_bounds__(KOR_THREADS) void k_kor_fused(int npts, const int32_t *__restrict__ wbound,
                                                            const int32_t *__restrict__ ptr,
                                                            const uint4 *__restrict__ e4,
                                                            const uint32_t *__restrict__ e1, double K,
                                                            const double *__restrict__ f,
                                                            const double *__restrict__ xi,
                                                            double *__restrict__ loss, double *__restrict__ Q,
                                                            double *__restrict__ mpart, const double *__restrict__ kmt)
{
#pragma clang fp contract(off)
    extern __shared__ double fs[];
    __shared__ double rpart[KOR_THREADS / 16][20];
    const int c = blockIdx.y;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int wg = blockIdx.x * KOR_WAVES + wave;
    const int ga = wbound[wg], gb = wbound[wg + 1];
    const int ea = ptr[ga], eb = ptr[gb];
    uint4 xa = make_uint4(0u, 0u, 0u, 0u), xb = xa;
    uint32_t ya = 0u, yb = 0u;
    if (ea + lane < eb) {
        xa = e4[ea + lane];
        ya = e1[ea + lane];
    }
    if (ea + WAVE + lane < eb) {
        xb = e4[ea + WAVE + lane];
        yb = e1[ea + WAVE + lane];
    }
    stage_f(fs, f + (size_t)c * npts, npts, kmt);
    __syncthreads();
    const double *ts = fs + kor_fpad(npts);
    double *Qc = Q + (size_t)c * npts;
    double *Lc = loss + (size_t)c * npts;
    double m[20];
#pragma unroll
    for (int i = 0; i < 20; i++) m[i] = 0.0;
    zero_empty(ga, gb, ptr, Qc, Lc, lane);

    int ckey = -1;  // open run carried from the previous chunk
    double cq = 0.0, cl = 0.0;
    for (int base = ea; base < eb; base += WAVE) {
        const uint4 x = xa;
        const uint32_t y = ya;
        xa = xb;
        ya = yb;
        const int e2 = base + 2 * WAVE + lane;
        if (e2 < eb) {
            xb = e4[e2];
            yb = e1[e2];
        }
        const bool valid = base + lane < eb;
        double q = 0.0, l = 0.0;
        int key = -2 - lane;
        if (valid) {
            double D, W2P, r;
            node_terms(x, fs, ts, K, D, W2P, r);
            const uint32_t role = (y >> 12) & 3u;
            if (role == 0u) {
                q = -1.0 * D;
                l = W2P;
            } else {
                q = (role == 1u ? (1.0 - r) : r) * D;
            }
            key = (int)(y >> 14);
        }
        seg_scan_dpp(q, l, key);
        if (key == ckey) {
            q += cq;
            l += cl;
        }
        // a run ends where the next lane's key differs; lane 63 compares with
        // the first entry of the next chunk
        const bool more = base + WAVE < eb;
        const int next_first = more ? (int)(__builtin_amdgcn_readfirstlane(ya) >> 14) : -1;
        int knext = dpp_i32_or<DPP_WAVE_SHL1, 0xF>(key, -1);
        if (lane == WAVE - 1) knext = next_first;
        if (valid && knext != key) {
            Qc[key] = q;
            Lc[key] = l;
            moments_add(m, xi, key, q, l);
        }
        if (more) {
            const int k63 = __builtin_amdgcn_readlane(key, WAVE - 1);
            ckey = k63 == next_first ? k63 : -1;
            cq = readlane_f64(q, WAVE - 1);
            cl = readlane_f64(l, WAVE - 1);
        }
    }
    block_moments<KOR_THREADS>(m, rpart, mpart + (size_t)c * 20 * gridDim.x, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// group sum, correction from the moment sums (+ explicit step)
// ---------------------------------------------------------------------------
// Every block reduces its cell's nparts x 20 moment sums in the same order
// and solves the same 5x5 system; then each thread sums its node's G group
// partials in group order and writes
//   Q = Q0 - L X.psi,  loss = L,  nu = (L/f)(1 + X.psi),  f <- f + dt Q.
// All global loads of a thread (its node's partials, f and centre, and its
// share of the moment sums) are issued together at the start, so the block
// waits for memory once.
#define APPLY_ROUNDS 2  // pairs of block sums per thread and moment in flight: 4 APPLY_THREADS partials a round
#define APPLY_THREADS KOR_APPLY_THREADS
// Block b of a cell (destinations b * APPLY_THREADS + tid).  Pointers are the cell's:
// mp [20][nparts], fc [npts], Lp / Qp [G][npts], outputs [npts], X [5].
template <bool CORR, int G>
__device__ __forceinline__ void apply_block(double (*red)[20], double *Xs, int npts, int nparts,
                                            const double *__restrict__ mp, const double *__restrict__ xi,
                                            const double *fc, const double *__restrict__ Lp,
                                            const double *__restrict__ Qp, double *__restrict__ Qout,
                                            double *__restrict__ Lout, double *__restrict__ nu_out,
                                            double *__restrict__ Xout, double dt, double *f_upd, int b)
{
    const int v = b * APPLY_THREADS + threadIdx.x;
    const int vc = v < npts ? v : npts - 1;  // clamped: the loads are unconditional
    const double x = xi[3 * vc], y = xi[3 * vc + 1], z = xi[3 * vc + 2], fv = fc[vc];
    double qg[G], lg[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        qg[g] = __builtin_nontemporal_load(&Qp[(size_t)g * npts + vc]);
        lg[g] = __builtin_nontemporal_load(&Lp[(size_t)g * npts + vc]);
    }
    if (CORR) {
        // moment partials: thread t adds the partial pairs t, t + APPLY_THREADS,
        // ... of every moment (coalesced, all loads in flight), then DPP row
        // sums; wave 0 adds the APPLY_THREADS / 16 row sums in a fixed order
        // (lane a: moment a) and solves (every block gets the same X)
        double m[20];
#pragma unroll
        for (int a = 0; a < 20; a++) m[a] = 0.0;
        if ((nparts & 1) == 0) {
            // pairs of partials per 16-byte load (even nparts: every moment
            // row starts 16-byte aligned and holds whole pairs)
            for (int p0 = 0; p0 < nparts; p0 += 2 * APPLY_ROUNDS * APPLY_THREADS) {
                double2 vv[APPLY_ROUNDS][20];
#pragma unroll
                for (int k = 0; k < APPLY_ROUNDS; k++) {
                    const int p = p0 + 2 * (k * APPLY_THREADS + threadIdx.x);
#pragma unroll
                    for (int a = 0; a < 20; a++)
                        vv[k][a] = p < nparts ? *(const double2 *)(mp + (size_t)a * nparts + p)
                                              : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int k = 0; k < APPLY_ROUNDS; k++)
#pragma unroll
                    for (int a = 0; a < 20; a++) m[a] += vv[k][a].x + vv[k][a].y;
            }
        } else {
            for (int p = threadIdx.x; p < nparts; p += APPLY_THREADS)
#pragma unroll
                for (int a = 0; a < 20; a++) m[a] += mp[(size_t)a * nparts + p];
        }
#pragma unroll
        for (int a = 0; a < 20; a++) m[a] += dpp_f64<DPP_QUAD_XOR1>(m[a]);
#pragma unroll
        for (int a = 0; a < 20; a++) m[a] += dpp_f64<DPP_QUAD_XOR2>(m[a]);
#pragma unroll
        for (int a = 0; a < 20; a++) m[a] += dpp_f64<DPP_ROW_HALF_MIRROR>(m[a]);
#pragma unroll
        for (int a = 0; a < 20; a++) m[a] += dpp_f64<DPP_ROW_MIRROR>(m[a]);
        if ((threadIdx.x & 15) == 0)
#pragma unroll
            for (int a = 0; a < 20; a++) red[threadIdx.x / 16][a] = m[a];
        __syncthreads();
        if (threadIdx.x < WAVE) {
            const int a = threadIdx.x < 20 ? threadIdx.x : 0;
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < APPLY_THREADS / 16; r++) t += red[r][a];
            double sm[20], X[5];
#pragma unroll
            for (int k = 0; k < 20; k++) sm[k] = readlane_f64(t, k);
            moments_solve(sm, X);
            if (threadIdx.x == 0)
#pragma unroll
                for (int k = 0; k < 5; k++) Xs[k] = X[k];
        }
    } else if (threadIdx.x < 5) {
        Xs[threadIdx.x] = 0.0;
    }
    __syncthreads();
    if (Xout && b == 0 && threadIdx.x < 5) Xout[threadIdx.x] = Xs[threadIdx.x];
    if (v < npts) {  // no early return: work-list callers may run further items after this one
        double q0 = 0.0, L = 0.0;
#pragma unroll
        for (int g = 0; g < G; g++) {
            q0 += qg[g];
            L += lg[g];
        }
        const double e = x * x + y * y + z * z;
        const double s = Xs[0] + Xs[1] * x + Xs[2] * y + Xs[3] * z + Xs[4] * e;
        const double q = q0 - L * s;
        if (Qout) Qout[v] = q;
        if (Lout) Lout[v] = L;
        if (nu_out) nu_out[v] = fv > 0.0 ? (L / fv) * (1.0 + s) : 0.0;
        if (dt != 0.0) f_upd[v] = fv + dt * q;
    }
}

template <bool CORR, int G>
__global__ __launch_bounds__(APPLY_THREADS) void k_kor_apply(int npts, int nparts, const double *__restrict__ mpart,
                                                             const double *__restrict__ xi, const double *f,
                                                             const double *__restrict__ Lpart,
                                                             const double *__restrict__ Qpart,
                                                             double *__restrict__ Qout, double *__restrict__ Lout,
                                                             double *__restrict__ nu_out, double *__restrict__ Xout,
                                                             double dt, double *f_upd)
{
    __shared__ double red[APPLY_THREADS / 16][20];
    __shared__ double Xs[5];
    const int c = blockIdx.y;
    const size_t o = (size_t)c * npts, po = (size_t)c * G * npts;
    apply_block<CORR, G>(red, Xs, npts, nparts, mpart + (size_t)c * 20 * nparts, xi, f + o, Lpart + po, Qpart + po,
                         Qout ? Qout + o : nullptr, Lout ? Lout + o : nullptr, nu_out ? nu_out + o : nullptr,
                         Xout ? Xout + 5 * c : nullptr, dt, f_upd ? f_upd + o : nullptr, blockIdx.x);
}

// ---------------------------------------------------------------------------
// uneven batches: work-list kernels
// ---------------------------------------------------------------------------
// Each kernel walks a static list of (class, cell, block, slot) items; block p
// takes items p, p + gridDim.x, ... (the launches give every item its own
// block; the loop keeps any smaller grid correct).  Every item runs exactly the body one
// block of the single-cell kernel runs, so per-cell results are bitwise those
// of a non-batched run.  Gather lists keep item i in group i % G, and the
// grid is a multiple of G, so each group stays on one XCD.  No atomics and no
// waits between blocks: every block finishes its own items and exits.
__global__ __launch_bounds__(KOR_THREADS) void k_batch_node(const KorClassDev *__restrict__ cls,
                                                             const int4 *__restrict__ items, int nitems)
{
    extern __shared__ double fs[];
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        if (i != (int)blockIdx.x) __syncthreads();  // the previous item's LDS reads are done
        const int4 it = items[i];
        const KorClassDev C = cls[it.x];
        node_block(fs, C.npts, C.nn, C.nrec, C.K, C.f + (size_t)it.y * C.npts, C.V + (size_t)it.w * C.nv,
                   it.z, C.node_blocks, C.kmt, false, C.cbase);
    }
}

__global__ __launch_bounds__(TILE_THREADS) void k_batch_gather(const KorClassDev *__restrict__ cls,
                                                                const int4 *__restrict__ items, int nitems)
{
    extern __shared__ double2 ql[];
    __shared__ double rpart[TILE_THREADS / 16][20];
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        if (i != (int)blockIdx.x) __syncthreads();
        const int4 it = items[i];
        const KorClassDev C = cls[it.x];
        const size_t po = (size_t)it.w * KOR_GROUPS * C.npts;
        gather_block(ql, rpart, C.npts, C.ntile, C.cap, C.tiles, C.gptr, C.ent, C.V + (size_t)it.w * C.nv, C.xi,
                     C.Lp + po, C.Qp + po, C.mp + (size_t)it.w * 20 * C.nparts, it.z, C.nparts);
    }
}

template <bool CORR>
__global__ __launch_bounds__(APPLY_THREADS) void k_batch_apply(const KorClassDev *__restrict__ cls,
                                                                const int4 *__restrict__ items, int nitems, double dt)
{
    __shared__ double red[APPLY_THREADS / 16][20];
    __shared__ double Xs[5];
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        if (i != (int)blockIdx.x) __syncthreads();
        const int4 it = items[i];
        const KorClassDev C = cls[it.x];
        const size_t o = (size_t)it.y * C.npts, po = (size_t)it.w * KOR_GROUPS * C.npts;
        apply_block<CORR, KOR_GROUPS>(red, Xs, C.npts, C.nparts, C.mp + (size_t)it.w * 20 * C.nparts, C.xi, C.f + o,
                                      C.Lp + po, C.Qp + po, C.Qout ? C.Qout + o : nullptr, nullptr, nullptr, nullptr,
                                      dt, dt != 0.0 ? C.f + o : nullptr, it.z);
    }
}

// The work-list kernels may take the largest dynamic LDS any class can need
// (f of KOR_NPTS_MAX nodes and the tables; KOR_TILE_CAP parked terms): the
// attribute is a per-kernel limit shared by every batch of the process, and
// each launch passes its own batch's size.
int batch_set_lds()
{
    const int lds_node = (int)sizeof(double) * (kor_fpad(KOR_NPTS_MAX) + KMT_SIZE);
    const int lds_gather = (int)sizeof(double2) * KOR_TILE_CAP;
    GDT_HIP(hipFuncSetAttribute((const void *)k_batch_node, hipFuncAttributeMaxDynamicSharedMemorySize, lds_node));
    GDT_HIP(hipFuncSetAttribute((const void *)k_batch_gather, hipFuncAttributeMaxDynamicSharedMemorySize, lds_gather));
    return GDT_OK;
}

int launch_batch_node(hipStream_t s, const KorClassDev *cls, const int4 *items, int nitems, int blocks, int lds)
{
    hipLaunchKernelGGL(k_batch_node, dim3(blocks), dim3(KOR_THREADS), lds, s, cls, items, nitems);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_batch_gather(hipStream_t s, const KorClassDev *cls, const int4 *items, int nitems, int blocks, int lds)
{
    hipLaunchKernelGGL(k_batch_gather, dim3(blocks), dim3(TILE_THREADS), lds, s, cls, items, nitems);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_batch_apply(hipStream_t s, bool correct, const KorClassDev *cls, const int4 *items, int nitems, int blocks,
                       double dt)
{
    if (correct)
        hipLaunchKernelGGL(k_batch_apply<true>, dim3(blocks), dim3(APPLY_THREADS), 0, s, cls, items, nitems, dt);
    else
        hipLaunchKernelGGL(k_batch_apply<false>, dim3(blocks), dim3(APPLY_THREADS), 0, s, cls, items, nitems, dt);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// x^r = exp(r log x) through the node kernels' table-driven log and exp
// (accuracy hook).
__global__ __launch_bounds__(256) void k_kmath_pow(const double *__restrict__ x, const double *__restrict__ r,
                                                   double *__restrict__ out, long n, const double *__restrict__ kmt)
{
    __shared__ double ts[KMT_SIZE];
    for (int i = threadIdx.x; i < KMT_SIZE; i += blockDim.x) ts[i] = kmt[i];
    __syncthreads();
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = kmt_exp(r[i] * kmt_log(x[i], ts), ts);
}

int launch_kmath_pow(gdt_context *h, const double *x, const double *r, double *out, long n)
{
    const long want = (n + 255) / 256;
    const int blocks = (int)(want < 1024 ? (want < 1 ? 1 : want) : 1024);
    hipLaunchKernelGGL(k_kmath_pow, dim3(blocks), dim3(256), 0, h->stream, x, r, out, n, h->kmath.as<double>());
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// Dynamic LDS of the kernels that stage f: f (padded to whole 128-double
// chunks) and the log / exp tables.
int kor_node_lds(const gdt_context *h) { return (int)(sizeof(double) * (kor_fpad(h->n_pnt) + KMT_SIZE)); }

int kor_blocks(gdt_context *h, int cells)
{
    (void)h;
    const int target = 256;  // one block per CU: one block's LDS per CU
    const int b = (target + cells - 1) / cells;
    // from KOR_GROUPS blocks per cell up: a multiple of KOR_GROUPS, so the node
    // kernel's ranges follow the gather groups' XCDs
    return b < KOR_GROUPS ? b : (b + KOR_GROUPS - 1) / KOR_GROUPS * KOR_GROUPS;
}

// Destination ranges per wave with balanced entry counts, for B blocks.
static int kor_wave_bounds(gdt_context *h, int B)
{
    if (h->wb_blocks == B) return GDT_OK;
    const int W = B * KOR_WAVES;
    std::vector<int32_t> wb(W + 1);
    const long ne = h->nent;
    for (int w = 0; w <= W; w++) {
        const long target = (long)((double)ne * w / W);
        wb[w] = (int32_t)(std::lower_bound(h->h_kptr.begin(), h->h_kptr.end() - 1, (int32_t)target) -
                          h->h_kptr.begin());
    }
    wb[0] = 0;
    wb[W] = h->n_pnt;
    GDT_CHECK(h->kwb.ensure(sizeof(int32_t) * (W + 1)) == 0, GDT_E_NOMEM, "wave-bound allocation failed");
    GDT_HIP(hipMemcpy(h->kwb.p, wb.data(), sizeof(int32_t) * (W + 1), hipMemcpyHostToDevice));
    h->wb_blocks = B;
    return GDT_OK;
}

// Outside graph capture: LDS limits and the wave partition for `cells`.
int kor_setup(gdt_context *h, int cells)
{
    const int lds = kor_node_lds(h);
    if (h->lds_attr < lds) {
        GDT_HIP(hipFuncSetAttribute((const void *)k_kor_node, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        GDT_HIP(hipFuncSetAttribute((const void *)k_kor_fused, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        GDT_HIP(hipFuncSetAttribute((const void *)k_kor_gather, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(double2) * KOR_TILE_CAP)));
        h->lds_attr = lds;
    }
    return kor_wave_bounds(h, kor_blocks(h, cells));
}

int launch_kor_node(gdt_context *h, const double *f, double *DW, int cells)
{
    GDT_CHECK_GRID_Y(cells, "kor_node: cells");
    const int B = kor_blocks(h, cells);
    GDT_LAUNCH_TIMED(h, "kor_node", k_kor_node, dim3(B, cells), dim3(KOR_THREADS), kor_node_lds(h), h->n_pnt, h->nn,
                     h->nv, h->kn4.as<uint4>(), h->kK, f, (double2 *)DW, h->kmath.as<double>(),
                     h->kcbase.as<int32_t>());
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int kor_gather_parts(gdt_context *h) { return KOR_GROUPS * h->kntile; }

// one block per tile (G * at most ktiles_max tiles per cell)
int kor_gather_blocks(gdt_context *h, int cells)
{
    (void)cells;
    return kor_gather_parts(h);
}

int launch_kor_gather(gdt_context *h, const double *V, double *Lpart, double *Qpart, double *mpart, int cells)
{
    GDT_CHECK_GRID_Y(cells, "kor_gather: cells");
    const int B = kor_gather_blocks(h, cells);
    GDT_LAUNCH_TIMED(h, "kor_gather", k_kor_gather, dim3(B, cells), dim3(TILE_THREADS),
                     (uint32_t)(sizeof(double2) * h->kcap), h->n_pnt, h->nv, h->kntile, h->kcap,
                     h->ktile.as<int4>(), h->kgptr.as<int32_t>(), h->kgent.as<uint2>(), (const double2 *)V,
                     h->xi.as<double>(), Lpart, Qpart, mpart);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_kor_fused(gdt_context *h, const double *f, double *loss, double *Q, double *mpart, int cells)
{
    GDT_CHECK_GRID_Y(cells, "kor_fused: cells");
    const int B = kor_blocks(h, cells);
    GDT_CHECK(h->wb_blocks == B, GDT_E_STATE, "fused: wave bounds for %d blocks not prepared", B);
    GDT_LAUNCH_TIMED(h, "kor_fused", k_kor_fused, dim3(B, cells), dim3(KOR_THREADS), kor_node_lds(h), h->n_pnt,
                     h->kwb.as<int32_t>(), h->kptr.as<int32_t>(), h->ke4.as<uint4>(), h->ke1.as<uint32_t>(), h->kK,
                     f, h->xi.as<double>(), loss, Q, mpart, h->kmath.as<double>());
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_kor_apply(gdt_context *h, bool correct, int G, int nparts, const double *mpart, const double *f,
                     const double *Lpart, const double *Qpart, double *Qout, double *Lout, double *nu_out, double *X,
                     int cells, double dt, double *f_upd)
{
    GDT_CHECK_GRID_Y(cells, "kor_apply: cells");
    GDT_CHECK(G == 1 || G == KOR_GROUPS, GDT_E_ARG, "kor_apply: %d partials per node", G);
    const dim3 grid((h->n_pnt + APPLY_THREADS - 1) / APPLY_THREADS, cells);
    auto kern = correct ? (G == 1 ? k_kor_apply<true, 1> : k_kor_apply<true, KOR_GROUPS>)
                        : (G == 1 ? k_kor_apply<false, 1> : k_kor_apply<false, KOR_GROUPS>);
    GDT_LAUNCH_TIMED(h, "kor_apply", kern, grid, dim3(APPLY_THREADS), 0u, h->n_pnt, nparts, mpart, h->xi.as<double>(), f, Lpart,
                     Qpart, Qout, Lout, nu_out, X, dt, f_upd);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}
