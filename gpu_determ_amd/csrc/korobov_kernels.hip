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
//   k_kor_node    node-parallel, f of the cell staged in LDS: D and 2wP of
//                 every node, once, written contiguously (double2 per node);
//   k_kor_gather  destination-sorted reduction, XCD-aware: the nodes are
//                 split into KOR_GROUPS contiguous ranges, each with its own
//                 destination-sorted CSR of (node | role, coefficient -1 /
//                 1-r / r) entries.  Workgroups are dealt to the XCDs round
//                 robin, so block b serves group b % KOR_GROUPS and every XCD
//                 gathers only its range's (D, 2wP) -- under 1 MB, resident in
//                 that XCD's L2.  A 16-lane quarter-wave owns a destination,
//                 sums its entries in a fixed order and writes the group's
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

#define KOR_THREADS 1024
#define KOR_WAVES (KOR_THREADS / WAVE)
#define KOR_GROUPS 8  // node groups of the gather: one per XCD

// D and 2 w P of one node from the LDS copy of f.  x = (a | b << 16,
// l | m << 16, ls | ms << 16, num | den << 16), D2 in the low 12 bits of y.
__device__ __forceinline__ void node_terms(const uint4 x, const uint32_t y, const double *fs, double K, double &D,
                                           double &W2P, double &r)
{
#pragma clang fp contract(off)
    const int num = (int)(short)(x.w & 0xFFFFu), den = (int)(short)(x.w >> 16);
    const double P = fs[x.x & 0xFFFFu] * fs[x.x >> 16];
    const double Pl = fs[x.y & 0xFFFFu] * fs[x.y >> 16];
    double R;
    r = 0.0;
    if (num == 0) {
        R = Pl;
    } else {
        r = (double)num / (double)den;
        const double Ps = fs[x.z & 0xFFFFu] * fs[x.z >> 16];
        if (Pl > 0.0 && Ps > 0.0)
            R = Pl * exp(r * log(Ps / Pl));
        else
            R = (1.0 - r) * Pl + r * Ps;
    }
    const double w = K * sqrt((double)(y & 0xFFFu));
    D = w * (P - R);
    W2P = 2.0 * (w * P);
}

// Stages f of cell c into LDS with 8 loads in flight per thread.
__device__ __forceinline__ void stage_f(double *fs, const double *fc, int npts)
{
    const int per = (npts + KOR_THREADS - 1) / KOR_THREADS;
    for (int k0 = 0; k0 < per; k0 += 8) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int t = (k0 + j) * KOR_THREADS + threadIdx.x;
            v[j] = t < npts ? fc[t] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int t = (k0 + j) * KOR_THREADS + threadIdx.x;
            if (t < npts) fs[t] = v[j];
        }
    }
}

// ---------------------------------------------------------------------------
// node values
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(KOR_THREADS) void k_kor_node(int npts, long nn, const uint4 *__restrict__ nrec,
                                                           const uint32_t *__restrict__ nd2, double K,
                                                           const double *__restrict__ f, double2 *__restrict__ DW)
{
    extern __shared__ double fs[];
    const int c = blockIdx.y;
    stage_f(fs, f + (size_t)c * npts, npts);
    __syncthreads();
    double2 *out = DW + (size_t)c * nn;
    for (long t = (long)blockIdx.x * KOR_THREADS + threadIdx.x; t < nn; t += (long)gridDim.x * KOR_THREADS) {
        double D, W2P, r;
        node_terms(nrec[t], nd2[t], fs, K, D, W2P, r);
        out[t] = make_double2(D, W2P);
    }
}

// ---------------------------------------------------------------------------
// segmented reduction helpers
// ---------------------------------------------------------------------------
// Inclusive scan of (q, l) over lanes whose keys are equal; keys form
// contiguous runs, so a lane off positions down with the same key is in the
// same run.
__device__ __forceinline__ void seg_scan(double &q, double &l, int key, int lane)
{
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const double qo = __shfl_up(q, off, WAVE);
        const double lo = __shfl_up(l, off, WAVE);
        const int ko = __shfl_up(key, off, WAVE);
        if (lane >= off && ko == key) {
            q += qo;
            l += lo;
        }
    }
}

// After a chunk's scan: the tail lane of every run holds the run's sum.
// Runs that end inside the chunk are complete (joined with the carry if they
// continue the previous chunk's open run); the chunk's last run stays open
// unless the wave's range ends here.  Completed destinations are written and
// added to the lane's moment sums.
struct SegState {
    int ckey = -1;
    double cq = 0.0, cl = 0.0;
};

__device__ __forceinline__ void seg_flush_carry(SegState &st, int kfirst, int lane, double *Qc, double *Lc,
                                                const double *xi, double (&m)[20])
{
    if (st.ckey >= 0 && kfirst != st.ckey) {
        if (lane == 0) {
            Qc[st.ckey] = st.cq;
            if (Lc) Lc[st.ckey] = st.cl;
            moments_add(m, xi, st.ckey, st.cq, st.cl);
        }
        st.ckey = -1;
    }
}

__device__ __forceinline__ void seg_emit(SegState &st, double q, double l, int key, bool valid, int nvalid,
                                         bool last_chunk, int lane, double *Qc, double *Lc, const double *xi,
                                         double (&m)[20])
{
    const int knext = __shfl_down(key, 1, WAVE);
    const bool tail = valid && (lane == nvalid - 1 || knext != key);
    const int klast = __shfl(key, nvalid - 1, WAVE);
    if (tail) {
        if (key == st.ckey) {
            q = st.cq + q;
            l = st.cl + l;
        }
        if (key != klast || last_chunk) {
            Qc[key] = q;
            if (Lc) Lc[key] = l;
            moments_add(m, xi, key, q, l);
        }
    }
    st.cq = __shfl(q, nvalid - 1, WAVE);
    st.cl = __shfl(l, nvalid - 1, WAVE);
    st.ckey = last_chunk ? -1 : klast;
}

__device__ __forceinline__ void block_moments(double (&m)[20], double (*wpart)[20], double *mpart, int c)
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    wave_sum_n<20>(m);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 20; q++) wpart[wave][q] = m[q];
    __syncthreads();
    if (threadIdx.x < 20) {
        double sm = 0.0;
        for (int w = 0; w < KOR_WAVES; w++) sm += wpart[w][threadIdx.x];
        mpart[((size_t)c * 20 + threadIdx.x) * gridDim.x + blockIdx.x] = sm;
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
// destination-sorted gather, one node group per XCD
// ---------------------------------------------------------------------------
#define QW 16  // lanes per destination (one DPP row)

__global__ __launch_bounds__(KOR_THREADS) void k_kor_gather(int npts, long nn, int dps,
                                                             const int32_t *__restrict__ gptr,
                                                             const uint32_t *__restrict__ ent,
                                                             const double *__restrict__ coef,
                                                             const double2 *__restrict__ DW,
                                                             const double *__restrict__ xi,
                                                             double *__restrict__ Lpart, double *__restrict__ Qpart,
                                                             double *__restrict__ mpart)
{
#pragma clang fp contract(off)
    __shared__ double wpart[KOR_WAVES][20];
    const int c = blockIdx.y;
    const int g = blockIdx.x % KOR_GROUPS, sub = blockIdx.x / KOR_GROUPS;
    const int d0 = sub * dps, d1 = min(npts, d0 + dps);
    const int32_t *ptr = gptr + (size_t)g * (npts + 1);
    const double2 *dw = DW + (size_t)c * nn;
    const size_t po = ((size_t)c * KOR_GROUPS + g) * npts;
    const int lane = threadIdx.x % WAVE, ql = threadIdx.x % QW, qwid = threadIdx.x / QW;
    double m[20];
#pragma unroll
    for (int q = 0; q < 20; q++) m[q] = 0.0;
    const int step = KOR_THREADS / QW;
    int dst = d0 + qwid;
    int e0 = dst < d1 ? ptr[dst] : 0, e1 = dst < d1 ? ptr[dst + 1] : 0;
    for (; dst < d1; dst += step) {
        // next destination's range, loaded ahead
        const int nd = dst + step;
        const int ne0 = nd < d1 ? ptr[nd] : 0, ne1 = nd < d1 ? ptr[nd + 1] : 0;
        double q = 0.0, l = 0.0;
        for (int e = e0 + ql; e < e1; e += 2 * QW) {
            const bool two = e + QW < e1;
            const uint32_t xa = ent[e];
            const uint32_t xb = two ? ent[e + QW] : 0u;
            const double ca = coef[e];
            const double cb = two ? coef[e + QW] : 0.0;
            const double2 va = dw[xa >> 2];
            const double2 vb = two ? dw[xb >> 2] : make_double2(0.0, 0.0);
            const double ta = ca * va.x;
            q += ta;
            if ((xa & 3u) == 0) l += va.y;
            if (two) {
                const double tb = cb * vb.x;
                q += tb;
                if ((xb & 3u) == 0) l += vb.y;
            }
        }
        q = row16_sum(q);
        l = row16_sum(l);
        if (ql == 0) {
            Qpart[po + dst] = q;
            Lpart[po + dst] = l;
            moments_add(m, xi, dst, q, l);
        }
        e0 = ne0;
        e1 = ne1;
    }
    const int wave = threadIdx.x / WAVE;
    wave_sum_n<20>(m);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 20; q++) wpart[wave][q] = m[q];
    __syncthreads();
    if (threadIdx.x < 20) {
        double sm = 0.0;
        for (int w = 0; w < KOR_WAVES; w++) sm += wpart[w][threadIdx.x];
        mpart[((size_t)c * 20 + threadIdx.x) * gridDim.x + blockIdx.x] = sm;
    }
}

// ---------------------------------------------------------------------------
// alternative: fused, per-entry recomputation
// ---------------------------------------------------------------------------
// One 20-byte record per CSR entry: the node's six indices, num | den, and
// D2 | role << 12 | destination << 14.
__global__ __launch_bounds__(KOR_THREADS) void k_kor_fused(int npts, const int32_t *__restrict__ wbound,
                                                            const int32_t *__restrict__ ptr,
                                                            const uint4 *__restrict__ e4,
                                                            const uint32_t *__restrict__ e1, double K,
                                                            const double *__restrict__ f,
                                                            const double *__restrict__ xi,
                                                            double *__restrict__ loss, double *__restrict__ Q,
                                                            double *__restrict__ mpart)
{
#pragma clang fp contract(off)
    extern __shared__ double fs[];
    __shared__ double wpart[KOR_WAVES][20];
    const int c = blockIdx.y;
    stage_f(fs, f + (size_t)c * npts, npts);
    __syncthreads();
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int wg = blockIdx.x * KOR_WAVES + wave;
    const int ga = wbound[wg], gb = wbound[wg + 1];
    const int ea = ptr[ga], eb = ptr[gb];
    double *Qc = Q + (size_t)c * npts;
    double *Lc = loss ? loss + (size_t)c * npts : nullptr;
    double m[20];
#pragma unroll
    for (int q = 0; q < 20; q++) m[q] = 0.0;
    zero_empty(ga, gb, ptr, Qc, Lc, lane);

    SegState st;
    uint4 xn = make_uint4(0, 0, 0, 0);
    uint32_t yn = 0;
    if (ea + lane < eb) {
        xn = e4[ea + lane];
        yn = e1[ea + lane];
    }
    for (int base = ea; base < eb; base += WAVE) {
        const uint4 x = xn;
        const uint32_t y = yn;
        const int e = base + lane;
        const bool valid = e < eb;
        if (base + WAVE + lane < eb) {
            xn = e4[base + WAVE + lane];
            yn = e1[base + WAVE + lane];
        }
        double q = 0.0, l = 0.0;
        int key = -2 - lane;
        if (valid) {
            double D, W2P, r;
            node_terms(x, y, fs, K, D, W2P, r);
            const uint32_t role = (y >> 12) & 3u;
            if (role == 0) {
                q = -1.0 * D;
                l = W2P;
            } else {
                q = (role == 1 ? (1.0 - r) : r) * D;
            }
            key = (int)(y >> 14);
        }
        seg_flush_carry(st, __shfl(key, 0, WAVE), lane, Qc, Lc, xi, m);
        seg_scan(q, l, key, lane);
        const int nvalid = min(WAVE, eb - base);
        seg_emit(st, q, l, key, valid, nvalid, base + WAVE >= eb, lane, Qc, Lc, xi, m);
    }
    block_moments(m, wpart, mpart, c);
}

// ---------------------------------------------------------------------------
// group sum, correction from the moment sums (+ explicit step)
// ---------------------------------------------------------------------------
// Every block reduces its cell's nparts x 20 moment sums in the same order
// (all loads issued at once) and solves the same 5x5 system; then each
// thread sums its node's G group partials in group order and writes
//   Q = Q0 - L X.psi,  loss = L,  nu = (L/f)(1 + X.psi),  f <- f + dt Q.
template <bool CORR>
__global__ __launch_bounds__(256) void k_kor_apply(int npts, int nparts, int G, const double *__restrict__ mpart,
                                                   const double *__restrict__ xi, const double *f,
                                                   const double *__restrict__ Lpart,
                                                   const double *__restrict__ Qpart, double *__restrict__ Qout,
                                                   double *__restrict__ Lout, double *__restrict__ nu_out,
                                                   double *__restrict__ Xout, double dt, double *f_upd)
{
    __shared__ double part[4][20];
    __shared__ double Xs[5];
    const int c = blockIdx.y;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t o = (size_t)c * npts + v;
    double x = 0.0, y = 0.0, z = 0.0, L = 0.0, fv = 0.0, q0 = 0.0;
    if (v < npts) {
        x = xi[3 * v];
        y = xi[3 * v + 1];
        z = xi[3 * v + 2];
        fv = f[o];
        for (int g = 0; g < G; g++) {
            const size_t po = ((size_t)c * G + g) * npts + v;
            q0 += Qpart[po];
            L += Lpart[po];
        }
    }
    if (CORR) {
        double m[20];
#pragma unroll
        for (int q = 0; q < 20; q++) m[q] = 0.0;
        for (int p0 = 0; p0 < nparts; p0 += blockDim.x) {
            const int p = p0 + threadIdx.x;
            if (p < nparts) {
#pragma unroll
                for (int q = 0; q < 20; q++) m[q] += mpart[((size_t)c * 20 + q) * nparts + p];
            }
        }
        const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
        wave_sum_n<20>(m);
        if (lane == 0)
#pragma unroll
            for (int q = 0; q < 20; q++) part[wave][q] = m[q];
        __syncthreads();
        if (threadIdx.x == 0) {
            double sm[20], X[5];
            for (int q = 0; q < 20; q++) sm[q] = part[0][q] + part[1][q] + part[2][q] + part[3][q];
            moments_solve(sm, X);
            for (int a = 0; a < 5; a++) Xs[a] = X[a];
        }
    } else if (threadIdx.x < 5) {
        Xs[threadIdx.x] = 0.0;
    }
    __syncthreads();
    if (Xout && blockIdx.x == 0 && threadIdx.x < 5) Xout[5 * c + threadIdx.x] = Xs[threadIdx.x];
    if (v >= npts) return;
    const double e = x * x + y * y + z * z;
    const double s = Xs[0] + Xs[1] * x + Xs[2] * y + Xs[3] * z + Xs[4] * e;
    const double q = q0 - L * s;
    if (Qout) Qout[o] = q;
    if (Lout) Lout[o] = L;
    if (nu_out) nu_out[o] = fv > 0.0 ? (L / fv) * (1.0 + s) : 0.0;
    if (dt != 0.0) f_upd[o] = fv + dt * q;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int kor_blocks(gdt_context *h, int cells)
{
    (void)h;
    const int target = 256;  // one block per CU: one block's LDS per CU
    const int b = (target + cells - 1) / cells;
    return b < 1 ? 1 : b;
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
    const int lds = (int)(sizeof(double) * h->n_pnt);
    if (h->lds_attr < lds) {
        GDT_HIP(hipFuncSetAttribute((const void *)k_kor_node, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        GDT_HIP(hipFuncSetAttribute((const void *)k_kor_fused, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        h->lds_attr = lds;
    }
    return kor_wave_bounds(h, kor_blocks(h, cells));
}

int launch_kor_node(gdt_context *h, const double *f, double *DW, int cells)
{
    const int B = kor_blocks(h, cells);
    KernelTimer kt(h, "kor_node");
    hipLaunchKernelGGL(k_kor_node, dim3(B, cells), dim3(KOR_THREADS), sizeof(double) * h->n_pnt, h->stream, h->n_pnt,
                       h->nn, h->kn4.as<uint4>(), h->kn1.as<uint32_t>(), h->kK, f, (double2 *)DW);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int kor_gather_blocks(gdt_context *h, int cells)
{
    // per group: about 256 blocks per cell in total (one round of waves)
    const int per = (256 / KOR_GROUPS + cells - 1) / cells;
    return KOR_GROUPS * (per < 1 ? 1 : per);
}

int launch_kor_gather(gdt_context *h, const double *DW, double *Lpart, double *Qpart, double *mpart, int cells)
{
    const int B = kor_gather_blocks(h, cells);
    const int dps = (h->n_pnt + B / KOR_GROUPS - 1) / (B / KOR_GROUPS);
    KernelTimer kt(h, "kor_gather");
    hipLaunchKernelGGL(k_kor_gather, dim3(B, cells), dim3(KOR_THREADS), 0, h->stream, h->n_pnt, h->nn, dps,
                       h->kgptr.as<int32_t>(), h->kgent.as<uint32_t>(), h->kgcoef.as<double>(), (const double2 *)DW,
                       h->xi.as<double>(), Lpart, Qpart, mpart);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_kor_fused(gdt_context *h, const double *f, double *loss, double *Q, double *mpart, int cells)
{
    const int B = kor_blocks(h, cells);
    GDT_CHECK(h->wb_blocks == B, GDT_E_STATE, "fused: wave bounds for %d blocks not prepared", B);
    KernelTimer kt(h, "kor_fused");
    hipLaunchKernelGGL(k_kor_fused, dim3(B, cells), dim3(KOR_THREADS), sizeof(double) * h->n_pnt, h->stream,
                       h->n_pnt, h->kwb.as<int32_t>(), h->kptr.as<int32_t>(), h->ke4.as<uint4>(),
                       h->ke1.as<uint32_t>(), h->kK, f, h->xi.as<double>(), loss, Q, mpart);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_kor_apply(gdt_context *h, bool correct, int G, int nparts, const double *mpart, const double *f,
                     const double *Lpart, const double *Qpart, double *Qout, double *Lout, double *nu_out, double *X,
                     int cells, double dt, double *f_upd)
{
    const dim3 grid((h->n_pnt + 255) / 256, cells);
    KernelTimer kt(h, "kor_apply");
    if (correct)
        hipLaunchKernelGGL(k_kor_apply<true>, grid, dim3(256), 0, h->stream, h->n_pnt, nparts, G, mpart,
                           h->xi.as<double>(), f, Lpart, Qpart, Qout, Lout, nu_out, X, dt, f_upd);
    else
        hipLaunchKernelGGL(k_kor_apply<false>, grid, dim3(256), 0, h->stream, h->n_pnt, nparts, G, mpart,
                           h->xi.as<double>(), f, Lpart, Qpart, Qout, Lout, nu_out, X, dt, f_upd);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}
