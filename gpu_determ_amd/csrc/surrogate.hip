// Learned surrogate of Q(f,f) with exact conservation projection and a
// plain-solver fallback (no reference counterpart).
//
// Network: a bilinear MLP on the node values of one velocity grid,
//   x = standardise(f / n),  h = (W1 x + b1) * (W2 x + b2)  (elementwise),
//   y = W3 h + b3,  Q_nn = n^2 unstandardise(y)
// (n = density: Q(f,f) is a quadratic form in f, so f / n -> Q / n^2 removes
// the density scale, and the hidden layer is a rank-H quadratic form of x --
// the operator's own structure, truncated).  bf16 weights and activations in the MFMA GEMMs (FP32 accumulation),
// FP32 master weights and Adam.  Training, forward and backward are the
// kernels below plus launch_gemm_nt; node counts are padded to 128.
//
// Scaling: per node the training mean is removed, and one global scale (the
// RMS of the per-node spreads) divides inputs and targets.
//
// Inference: Q_nn, then the FP64 projection onto zero mass / momentum /
// energy moments (weights f: Q = Q_nn - f X.psi with the 5x5 solve, the
// correction kernel), then per cell the residual
//   r = |Q - Q_nn|_2 / |Q|_2   (how far the network was from conserving),
// raised to at least 1 when the entropy production sum Q log f is positive; cells with
// r above the threshold are recomputed by the plain solver (the grid's
// Korobov stencil, corrected, in FP64 cell tiles).  With a partner network
// (gdt_surrogate_set_partner) Q_nn is the mean of the two predictions and r
// is raised to |Q_a - Q_b|_2 / |Q|_2: the two networks' disagreement follows
// their true error against the plain solver far more closely than the
// projection's change does (log-correlation over cells 0.97 against 0.65-0.8
// at 24^3).
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "gdt_internal.h"
#include "kernel_common.h"

__device__ __forceinline__ uint16_t sg_bf16(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// per-cell density sum f d3v (one block per cell)
__global__ __launch_bounds__(256) void k_sg_density(int n, const double *__restrict__ f, double d3v,
                                                    double *__restrict__ dens)
{
    __shared__ double part[4];
    const double *fc = f + (size_t)blockIdx.x * n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += fc[i];
    s = wave_sum(s);
    if (threadIdx.x % WAVE == 0) part[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) dens[blockIdx.x] = d3v * ((part[0] + part[1]) + (part[2] + part[3]));
}

// out[c][j] = (v[c][j] / dens[c]^pw - mu[j]) / sd[j] for j < n, 0 for n <= j < np
// (mu, sd NULL: unstandardised)
__global__ void k_sg_scale(long rows, int n, int np, const double *__restrict__ v, const double *__restrict__ dens,
                           int pw, const float *__restrict__ mu, const float *__restrict__ sd, float *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * np) return;
    const long c = t / np;
    const int j = (int)(t % np);
    float r = 0.0f;
    if (j < n) {
        const double d = pw == 1 ? dens[c] : dens[c] * dens[c];
        double x = v[c * n + j] / d;
        if (mu) x = (x - (double)mu[j]) / (double)sd[j];
        r = (float)x;
    }
    out[t] = r;
}

// column mean / standard deviation over rows (two passes, FP64 sums): a block
// of 256 threads takes 8 columns; thread (k, c) sums rows k, k + 32, ... of
// column c and the 32 row-slice sums are added in a fixed tree order
// (reproducible).  Columns n .. np-1 (padding) get mean 0, deviation 1.
// Grid: ceil(np / 8) blocks.
__global__ __launch_bounds__(256) void k_sg_stats(long rows, int n, int np, const float *__restrict__ x,
                                                  float *__restrict__ mu, float *__restrict__ sd)
{
    __shared__ double part[32][9];
    __shared__ double mean[8];
    const int cl = threadIdx.x & 7, k = threadIdx.x >> 3, j = blockIdx.x * 8 + cl;
    const bool live = j < n;
    double s = 0.0;
    if (live)
        for (long r = k; r < rows; r += 32) s += x[r * np + j];
    part[k][cl] = s;
    __syncthreads();
    for (int h = 16; h > 0; h >>= 1) {
        if (k < h) part[k][cl] += part[k + h][cl];
        __syncthreads();
    }
    if (k == 0) mean[cl] = part[0][cl] / (double)rows;
    __syncthreads();
    const double m = mean[cl];
    double q = 0.0;
    if (live)
        for (long r = k; r < rows; r += 32) {
            const double d = x[r * np + j] - m;
            q += d * d;
        }
    part[k][cl] = q;
    __syncthreads();
    for (int h = 16; h > 0; h >>= 1) {
        if (k < h) part[k][cl] += part[k + h][cl];
        __syncthreads();
    }
    if (k == 0 && j < np) {
        mu[j] = live ? (float)m : 0.0f;
        sd[j] = live ? (float)std::max(sqrt(part[0][cl] / (double)rows), 1e-12) : 1.0f;
    }
}

// standardise in place: x = (x - mu) / sd
__global__ void k_sg_standardise(long count, int np, float *__restrict__ x, const float *__restrict__ mu,
                                 const float *__restrict__ sd)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int j = (int)(t % np);
    x[t] = (x[t] - mu[j]) / sd[j];
}

// rows idx of src (fp32 [*][cols]) -> dst bf16 [B][cols] (and fp32 copy)
__global__ void k_sg_gather(int B, int cols, const int *__restrict__ idx, const float *__restrict__ src,
                            uint16_t *__restrict__ dh, float *__restrict__ df)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * cols) return;
    const int r = (int)(t / cols), j = (int)(t % cols);
    const float v = src[(long)idx[r] * cols + j];
    if (dh) dh[t] = sg_bf16(v);
    if (df) df[t] = v;
}

// bf16 [R][C] -> [C][R], 32 x 32 tiles
__global__ void k_sg_transpose(int R, int C, const uint16_t *__restrict__ src, uint16_t *__restrict__ dst)
{
    __shared__ uint16_t tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < R && c < C) tile[k][threadIdx.x] = src[(long)r * C + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < R && c < C) dst[(long)c * R + r] = tile[threadIdx.x][k];
    }
}

// hidden layer: u = p1 + b1, w = p2 + b2 (kept for the backward pass),
// h = u w (bf16: the next GEMM operand)
__global__ void k_sg_bilinear(int B, int H, float *__restrict__ p1, const float *__restrict__ b1,
                              float *__restrict__ p2, const float *__restrict__ b2, uint16_t *__restrict__ hh)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * H) return;
    const float u = p1[t] + b1[t % H], w = p2[t] + b2[t % H];
    p1[t] = u;
    p2[t] = w;
    hh[t] = sg_bf16(u * w);
}

// y += b3; dY = 2 (y - target) / (B n) on the n real columns (0 on padding);
// loss partial sums per block
__global__ __launch_bounds__(256) void k_sg_out_grad(int B, int n, int np, float *__restrict__ y,
                                                     const float *__restrict__ b3, const float *__restrict__ target,
                                                     uint16_t *__restrict__ dyh, double *__restrict__ loss_part)
{
    __shared__ double part[4];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (t < (long)B * np) {
        const int j = (int)(t % np);
        const float v = y[t] + b3[j];
        y[t] = v;
        float g = 0.0f;
        if (j < n) {
            const float d = v - target[t];
            l = (double)d * d;
            g = 2.0f * d / (float)((long)B * n);
        }
        dyh[t] = sg_bf16(g);
    }
    l = wave_sum(l);
    if (threadIdx.x % WAVE == 0) part[threadIdx.x / WAVE] = l;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// d(u) = dh w, d(w) = dh u -> bf16 (GEMM operands) and fp32 (bias sums)
__global__ void k_sg_dbilinear(int count, const float *__restrict__ dh, const float *__restrict__ u,
                               const float *__restrict__ w, uint16_t *__restrict__ duh, float *__restrict__ duf,
                               uint16_t *__restrict__ dwh, float *__restrict__ dwf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float a = dh[t] * w[t], b = dh[t] * u[t];
    duh[t] = sg_bf16(a);
    duf[t] = a;
    dwh[t] = sg_bf16(b);
    dwf[t] = b;
}

// column sums over B rows of a bf16 or fp32 [B][C] array (bias gradients):
// a block of 256 threads takes 8 columns; thread (k, c) adds rows k, k + 32,
// ... of column c, and the 32 row-slice sums are added in a fixed tree order
// (reproducible).  Grid: ceil(C / 8) blocks.
__global__ __launch_bounds__(256) void k_sg_colsum(int B, int C, const uint16_t *__restrict__ xh,
                                                   const float *__restrict__ xf, float *__restrict__ out)
{
    __shared__ float part[32][9];
    const int cl = threadIdx.x & 7, k = threadIdx.x >> 3, j = blockIdx.x * 8 + cl;
    float s = 0.0f;
    if (j < C)
        for (int r = k; r < B; r += 32)
            s += xh ? __uint_as_float((uint32_t)xh[(long)r * C + j] << 16) : xf[(long)r * C + j];
    part[k][cl] = s;
    __syncthreads();
    for (int h = 16; h > 0; h >>= 1) {
        if (k < h) part[k][cl] += part[k + h][cl];
        __syncthreads();
    }
    if (k == 0 && j < C) out[j] = part[0][cl];
}

// Adam step on FP32 master weights; refreshes the bf16 copy
__global__ void k_sg_adam(long count, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                          uint16_t *__restrict__ wh)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    const float wt = w[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
    w[t] = wt;
    if (wh) wh[t] = sg_bf16(wt);
}

// Q_nn[c][j] = dens[c]^2 (y sd + mu) for the n real columns (FP64)
__global__ void k_sg_unscale(long rows, int n, int np, const float *__restrict__ y, const float *__restrict__ mu,
                             const float *__restrict__ sd, const double *__restrict__ dens, double *__restrict__ Q)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * n) return;
    const long c = t / n;
    const int j = (int)(t % n);
    const double d = dens[c];
    Q[t] = d * d * ((double)y[c * np + j] * (double)sd[j] + (double)mu[j]);
}

// Two networks: Qa <- (Qa + Qb) / 2 (the prediction kept), dis2[c] = sum
// (Qa - Qb)^2 over the cell (one block per cell, fixed pairing order).
__global__ __launch_bounds__(256) void k_sg_pair(int n, double *__restrict__ Qa, const double *__restrict__ Qb,
                                                 double *__restrict__ dis2)
{
    __shared__ double part[4];
    const size_t o = (size_t)blockIdx.x * n;
    double d2 = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double a = Qa[o + i], b = Qb[o + i], d = a - b;
        d2 += d * d;
        Qa[o + i] = 0.5 * (a + b);
    }
    d2 = wave_sum(d2);
    if (threadIdx.x % WAVE == 0) part[threadIdx.x / WAVE] = d2;
    __syncthreads();
    if (threadIdx.x == 0) dis2[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// per cell: r = |Q - Qnn| / |Q|, raised to |Qa - Qb| / |Q| when two networks
// predict (dis2 non-NULL), and to 1 if sum Q log f > 0; flag r > thr
__global__ __launch_bounds__(256) void k_sg_residual(int n, const double *__restrict__ Q,
                                                     const double *__restrict__ Qnn, const double *__restrict__ f,
                                                     const double *__restrict__ dis2, double thr,
                                                     double *__restrict__ res, int *__restrict__ flag)
{
    __shared__ double part[4][3];
    const size_t o = (size_t)blockIdx.x * n;
    double d2 = 0.0, q2 = 0.0, ent = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double q = Q[o + i], e = q - Qnn[o + i], fv = f[o + i];
        d2 += e * e;
        q2 += q * q;
        if (fv > 0.0) ent += q * log(fv);
    }
    d2 = wave_sum(d2);
    q2 = wave_sum(q2);
    ent = wave_sum(ent);
    if (threadIdx.x % WAVE == 0) {
        part[threadIdx.x / WAVE][0] = d2;
        part[threadIdx.x / WAVE][1] = q2;
        part[threadIdx.x / WAVE][2] = ent;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double D = (part[0][0] + part[1][0]) + (part[2][0] + part[3][0]);
        const double S = (part[0][1] + part[1][1]) + (part[2][1] + part[3][1]);
        const double E = (part[0][2] + part[1][2]) + (part[2][2] + part[3][2]);
        // fails closed: a NaN or Inf anywhere in the cell's sums (a zero
        // density, an overflow in the network), or a zero Q against a
        // nonzero prediction, sends the cell to the plain solver
        double r = S > 0.0 ? sqrt(D / S) : (D > 0.0 ? 1.0 : 0.0);
        if (dis2) {
            const double P = dis2[blockIdx.x];
            const double rd = S > 0.0 ? sqrt(P / S) : (P > 0.0 ? 1.0 : 0.0);
            r = rd > r ? rd : r;
            if (!isfinite(P)) r = INFINITY;
        }
        if (E > 0.0 && !(r >= 1.0)) r = 1.0;  // entropy produced: at least 1
        const bool finite = isfinite(D) && isfinite(S) && isfinite(E) && isfinite(r);
        if (!finite) r = INFINITY;
        res[blockIdx.x] = r;
        flag[blockIdx.x] = (finite && r <= thr) ? 0 : 1;
    }
}

// rows listed in idx: dst[k] = src[idx[k]] (copy_in) or dst[idx[k]] = src[k]
__global__ void k_sg_rows(int k_rows, int n, const int *__restrict__ idx, const double *__restrict__ src,
                          double *__restrict__ dst, int scatter)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)k_rows * n) return;
    const int k = (int)(t / n), j = (int)(t % n);
    if (scatter)
        dst[(long)idx[k] * n + j] = src[t];
    else
        dst[t] = src[(long)idx[k] * n + j];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct gdt_surrogate {
    gdt_context *h = nullptr;
    int n = 0, np = 0, H = 0, B = 0;
    uint64_t seed = 0;
    long step = 0;
    bool trained = false;
    gdt_surrogate *partner = nullptr;  // second network of the gate (gdt_surrogate_set_partner)
    // FP32 master weights (W1, W2 [H][np], W3 [np][H]), Adam moments, bf16
    // copies (W3 also transposed for the backward pass)
    DevBuf W1, W2, W3, b1, b2, b3;
    DevBuf mW1, mW2, mW3, mb1, mb2, mb3, vW1, vW2, vW3, vb1, vb2, vb3;
    DevBuf W1h, W2h, W3h, W3th;
    DevBuf mux, sdx, muy, sdy;
    // batch scratch (sized for B rows)
    DevBuf X, Xt, T, P1, P2, Hh, Hth, Y, dYh, dYth, dH, dUh, dUf, dUth, dWh, dWf, dWth, gW1, gW2, gW3, gb1, gb2, gb3,
        lossp, idx;
    // inference scratch
    DevBuf dens, Qnn, res, flg, sub_f, sub_Q, Qnn2, dis2;
    DevBuf gws;  // split-K partials of the skinny GEMMs
};

#define SG_FALLBACK_CHUNK 4096  // cells per plain-solver call of the fallback

static int sg_blocks(long n) { return (int)std::max(1L, (n + 255) / 256); }

// C = A Bm^T (bf16 operands, FP32 C [M][N]); skinny products split K into
// the surrogate's GEMM workspace.
static int sg_gemm(gdt_surrogate *s, int M, int N, int K, const DevBuf &A, const DevBuf &Bm, DevBuf &C)
{
    const size_t need = (size_t)gemm_k_splits(GDT_GEMM_BF16, M, N, K) * M * N;
    float *ws = nullptr;
    if (need > (size_t)M * N) {
        GDT_CHECK(s->gws.ensure(sizeof(float) * need) == 0, GDT_E_NOMEM, "surrogate: GEMM workspace allocation failed");
        ws = s->gws.as<float>();
    }
    return launch_gemm_nt(s->h->stream, GDT_GEMM_BF16, M, N, K, A.p, nullptr, Bm.p, nullptr, C.as<float>(), N, 0, ws,
                          ws ? need : 0);
}

static int sg_transpose(gdt_surrogate *s, int R, int C, const DevBuf &src, DevBuf &dst)
{
    hipLaunchKernelGGL(k_sg_transpose, dim3((C + 31) / 32, (R + 31) / 32), dim3(32, 8), 0, s->h->stream, R, C,
                       src.as<uint16_t>(), dst.as<uint16_t>());
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

static int sg_alloc_batch(gdt_surrogate *s, int B)
{
    if (s->B == B) return GDT_OK;
    const size_t np = s->np, H = s->H;
    bool ok = s->X.ensure(2 * B * np) == 0 && s->Xt.ensure(2 * B * np) == 0 && s->T.ensure(4 * B * np) == 0 &&
              s->P1.ensure(4 * B * H) == 0 && s->P2.ensure(4 * B * H) == 0 && s->Hh.ensure(2 * B * H) == 0 &&
              s->Hth.ensure(2 * B * H) == 0 && s->Y.ensure(4 * B * np) == 0 && s->dYh.ensure(2 * B * np) == 0 &&
              s->dYth.ensure(2 * B * np) == 0 && s->dH.ensure(4 * B * H) == 0 && s->dUh.ensure(2 * B * H) == 0 &&
              s->dUf.ensure(4 * B * H) == 0 && s->dUth.ensure(2 * B * H) == 0 && s->dWh.ensure(2 * B * H) == 0 &&
              s->dWf.ensure(4 * B * H) == 0 && s->dWth.ensure(2 * B * H) == 0 &&
              s->lossp.ensure(8 * (size_t)sg_blocks((long)B * np)) == 0 && s->idx.ensure(4 * (size_t)B) == 0;
    GDT_CHECK(ok, GDT_E_NOMEM, "surrogate: batch scratch allocation failed");
    s->B = B;
    return GDT_OK;
}

// Forward of B rows already in s->X (bf16 [B][np]); Y fp32 [B][np] without b3.
static int sg_forward(gdt_surrogate *s, int B)
{
    const int np = s->np, H = s->H;
    int rc;
    if ((rc = sg_gemm(s, B, H, np, s->X, s->W1h, s->P1))) return rc;
    if ((rc = sg_gemm(s, B, H, np, s->X, s->W2h, s->P2))) return rc;
    hipLaunchKernelGGL(k_sg_bilinear, dim3(sg_blocks((long)B * H)), dim3(256), 0, s->h->stream, B, H,
                       s->P1.as<float>(), (const float *)s->b1.p, s->P2.as<float>(), (const float *)s->b2.p,
                       s->Hh.as<uint16_t>());
    if ((rc = sg_gemm(s, B, np, H, s->Hh, s->W3h, s->Y))) return rc;
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

static int sg_adam(gdt_surrogate *s, long count, DevBuf &w, DevBuf &g, DevBuf &m, DevBuf &v, DevBuf *wh, float lr)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)s->step), c2 = 1.0f - powf(b2, (float)s->step);
    hipLaunchKernelGGL(k_sg_adam, dim3(sg_blocks(count)), dim3(256), 0, s->h->stream, count, w.as<float>(),
                       g.as<float>(), m.as<float>(), v.as<float>(), lr, b1, b2, eps, c1, c2,
                       wh ? wh->as<uint16_t>() : nullptr);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int gdt_surrogate_create(gdt_surrogate_handle *out, gdt_handle h, int hidden, uint64_t seed)
{
    GDT_CHECK(out && h && hidden > 0 && hidden % 128 == 0, GDT_E_ARG,
              "gdt_surrogate_create: bad arguments (hidden %d must be a positive multiple of 128)", hidden);
    *out = nullptr;
    gdt_surrogate *s = new gdt_surrogate();
    s->h = h;
    s->n = h->n_pnt;
    s->np = (h->n_pnt + 127) / 128 * 128;
    s->H = hidden;
    s->seed = seed;
    const size_t np = s->np, H = s->H, w = H * np;
    bool ok = s->W1.ensure(4 * w) == 0 && s->W2.ensure(4 * w) == 0 && s->W3.ensure(4 * w) == 0 &&
              s->mW1.ensure(4 * w) == 0 && s->mW2.ensure(4 * w) == 0 && s->mW3.ensure(4 * w) == 0 &&
              s->vW1.ensure(4 * w) == 0 && s->vW2.ensure(4 * w) == 0 && s->vW3.ensure(4 * w) == 0 &&
              s->b1.ensure(4 * H) == 0 && s->b2.ensure(4 * H) == 0 && s->b3.ensure(4 * np) == 0 &&
              s->mb1.ensure(4 * H) == 0 && s->mb2.ensure(4 * H) == 0 && s->mb3.ensure(4 * np) == 0 &&
              s->vb1.ensure(4 * H) == 0 && s->vb2.ensure(4 * H) == 0 && s->vb3.ensure(4 * np) == 0 &&
              s->W1h.ensure(2 * w) == 0 && s->W2h.ensure(2 * w) == 0 && s->W3h.ensure(2 * w) == 0 &&
              s->W3th.ensure(2 * w) == 0 && s->mux.ensure(4 * np) == 0 && s->sdx.ensure(4 * np) == 0 &&
              s->muy.ensure(4 * np) == 0 && s->sdy.ensure(4 * np) == 0 && s->gW1.ensure(4 * w) == 0 &&
              s->gW2.ensure(4 * w) == 0 && s->gW3.ensure(4 * w) == 0 && s->gb1.ensure(4 * H) == 0 &&
              s->gb2.ensure(4 * H) == 0 && s->gb3.ensure(4 * np) == 0;
    if (!ok) {
        gdt_surrogate_destroy(s);
        gdt_set_error(GDT_E_NOMEM, "gdt_surrogate_create: allocation failed");
        return GDT_E_NOMEM;
    }
    // initial weights: N(0, 1 / fan_in) on the real columns, zero biases and
    // Adam moments
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    std::vector<float> w1(w, 0.0f), w2(w, 0.0f), w3(w, 0.0f);
    for (size_t j = 0; j < H; j++)
        for (int k = 0; k < s->n; k++) {
            w1[j * np + k] = nd(rng) / sqrtf((float)s->n);
            w2[j * np + k] = nd(rng) / sqrtf((float)s->n);
        }
    for (int o = 0; o < s->n; o++)
        for (size_t k = 0; k < H; k++) w3[o * H + k] = nd(rng) / sqrtf((float)H);
    int rc = GDT_OK;
    auto up = [&](DevBuf &d, const std::vector<float> &v) {
        if (hipMemcpy(d.p, v.data(), 4 * v.size(), hipMemcpyHostToDevice) != hipSuccess) rc = GDT_E_HIP;
    };
    up(s->W1, w1);
    up(s->W2, w2);
    up(s->W3, w3);
    DevBuf *zero[] = {&s->b1, &s->b2, &s->b3, &s->mW1, &s->mW2, &s->mW3, &s->mb1, &s->mb2, &s->mb3,
                      &s->vW1, &s->vW2, &s->vW3, &s->vb1, &s->vb2, &s->vb3};
    for (DevBuf *d : zero)
        if (hipMemset(d->p, 0, d->bytes) != hipSuccess) rc = GDT_E_HIP;
    if (rc) {
        gdt_surrogate_destroy(s);
        gdt_set_error(GDT_E_HIP, "gdt_surrogate_create: upload failed");
        return rc;
    }
    // bf16 copies: an Adam step with zero gradients and zero rate
    s->step = 1;
    if ((rc = sg_adam(s, (long)w, s->W1, s->mW1, s->mW1, s->vW1, &s->W1h, 0.0f)) ||
        (rc = sg_adam(s, (long)w, s->W2, s->mW2, s->mW2, s->vW2, &s->W2h, 0.0f)) ||
        (rc = sg_adam(s, (long)w, s->W3, s->mW3, s->mW3, s->vW3, &s->W3h, 0.0f)) ||
        (rc = sg_transpose(s, np, H, s->W3h, s->W3th))) {
        gdt_surrogate_destroy(s);
        return rc;
    }
    s->step = 0;
    GDT_HIP(hipStreamSynchronize(h->stream));
    *out = s;
    return GDT_OK;
}

int gdt_surrogate_destroy(gdt_surrogate_handle s)
{
    if (!s) return GDT_OK;
    DevBuf *all[] = {&s->W1, &s->W2, &s->W3, &s->b1, &s->b2, &s->b3, &s->mW1, &s->mW2, &s->mW3, &s->mb1, &s->mb2,
                     &s->mb3, &s->vW1, &s->vW2, &s->vW3, &s->vb1, &s->vb2, &s->vb3, &s->W1h, &s->W2h, &s->W3h,
                     &s->W3th, &s->mux, &s->sdx, &s->muy, &s->sdy, &s->X, &s->Xt, &s->T, &s->P1, &s->P2, &s->Hh,
                     &s->Hth, &s->Y, &s->dYh, &s->dYth, &s->dH, &s->dUh, &s->dUf, &s->dUth, &s->dWh, &s->dWf,
                     &s->dWth, &s->gW1, &s->gW2, &s->gW3, &s->gb1, &s->gb2, &s->gb3, &s->lossp, &s->idx, &s->dens,
                     &s->Qnn, &s->res, &s->flg, &s->sub_f, &s->sub_Q, &s->Qnn2, &s->dis2, &s->gws};
    for (DevBuf *d : all) d->release();
    delete s;
    return GDT_OK;
}

int gdt_surrogate_train(gdt_surrogate_handle s, const double *f, const double *Q, long ns, int epochs, int batch,
                        double lr, double *loss_out)
{
    GDT_CHECK(s && f && Q && ns >= batch && epochs >= 0 && batch > 0 && batch % 128 == 0, GDT_E_ARG,
              "gdt_surrogate_train: bad arguments (batch %d must be a positive multiple of 128, <= %ld samples)",
              batch, ns);
    hipStream_t st = s->h->stream;
    const int n = s->n, np = s->np, H = s->H, B = batch;
    const long w = (long)H * np;
    int rc = sg_alloc_batch(s, B);
    if (rc) return rc;
    // standardised inputs and targets of all samples (fp32 [ns][np])
    DevBuf Xall, Yall, dens;
    GDT_CHECK(Xall.ensure(4 * (size_t)ns * np) == 0 && Yall.ensure(4 * (size_t)ns * np) == 0 &&
                  dens.ensure(8 * (size_t)ns) == 0,
              GDT_E_NOMEM, "gdt_surrogate_train: sample allocation failed (%ld samples)", ns);
    hipLaunchKernelGGL(k_sg_density, dim3((unsigned)ns), dim3(256), 0, st, n, f, s->h->d3v, dens.as<double>());
    hipLaunchKernelGGL(k_sg_scale, dim3(sg_blocks(ns * np)), dim3(256), 0, st, ns, n, np, f,
                       (const double *)dens.p, 1, (const float *)nullptr, (const float *)nullptr, Xall.as<float>());
    hipLaunchKernelGGL(k_sg_scale, dim3(sg_blocks(ns * np)), dim3(256), 0, st, ns, n, np, Q,
                       (const double *)dens.p, 2, (const float *)nullptr, (const float *)nullptr, Yall.as<float>());
    hipLaunchKernelGGL(k_sg_stats, dim3((np + 7) / 8), dim3(256), 0, st, ns, n, np, Xall.as<float>(),
                       s->mux.as<float>(), s->sdx.as<float>());
    hipLaunchKernelGGL(k_sg_stats, dim3((np + 7) / 8), dim3(256), 0, st, ns, n, np, Yall.as<float>(),
                       s->muy.as<float>(), s->sdy.as<float>());
    // one scale for all nodes (per-node centring only): the training loss is
    // then the plain L2 error of the centred Q / n^2, the metric the surrogate
    // is judged by, instead of weighting every tail node like the bulk
    for (DevBuf *sd : {&s->sdx, &s->sdy}) {
        std::vector<float> hs(np);
        GDT_HIP(hipMemcpyAsync(hs.data(), sd->p, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
        GDT_HIP(hipStreamSynchronize(st));
        double q = 0.0;
        for (int j = 0; j < n; j++) q += (double)hs[j] * hs[j];
        const float g = (float)std::max(sqrt(q / n), 1e-30);
        for (int j = 0; j < np; j++) hs[j] = g;
        GDT_HIP(hipMemcpyAsync(sd->p, hs.data(), 4 * (size_t)np, hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_sg_standardise, dim3(sg_blocks(ns * np)), dim3(256), 0, st, ns * np, np, Xall.as<float>(),
                       s->mux.as<float>(), s->sdx.as<float>());
    hipLaunchKernelGGL(k_sg_standardise, dim3(sg_blocks(ns * np)), dim3(256), 0, st, ns * np, np, Yall.as<float>(),
                       s->muy.as<float>(), s->sdy.as<float>());
    GDT_HIP(hipGetLastError());
    std::mt19937_64 rng(s->seed + 17);
    std::vector<int> perm(ns);
    for (long i = 0; i < ns; i++) perm[i] = (int)i;
    const long steps_per_epoch = ns / B;
    const int nlp = sg_blocks((long)B * np);
    std::vector<double> lp(nlp);
    double last_loss = 0.0;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        for (long it = 0; it < steps_per_epoch; it++) {
            GDT_HIP(hipMemcpyAsync(s->idx.p, perm.data() + it * B, 4 * (size_t)B, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_sg_gather, dim3(sg_blocks((long)B * np)), dim3(256), 0, st, B, np,
                               (const int *)s->idx.p, (const float *)Xall.p, s->X.as<uint16_t>(), (float *)nullptr);
            hipLaunchKernelGGL(k_sg_gather, dim3(sg_blocks((long)B * np)), dim3(256), 0, st, B, np,
                               (const int *)s->idx.p, (const float *)Yall.p, (uint16_t *)nullptr, s->T.as<float>());
            if ((rc = sg_transpose(s, B, np, s->X, s->Xt))) return rc;
            if ((rc = sg_forward(s, B))) return rc;
            hipLaunchKernelGGL(k_sg_out_grad, dim3(nlp), dim3(256), 0, st, B, n, np, s->Y.as<float>(),
                               (const float *)s->b3.p, (const float *)s->T.p, s->dYh.as<uint16_t>(),
                               s->lossp.as<double>());
            if ((rc = sg_transpose(s, B, np, s->dYh, s->dYth))) return rc;
            if ((rc = sg_transpose(s, B, H, s->Hh, s->Hth))) return rc;
            // output layer: gW3 = dY^T h, gb3 = sum dY, dh = dY W3
            if ((rc = sg_gemm(s, np, H, B, s->dYth, s->Hth, s->gW3))) return rc;
            hipLaunchKernelGGL(k_sg_colsum, dim3((np + 7) / 8), dim3(256), 0, st, B, np,
                               (const uint16_t *)s->dYh.p, (const float *)nullptr, s->gb3.as<float>());
            if ((rc = sg_gemm(s, B, H, np, s->dYh, s->W3th, s->dH))) return rc;
            // bilinear layer: du = dh w, dw = dh u
            hipLaunchKernelGGL(k_sg_dbilinear, dim3(sg_blocks((long)B * H)), dim3(256), 0, st, B * H,
                               (const float *)s->dH.p, (const float *)s->P1.p, (const float *)s->P2.p,
                               s->dUh.as<uint16_t>(), s->dUf.as<float>(), s->dWh.as<uint16_t>(), s->dWf.as<float>());
            if ((rc = sg_transpose(s, B, H, s->dUh, s->dUth))) return rc;
            if ((rc = sg_transpose(s, B, H, s->dWh, s->dWth))) return rc;
            if ((rc = sg_gemm(s, H, np, B, s->dUth, s->Xt, s->gW1))) return rc;
            if ((rc = sg_gemm(s, H, np, B, s->dWth, s->Xt, s->gW2))) return rc;
            hipLaunchKernelGGL(k_sg_colsum, dim3((H + 7) / 8), dim3(256), 0, st, B, H, (const uint16_t *)nullptr,
                               (const float *)s->dUf.p, s->gb1.as<float>());
            hipLaunchKernelGGL(k_sg_colsum, dim3((H + 7) / 8), dim3(256), 0, st, B, H, (const uint16_t *)nullptr,
                               (const float *)s->dWf.p, s->gb2.as<float>());
            GDT_HIP(hipGetLastError());
            // Adam
            s->step++;
            const float lrf = (float)lr;
            if ((rc = sg_adam(s, w, s->W1, s->gW1, s->mW1, s->vW1, &s->W1h, lrf)) ||
                (rc = sg_adam(s, w, s->W2, s->gW2, s->mW2, s->vW2, &s->W2h, lrf)) ||
                (rc = sg_adam(s, w, s->W3, s->gW3, s->mW3, s->vW3, &s->W3h, lrf)) ||
                (rc = sg_adam(s, H, s->b1, s->gb1, s->mb1, s->vb1, nullptr, lrf)) ||
                (rc = sg_adam(s, H, s->b2, s->gb2, s->mb2, s->vb2, nullptr, lrf)) ||
                (rc = sg_adam(s, np, s->b3, s->gb3, s->mb3, s->vb3, nullptr, lrf)) ||
                (rc = sg_transpose(s, np, H, s->W3h, s->W3th)))
                return rc;
            if (it == steps_per_epoch - 1) {
                GDT_HIP(hipMemcpyAsync(lp.data(), s->lossp.p, 8 * nlp, hipMemcpyDeviceToHost, st));
                GDT_HIP(hipStreamSynchronize(st));
                double sm = 0.0;
                for (double x : lp) sm += x;
                last_loss = sm / ((double)B * n);
            }
        }
    }
    GDT_HIP(hipStreamSynchronize(st));
    Xall.release();
    Yall.release();
    dens.release();
    s->trained = true;
    if (loss_out) *loss_out = last_loss;
    return GDT_OK;
}

// Network prediction Q_nn (FP64 [cells][n]) for device f; densities in s->dens.
static int sg_predict(gdt_surrogate *s, const double *f, int cells, double *Qnn)
{
    hipStream_t st = s->h->stream;
    const int n = s->n, np = s->np, B = s->B > 0 ? s->B : 256;
    int rc = sg_alloc_batch(s, B);
    if (rc) return rc;
    GDT_CHECK(s->dens.ensure(8 * (size_t)cells) == 0, GDT_E_NOMEM, "surrogate: allocation failed");
    hipLaunchKernelGGL(k_sg_density, dim3(cells), dim3(256), 0, st, n, f, s->h->d3v, s->dens.as<double>());
    DevBuf xin;
    GDT_CHECK(xin.ensure(4 * (size_t)B * np) == 0, GDT_E_NOMEM, "surrogate: allocation failed");
    std::vector<int> ident(B);
    for (int c0 = 0; c0 < cells; c0 += B) {
        const int k = std::min(B, cells - c0);
        // standardised inputs of rows c0 .. c0 + k (padding rows zero)
        GDT_HIP(hipMemsetAsync(xin.p, 0, 4 * (size_t)B * np, st));
        hipLaunchKernelGGL(k_sg_scale, dim3(sg_blocks((long)k * np)), dim3(256), 0, st, (long)k, n, np,
                           f + (size_t)c0 * n, (const double *)s->dens.p + c0, 1, (const float *)s->mux.p,
                           (const float *)s->sdx.p, xin.as<float>());
        for (int i = 0; i < B; i++) ident[i] = i;
        GDT_HIP(hipMemcpyAsync(s->idx.p, ident.data(), 4 * (size_t)B, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_sg_gather, dim3(sg_blocks((long)B * np)), dim3(256), 0, st, B, np,
                           (const int *)s->idx.p, (const float *)xin.p, s->X.as<uint16_t>(), (float *)nullptr);
        if ((rc = sg_forward(s, B))) return rc;
        // y + b3 (the gradient output is unused here)
        hipLaunchKernelGGL(k_sg_out_grad, dim3(sg_blocks((long)B * np)), dim3(256), 0, st, B, n, np,
                           s->Y.as<float>(), (const float *)s->b3.p, (const float *)s->Y.p, s->dYh.as<uint16_t>(),
                           s->lossp.as<double>());
        hipLaunchKernelGGL(k_sg_unscale, dim3(sg_blocks((long)k * n)), dim3(256), 0, st, (long)k, n, np,
                           (const float *)s->Y.p, (const float *)s->muy.p, (const float *)s->sdy.p,
                           (const double *)s->dens.p + c0, Qnn + (size_t)c0 * n);
        GDT_HIP(hipGetLastError());
    }
    GDT_HIP(hipStreamSynchronize(st));
    xin.release();
    return GDT_OK;
}

int gdt_surrogate_predict(gdt_surrogate_handle s, const double *f, double *Q, int cells)
{
    GDT_CHECK(s && f && Q && cells > 0, GDT_E_ARG, "gdt_surrogate_predict: bad arguments");
    return sg_predict(s, f, cells, Q);
}

int gdt_surrogate_infer(gdt_surrogate_handle s, const double *f, double *Q, int cells, double threshold,
                        int *fallback, int *n_fallback)
{
    GDT_CHECK(s && f && Q && cells > 0, GDT_E_ARG, "gdt_surrogate_infer: bad arguments");
    gdt_context *h = s->h;
    hipStream_t st = h->stream;
    const int n = s->n;
    GDT_CHECK(s->Qnn.ensure(8 * (size_t)cells * n) == 0 && s->res.ensure(8 * (size_t)cells) == 0 &&
                  s->flg.ensure(4 * (size_t)cells) == 0,
              GDT_E_NOMEM, "gdt_surrogate_infer: allocation failed");
    int rc = sg_predict(s, f, cells, s->Qnn.as<double>());
    if (rc) return rc;
    const double *dis2 = nullptr;
    if (s->partner) {
        // second network: keep the mean, measure the disagreement
        GDT_CHECK(s->Qnn2.ensure(8 * (size_t)cells * n) == 0 && s->dis2.ensure(8 * (size_t)cells) == 0, GDT_E_NOMEM,
                  "gdt_surrogate_infer: allocation failed");
        if ((rc = sg_predict(s->partner, f, cells, s->Qnn2.as<double>()))) return rc;
        hipLaunchKernelGGL(k_sg_pair, dim3(cells), dim3(256), 0, st, n, s->Qnn.as<double>(),
                           (const double *)s->Qnn2.p, s->dis2.as<double>());
        GDT_HIP(hipGetLastError());
        dis2 = s->dis2.as<double>();
    }
    // exact conservation: Q = Q_nn - f X.psi (weights f), FP64 5x5 per cell
    if ((rc = launch_correct(h, GDT_METHOD_KOROBOV, f, f, s->Qnn.as<double>(), nullptr, Q, nullptr, cells, 0.0,
                             nullptr)))
        return rc;
    hipLaunchKernelGGL(k_sg_residual, dim3(cells), dim3(256), 0, st, n, (const double *)Q, (const double *)s->Qnn.p,
                       f, dis2, threshold, s->res.as<double>(), s->flg.as<int>());
    GDT_HIP(hipGetLastError());
    std::vector<int> flg(cells);
    GDT_HIP(hipMemcpyAsync(flg.data(), s->flg.p, 4 * (size_t)cells, hipMemcpyDeviceToHost, st));
    GDT_HIP(hipStreamSynchronize(st));
    std::vector<int> idx;
    for (int c = 0; c < cells; c++)
        if (flg[c]) idx.push_back(c);
    const int k = (int)idx.size();
    if (fallback) GDT_HIP(hipMemcpy(fallback, s->flg.p, 4 * (size_t)cells, hipMemcpyDeviceToDevice));
    if (n_fallback) *n_fallback = k;
    if (k == 0) return GDT_OK;
    // plain solver on the flagged cells: the corrected Korobov operator in
    // FP64 cell tiles (the many-cell form of the plain method, <= 1e-12 of the
    // FP64 oracle; a cell's result does not depend on the cells beside it), in
    // chunks of SG_FALLBACK_CHUNK cells so its scratch stays bounded
    GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_surrogate_infer: %d cells need the plain solver, no stencil built", k);
    const int ch = std::min(k, SG_FALLBACK_CHUNK);
    DevBuf didx;
    GDT_CHECK(didx.ensure(4 * (size_t)k) == 0 && s->sub_f.ensure(8 * (size_t)k * n) == 0 &&
                  s->sub_Q.ensure(8 * (size_t)k * n) == 0,
              GDT_E_NOMEM, "gdt_surrogate_infer: fallback allocation failed");
    if ((rc = ct_prepare(h, ch, false))) return rc;
    GDT_HIP(hipMemcpyAsync(didx.p, idx.data(), 4 * (size_t)k, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_sg_rows, dim3(sg_blocks((long)k * n)), dim3(256), 0, st, k, n, (const int *)didx.p, f,
                       s->sub_f.as<double>(), 0);
    GDT_HIP(hipGetLastError());
    for (int c0 = 0; c0 < k; c0 += ch) {
        const int kc = std::min(ch, k - c0);
        if ((rc = ct_run(h, false, s->sub_f.as<double>() + (size_t)c0 * n, kc, true,
                         s->sub_Q.as<double>() + (size_t)c0 * n, nullptr, nullptr, 0.0, nullptr)))
            return rc;
    }
    hipLaunchKernelGGL(k_sg_rows, dim3(sg_blocks((long)k * n)), dim3(256), 0, st, k, n, (const int *)didx.p,
                       (const double *)s->sub_Q.p, Q, 1);
    GDT_HIP(hipGetLastError());
    GDT_HIP(hipStreamSynchronize(st));
    didx.release();
    return GDT_OK;
}

int gdt_surrogate_set_partner(gdt_surrogate_handle s, gdt_surrogate_handle partner)
{
    GDT_CHECK(s && partner != s, GDT_E_ARG, "gdt_surrogate_set_partner: bad arguments");
    GDT_CHECK(!partner || (partner->h == s->h && partner->n == s->n), GDT_E_ARG,
              "gdt_surrogate_set_partner: the partner must be built on the same context");
    s->partner = partner;
    return GDT_OK;
}

int gdt_surrogate_residuals(gdt_surrogate_handle s, double *res, int cells)
{
    GDT_CHECK(s && res && cells > 0 && s->res.bytes >= 8 * (size_t)cells, GDT_E_ARG,
              "gdt_surrogate_residuals: bad arguments");
    GDT_HIP(hipMemcpy(res, s->res.p, 8 * (size_t)cells, hipMemcpyDeviceToHost));
    return GDT_OK;
}

int gdt_surrogate_export(gdt_surrogate_handle s, float *W1, float *b1, float *W2, float *b2, float *W3, float *b3,
                         float *norm)
{
    GDT_CHECK(s, GDT_E_ARG, "gdt_surrogate_export: null handle");
    const size_t np = s->np, H = s->H;
    auto dn = [&](float *dst, const DevBuf &src, size_t count) -> int {
        if (dst) GDT_HIP(hipMemcpy(dst, src.p, 4 * count, hipMemcpyDeviceToHost));
        return GDT_OK;
    };
    int rc;
    if ((rc = dn(W1, s->W1, H * np)) || (rc = dn(b1, s->b1, H)) || (rc = dn(W2, s->W2, H * np)) ||
        (rc = dn(b2, s->b2, H)) || (rc = dn(W3, s->W3, np * H)) || (rc = dn(b3, s->b3, np)))
        return rc;
    if (norm) {
        if ((rc = dn(norm, s->mux, np)) || (rc = dn(norm + np, s->sdx, np)) || (rc = dn(norm + 2 * np, s->muy, np)) ||
            (rc = dn(norm + 3 * np, s->sdy, np)))
            return rc;
    }
    return GDT_OK;
}

int gdt_surrogate_info(gdt_surrogate_handle s, int *n, int *np, int *hidden, long *params)
{
    GDT_CHECK(s, GDT_E_ARG, "gdt_surrogate_info: null handle");
    if (n) *n = s->n;
    if (np) *np = s->np;
    if (hidden) *hidden = s->H;
    if (params) *params = 3L * s->n * s->H + 2L * s->H + s->n;
    return GDT_OK;
}
