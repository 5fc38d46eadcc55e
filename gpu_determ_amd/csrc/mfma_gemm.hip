// Dense contractions on the matrix cores (gfx950 MFMA), FP32 accumulation:
//   C[M][N] (+)= sum_k A[M][K] * B[N][K]          ("NT": both operands K-contiguous)
// Operand modes:
//   GDT_GEMM_BF16        A, B bf16
//   GDT_GEMM_BF16_SPLIT  A = Ah + Al, B = Bh + Bl (bf16 pairs from one FP32/FP64
//                        value); C = Ah Bh + Ah Bl + Al Bh (the Al Bl term is
//                        below FP32 rounding): ~16 mantissa bits per operand
//   GDT_GEMM_F32         A, B fp32 (v_mfma_f32_32x32x2_f32)
// Block tile 128 x 128, 256 threads = 4 waves, each wave 64 x 64 = 2 x 2
// MFMA 32 x 32 tiles; K staged through LDS in slices of 32 (bf16) or 16
// (fp32), rows padded against bank conflicts.  Shapes must be multiples of
// the tiles (callers pad).  Split K: when the tile grid is small and K long
// (skinny products such as batch x nodes x hidden) and the caller passes a
// workspace, gridDim.z blocks each take a K range and write an FP32 partial
// tile; k_gemm_ksum adds the partials in z order (reproducible).
#include <algorithm>

#include "gdt_internal.h"
#include "kernel_common.h"

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define GM_T 256
#define GM_B 128
#define GM_KB 32      // bf16 K slice
#define GM_KF 16      // fp32 K slice
#define GM_PADB 40    // bf16 row pitch in LDS (elements): 80 B
#define GM_PADF 17    // fp32 row pitch in LDS (elements)

// C layout of a 32 x 32 MFMA tile: lane l holds column l % 32, rows
// (i / 4) * 8 + (l / 32) * 4 + i % 4 for i < 16.
__device__ __forceinline__ void store_tile(float *C, int ldc, int r0, int c0, const f32x16 &acc, bool accumulate)
{
    const int l = threadIdx.x % WAVE;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int r = r0 + (i / 4) * 8 + (l / 32) * 4 + i % 4;
        float *p = C + (size_t)r * ldc + c0 + l % 32;
        *p = accumulate ? *p + acc[i] : acc[i];
    }
}

template <bool SPLIT>
__global__ __launch_bounds__(GM_T) void k_gemm_bf16(int M, int N, int K, const uint16_t *__restrict__ Ah,
                                                     const uint16_t *__restrict__ Al,
                                                     const uint16_t *__restrict__ Bh,
                                                     const uint16_t *__restrict__ Bl, float *__restrict__ C, int ldc,
                                                     int accumulate, int kc, float *__restrict__ part)
{
    constexpr int NP = SPLIT ? 2 : 1;
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
    if (gridDim.z > 1) {  // partial tile of this K range
        C = part + (size_t)blockIdx.z * M * N;
        ldc = N;
        accumulate = 0;
    }
    __shared__ uint16_t sA[NP][GM_B * GM_PADB];
    __shared__ uint16_t sB[NP][GM_B * GM_PADB];
    const int m0 = blockIdx.y * GM_B, n0 = blockIdx.x * GM_B;
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    const int wm = (w / 2) * 64, wn = (w % 2) * 64;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;
    for (int k0 = kb; k0 < ke; k0 += GM_KB) {
        // global -> LDS: 128 rows x 32 k = 512 chunks of 8 bf16 per operand
        const int q0 = threadIdx.x, q1 = threadIdx.x + GM_T;
        const int r0 = q0 / 4, c0 = (q0 % 4) * 8, r1 = q1 / 4, c1 = (q1 % 4) * 8;
        const uint4 ah0 = *(const uint4 *)(Ah + (size_t)(m0 + r0) * K + k0 + c0);
        const uint4 ah1 = *(const uint4 *)(Ah + (size_t)(m0 + r1) * K + k0 + c1);
        const uint4 bh0 = *(const uint4 *)(Bh + (size_t)(n0 + r0) * K + k0 + c0);
        const uint4 bh1 = *(const uint4 *)(Bh + (size_t)(n0 + r1) * K + k0 + c1);
        uint4 al0 = ah0, al1 = ah1, bl0 = bh0, bl1 = bh1;
        if (SPLIT) {
            al0 = *(const uint4 *)(Al + (size_t)(m0 + r0) * K + k0 + c0);
            al1 = *(const uint4 *)(Al + (size_t)(m0 + r1) * K + k0 + c1);
            bl0 = *(const uint4 *)(Bl + (size_t)(n0 + r0) * K + k0 + c0);
            bl1 = *(const uint4 *)(Bl + (size_t)(n0 + r1) * K + k0 + c1);
        }
        __syncthreads();  // the previous slice's LDS reads are done
        *(uint4 *)(&sA[0][r0 * GM_PADB + c0]) = ah0;
        *(uint4 *)(&sA[0][r1 * GM_PADB + c1]) = ah1;
        *(uint4 *)(&sB[0][r0 * GM_PADB + c0]) = bh0;
        *(uint4 *)(&sB[0][r1 * GM_PADB + c1]) = bh1;
        if (SPLIT) {
            *(uint4 *)(&sA[NP - 1][r0 * GM_PADB + c0]) = al0;
            *(uint4 *)(&sA[NP - 1][r1 * GM_PADB + c1]) = al1;
            *(uint4 *)(&sB[NP - 1][r0 * GM_PADB + c0]) = bl0;
            *(uint4 *)(&sB[NP - 1][r1 * GM_PADB + c1]) = bl1;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GM_KB; kk += 16) {
            bf16x8 a[NP][2], b[NP][2];
#pragma unroll
            for (int p = 0; p < NP; p++)
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    a[p][i] = *(const bf16x8 *)(&sA[p][(wm + i * 32 + l % 32) * GM_PADB + kk + (l / 32) * 8]);
                    b[p][i] = *(const bf16x8 *)(&sB[p][(wn + i * 32 + l % 32) * GM_PADB + kk + (l / 32) * 8]);
                }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
                    if (SPLIT) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[NP - 1][j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[NP - 1][i], b[0][j], acc[i][j], 0, 0, 0);
                    }
                }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) store_tile(C, ldc, m0 + wm + i * 32, n0 + wn + j * 32, acc[i][j], accumulate != 0);
}

__global__ __launch_bounds__(GM_T) void k_gemm_f32(int M, int N, int K, const float *__restrict__ A,
                                                    const float *__restrict__ B, float *__restrict__ C, int ldc,
                                                    int accumulate, int kc, float *__restrict__ part)
{
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
    if (gridDim.z > 1) {  // partial tile of this K range
        C = part + (size_t)blockIdx.z * M * N;
        ldc = N;
        accumulate = 0;
    }
    __shared__ float sA[GM_B * GM_PADF];
    __shared__ float sB[GM_B * GM_PADF];
    const int m0 = blockIdx.y * GM_B, n0 = blockIdx.x * GM_B;
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    const int wm = (w / 2) * 64, wn = (w % 2) * 64;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;
    for (int k0 = kb; k0 < ke; k0 += GM_KF) {
        // 128 rows x 16 k = 512 chunks of 4 floats per operand
        float4 ra[2], rb[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int q = threadIdx.x + i * GM_T, row = q / 4, col = (q % 4) * 4;
            ra[i] = *(const float4 *)(A + (size_t)(m0 + row) * K + k0 + col);
            rb[i] = *(const float4 *)(B + (size_t)(n0 + row) * K + k0 + col);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int q = threadIdx.x + i * GM_T, row = q / 4, col = (q % 4) * 4;
            float *pa = &sA[row * GM_PADF + col], *pb = &sB[row * GM_PADF + col];
            pa[0] = ra[i].x;
            pa[1] = ra[i].y;
            pa[2] = ra[i].z;
            pa[3] = ra[i].w;
            pb[0] = rb[i].x;
            pb[1] = rb[i].y;
            pb[2] = rb[i].z;
            pb[3] = rb[i].w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GM_KF; kk += 2) {
            float a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                a[i] = sA[(wm + i * 32 + l % 32) * GM_PADF + kk + l / 32];
                b[i] = sB[(wn + i * 32 + l % 32) * GM_PADF + kk + l / 32];
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) store_tile(C, ldc, m0 + wm + i * 32, n0 + wn + j * 32, acc[i][j], accumulate != 0);
}

// Round-to-nearest-even FP32 -> bf16 (finite inputs).
__device__ __forceinline__ uint16_t bf16_rne(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// FP64 -> bf16 pair (hi = bf16(x), lo = bf16(x - hi)) or FP32.
__global__ void k_split_f64(const double *__restrict__ x, long n, uint16_t *__restrict__ hi, uint16_t *__restrict__ lo,
                            float *__restrict__ f32)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = x[i];
        if (f32) f32[i] = (float)v;
        if (hi) {
            const uint16_t h = bf16_rne((float)v);
            hi[i] = h;
            if (lo) lo[i] = bf16_rne((float)(v - (double)bf16_to_f32(h)));
        }
    }
}

int launch_split_f64(hipStream_t s, const double *x, long n, uint16_t *hi, uint16_t *lo, float *f32)
{
    const long blocks = std::min<long>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_split_f64, dim3((unsigned)std::max<long>(1, blocks)), dim3(256), 0, s, x, n, hi, lo, f32);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// C (+)= sum_z part[z] in z order, elementwise.
__global__ void k_gemm_ksum(int M, int N, int KS, const float *__restrict__ part, float *__restrict__ C, int ldc,
                            int accumulate)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long MN = (long)M * N;
    if (t >= MN) return;
    float q = 0.0f;
    for (int z = 0; z < KS; z++) q += part[z * MN + t];
    float *p = C + (t / N) * ldc + t % N;
    *p = accumulate ? *p + q : q;
}

int gemm_k_splits(int mode, int M, int N, int K)
{
    const int slice = mode == GDT_GEMM_F32 ? GM_KF : GM_KB;
    const int blocks = (M / GM_B) * (N / GM_B), slices = K / slice;
    if (blocks >= 256 || slices < 16) return 1;
    // at least 8 slices per block, about 512 blocks in all
    const int ks = std::min(slices / 8, (512 + blocks - 1) / blocks);
    const int per = (slices + ks - 1) / ks;
    return (slices + per - 1) / per;  // every split non-empty
}

int launch_gemm_nt(hipStream_t s, int mode, int M, int N, int K, const void *Ah, const void *Al, const void *Bh,
                   const void *Bl, float *C, int ldc, int accumulate, float *ws, size_t ws_floats)
{
    GDT_CHECK(M > 0 && N > 0 && K > 0 && M % GM_B == 0 && N % GM_B == 0, GDT_E_ARG,
              "gemm: M, N must be positive multiples of %d (M %d, N %d)", GM_B, M, N);
    const int slice = mode == GDT_GEMM_F32 ? GM_KF : GM_KB;
    GDT_CHECK(K % slice == 0, GDT_E_ARG, "gemm: K %d not a multiple of %d", K, slice);
    int ks = ws ? gemm_k_splits(mode, M, N, K) : 1;
    if ((size_t)ks * M * N > ws_floats) ks = 1;
    const int kc = ((K / slice + ks - 1) / ks) * slice;
    const dim3 grid(N / GM_B, M / GM_B, ks);
    if (mode == GDT_GEMM_F32) {
        hipLaunchKernelGGL(k_gemm_f32, grid, dim3(GM_T), 0, s, M, N, K, (const float *)Ah, (const float *)Bh, C, ldc,
                           accumulate, kc, ws);
    } else if (mode == GDT_GEMM_BF16 || mode == GDT_GEMM_BF16_SPLIT) {
        if (mode == GDT_GEMM_BF16_SPLIT) {
            GDT_CHECK(Al && Bl, GDT_E_ARG, "gemm bf16 split: low parts missing");
            hipLaunchKernelGGL(k_gemm_bf16<true>, grid, dim3(GM_T), 0, s, M, N, K, (const uint16_t *)Ah,
                               (const uint16_t *)Al, (const uint16_t *)Bh, (const uint16_t *)Bl, C, ldc, accumulate,
                               kc, ws);
        } else {
            hipLaunchKernelGGL(k_gemm_bf16<false>, grid, dim3(GM_T), 0, s, M, N, K, (const uint16_t *)Ah,
                               (const uint16_t *)Ah, (const uint16_t *)Bh, (const uint16_t *)Bh, C, ldc, accumulate,
                               kc, ws);
        }
    } else {
        GDT_CHECK(false, GDT_E_ARG, "gemm: unknown mode %d", mode);
    }
    GDT_HIP(hipGetLastError());
    if (ks > 1) {
        const long mn = (long)M * N;
        hipLaunchKernelGGL(k_gemm_ksum, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, M, N, ks,
                           (const float *)ws, C, ldc, accumulate);
        GDT_HIP(hipGetLastError());
    }
    return GDT_OK;
}
