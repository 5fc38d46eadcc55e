// Batched analytic-matrix collision operator on the matrix cores, 16^3 box
// (evolve_direct_colisions / evolve_inverse_colisions, collision_integral.cu:
// 32-188, for many cells at once), FP32 accumulation, FP64 per-cell sums.
//
// Loss: nu[c][k] = sum_j b(|k - j|) f[c][j] -- one GEMM, F (cells x N) times
// the Toeplitz matrix B (N x N) built from the b table.
//
// Gain: with s = v + v1 and g = |v - v1| per axis (same parity as s), the
// reference sum is  gain[v] = sum_v1 G[g, s],
//   G[g, s] = sum_n1 a[n1, g] T[n1, s],
//   T[n1, s] = f(A)f(B) + f(A_x,B_y,B_z)f(B_x,A_y,A_z) + f(A_x,A_y,B_z)f(B_x,B_y,A_z)
//              + f(A_x,B_y,A_z)f(B_x,A_y,B_z),
// A = (s + n)/2, B = (s - n)/2 with n = 2 n1 + parity(s) (term dropped when
// A > 15 or s - n < 0 on an axis, :88-131).  Per parity class of s, G is a
// GEMM: (512 g x 512 n1) times (512 n1 x S_class s).  A block owns one cell
// and one class: f in LDS (FP64), the a operand streamed in K slices, T built
// on the fly for a tile of 128 s (one ix, eight iy rows of 16 iz, padded where
// the class has 15), 8 waves of 64 g x 128 s MFMA tiles.  The tile's G is then
// read back by output: thread t owns the outputs v = t + 512 k (k < 8) and,
// chunk by chunk (two iy rows of the tile staged in LDS), adds G[|2v - s|, s]
// for every column s whose partner v1 = s - v is on the grid (the pairs
// (v, v1) of the reference with v + v1 = s), in FP64 registers and in a fixed
// order: no atomics, bitwise reproducible.  The block writes its class's
// partial gain; the 8 partials are summed in FP64.
//
// Operand precision (mode): bf16, bf16 split into hi + lo (C = Ah Bh + Ah Bl
// + Al Bh), or fp32 MFMA.  The conservation correction and Q = gain - f nu
// run in FP64.
#include "gdt_internal.h"
#include "kernel_common.h"

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define MX_N 16           // grid nodes per axis
#define MX_NB 4096        // nodes
#define MX_G 512          // relative velocities per parity class (rows)
#define MX_K 512          // boxes n1 per class (contraction)
#define MX_S 128          // s per tile (columns)
#define MX_T 512          // threads: 8 waves of 64 g rows
#define MX_KB 32          // bf16 K slice
#define MX_KF 16          // fp32 K slice
#define MX_PB 40          // bf16 LDS row pitch (elements)
static_assert(MX_T == MX_KB * 16, "a bf16 T slice gives each thread one (n1, iz) pair");
#define MX_PF 17          // fp32 LDS row pitch (elements)
#define MX_GP 33          // LDS row pitch of a staged G chunk (floats)

// Bytes of the operand region: the operand tiles of the mode, or a staged G
// chunk (512 rows x 32 columns), whichever is larger.
__host__ __device__ constexpr size_t mx_ops_bytes(int mode)
{
    const size_t opnd = mode == 2 ? sizeof(float) * (MX_G + MX_S) * MX_PF
                                  : sizeof(uint16_t) * (mode == 1 ? 2 : 1) * (MX_G + MX_S) * MX_PB;
    const size_t gch = sizeof(float) * MX_G * MX_GP;
    return opnd > gch ? opnd : gch;
}

__device__ __forceinline__ uint16_t mx_bf16(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float mx_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// T[n1, s] by parts: for a thread's fixed n1 and tile column iz (sx fixed by
// the tile), the x and z halves of A = (s + n)/2 and B = (s - n)/2 and their
// drop tests are computed once; mx_Ty adds the y half of row iy.
struct MxTxz {
    int xa, xb, za, zb, ny;
    bool ok;
};

__device__ __forceinline__ MxTxz mx_txz(int p, int sx, int iz, int n1)
{
    const int px = (p >> 2) & 1, pz = p & 1;
    const int nx = 2 * (n1 >> 6) + px, nz = 2 * (n1 & 7) + pz, sz = 2 * iz + pz;
    const int ax = (sx + nx) >> 1, mx = sx - nx, az = (sz + nz) >> 1, mz = sz - nz;
    MxTxz t;
    t.ok = iz < MX_N - pz && ax <= MX_N - 1 && mx >= 0 && az <= MX_N - 1 && mz >= 0;
    t.xa = ax * MX_N * MX_N;
    t.xb = (mx >> 1) * MX_N * MX_N;
    t.za = az;
    t.zb = mz >> 1;
    t.ny = 2 * ((n1 >> 3) & 7) + ((p >> 1) & 1);
    return t;
}

__device__ __forceinline__ double mx_Ty(const double *fs, const MxTxz &t, int iy, int py)
{
    const int sy = 2 * iy + py, ay = (sy + t.ny) >> 1, my = sy - t.ny;
    if (!t.ok || iy >= MX_N - py || ay > MX_N - 1 || my < 0) return 0.0;
    const int ya = ay * MX_N, yb = (my >> 1) * MX_N;
    return fs[t.xa + ya + t.za] * fs[t.xb + yb + t.zb] + fs[t.xa + yb + t.zb] * fs[t.xb + ya + t.za] +
           fs[t.xa + ya + t.zb] * fs[t.xb + yb + t.za] + fs[t.xa + yb + t.za] * fs[t.xb + ya + t.zb];
}

// MODE: 0 bf16, 1 bf16 split, 2 fp32.  Grid (8 classes, cells); AwT[p][g][n1]
// (hi, lo for split; fp32 in Af for MODE 2); out gpart[c][p][v] (FP32).
template <int MODE>
__global__ __launch_bounds__(MX_T) void k_mx_gain(const uint16_t *__restrict__ Awh, const uint16_t *__restrict__ Awl,
                                                  const float *__restrict__ Awf, const double *__restrict__ f,
                                                  float *__restrict__ gpart)
{
    extern __shared__ __align__(16) unsigned char mx_lds[];
    constexpr int NP = MODE == 1 ? 2 : 1;
    double *fs = (double *)mx_lds;                       // 32 KB
    unsigned char *ops = mx_lds + 8 * MX_NB;             // operand tiles, then a G chunk
    uint16_t *sA = (uint16_t *)ops;                      // [NP][512][MX_PB]
    uint16_t *sB = sA + NP * MX_G * MX_PB;               // [NP][128][MX_PB]
    float *fA = (float *)ops;                            // [512][MX_PF]
    float *fB = fA + MX_G * MX_PF;                       // [128][MX_PF]
    float *Gc = (float *)ops;                            // [512][MX_GP]: 32 columns of G
    const int p = blockIdx.x, c = blockIdx.y;
    const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    const double *fc = f + (size_t)c * MX_NB;
    for (int i = threadIdx.x; i < MX_NB; i += MX_T) fs[i] = fc[i];
    const int px = (p >> 2) & 1, py = (p >> 1) & 1, pz = p & 1;
    const int NT = 2 * (MX_N - px);  // tiles: one ix, eight iy rows each
    const size_t aoff = (size_t)p * MX_G * MX_K;
    // outputs v = threadIdx.x + 512 k: vx = 2 k + w / 4 (the same in a wave),
    // vy = (threadIdx.x / 16) % 16, vz = threadIdx.x % 16
    const int vy = (threadIdx.x >> 4) & 15, vz = threadIdx.x & 15;
    double gacc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) gacc[k] = 0.0;
    // the iz of the outputs' partners v1 = s - v on the grid: sz - vz in
    // [0, 15], i.e. iz in [izlo, izlo + 7] (at most 8 values of step 2); per
    // m the LDS offset of G[gz, iz] in a staged row and whether it is on the
    // grid (off-grid m read a valid slot and add an exact zero, so the eight
    // reads of a row go out together with no branch between them)
    const int izlo = (vz - pz + 1) >> 1;
    int zoff[8];
    bool zon[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int iz = izlo + m, sz = 2 * iz + pz;
        zon[m] = iz < MX_N - pz && sz - vz <= 15;
        zoff[m] = zon[m] ? (abs(2 * vz - sz) >> 1) * MX_GP + iz : 0;
    }
    for (int t = 0; t < NT; t++) {
        // tile t: sx = 2 (t / 2) + px; column c: iy = 8 (t % 2) + c / 16,
        // iz = c % 16 (off the class where i_a >= 16 - p_a: T = 0)
        const int sx_t = 2 * (t >> 1) + px;
        // On the x axis both n (the K index n1 = 64 (n_x >> 1) + ...) and the
        // G rows read back (row block = wave w = |g_x| >> 1) are bounded by
        // min(sx, 30 - sx): T[n1, s] is zero for every column of the tile past
        // n_x = min(sx, 30 - sx) (the drop rule A_x <= 15, s_x - n_x >= 0), and
        // no pair (v, v1) on the grid has |g_x| beyond it.  So the K loop stops
        // at kx + 1 blocks of 64 n1, and waves past kx neither load their a
        // rows nor run MFMAs (about half the work; the slices skipped would
        // add exact zeros, and the rows skipped are never read).
        const int kx = (min(sx_t, 2 * (MX_N - 1) - sx_t) - px) >> 1;
        const int kend = 64 * (kx + 1);
        const bool live = w <= kx;
        f32x16 acc[2][4];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[i][j][e] = 0.0f;
        constexpr int KB = MODE == 2 ? MX_KF : MX_KB;
        for (int k0 = 0; k0 < kend; k0 += KB) {
            __syncthreads();  // previous slice's operand reads are done
            if (MODE != 2) {
                // a slice: 512 rows x 32 n1 = 2048 chunks of 8 bf16 per part
#pragma unroll
                for (int pp = 0; pp < NP; pp++) {
                    const uint16_t *src = pp ? Awl : Awh;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int q = threadIdx.x + i * MX_T, row = q / 4, col = (q % 4) * 8;
                        if (row < kend)
                            *(uint4 *)(&sA[(pp * MX_G + row) * MX_PB + col]) =
                                *(const uint4 *)(src + aoff + (size_t)row * MX_K + k0 + col);
                    }
                }
                // T slice: 128 s x 32 n1, 8 per thread.  Thread q holds n1 =
                // k0 + q / 16 and the column iz = q % 16 of the tile's eight iy
                // rows: n1, sx and sz (the x and z halves of A = (s + n) / 2
                // and B = (s - n) / 2) are fixed; row by row sy grows by 2, so
                // Ay and By step by one.  Consecutive lanes take consecutive iz,
                // so their reads of f are consecutive in z (0.8% faster than n1
                // fastest, the same bits).
                {
                    const int nk = threadIdx.x / MX_N, iz = threadIdx.x % MX_N;
                    const MxTxz tx = mx_txz(p, sx_t, iz, k0 + nk);
                    const int iy0 = 8 * (t & 1);
#pragma unroll 1
                    for (int i = 0; i < 8; i++) {
                        const double tv = mx_Ty(fs, tx, iy0 + i, py);
                        const int sc = 16 * i + iz;
                        const uint16_t h = mx_bf16((float)tv);
                        sB[sc * MX_PB + nk] = h;
                        if (NP == 2) sB[(MX_S + sc) * MX_PB + nk] = mx_bf16((float)(tv - (double)mx_f32(h)));
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int q = threadIdx.x + i * MX_T, row = q / 4, col = (q % 4) * 4;
                    if (row >= kend) continue;
                    const float4 v = *(const float4 *)(Awf + aoff + (size_t)row * MX_K + k0 + col);
                    float *d = &fA[row * MX_PF + col];
                    d[0] = v.x;
                    d[1] = v.y;
                    d[2] = v.z;
                    d[3] = v.w;
                }
                // 128 s x 16 n1, 4 per thread: n1 = k0 + q % 16 and the column
                // iz = (q / 16) % 16 fixed, rows iy0 + q / 256 + 2 i
                const MxTxz tx = mx_txz(p, sx_t, (threadIdx.x >> 4) & 15, k0 + (threadIdx.x & 15));
#pragma unroll 1
                for (int i = 0; i < 4; i++) {
                    const int sc = (threadIdx.x >> 4) + 32 * i, nk = threadIdx.x & 15;
                    const double tv = mx_Ty(fs, tx, 8 * (t & 1) + (threadIdx.x >> 8) + 2 * i, py);
                    fB[sc * MX_PF + nk] = (float)tv;
                }
            }
            __syncthreads();
            if (!live) continue;
            if (MODE != 2) {
#pragma unroll
                for (int kk = 0; kk < MX_KB; kk += 16) {
                    bf16x8 a[NP][2], b[NP][4];
#pragma unroll
                    for (int pp = 0; pp < NP; pp++) {
#pragma unroll
                        for (int i = 0; i < 2; i++)
                            a[pp][i] = *(const bf16x8 *)(&sA[(pp * MX_G + w * 64 + i * 32 + l % 32) * MX_PB + kk +
                                                             (l / 32) * 8]);
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            b[pp][j] = *(const bf16x8 *)(&sB[(pp * MX_S + j * 32 + l % 32) * MX_PB + kk + (l / 32) * 8]);
                    }
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
                            if (NP == 2) {
                                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[NP - 1][j], acc[i][j],
                                                                                   0, 0, 0);
                                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[NP - 1][i], b[0][j], acc[i][j],
                                                                                   0, 0, 0);
                            }
                        }
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < MX_KF; kk += 2) {
                    float a[2], b[4];
#pragma unroll
                    for (int i = 0; i < 2; i++) a[i] = fA[(w * 64 + i * 32 + l % 32) * MX_PF + kk + l / 32];
#pragma unroll
                    for (int j = 0; j < 4; j++) b[j] = fB[(j * 32 + l % 32) * MX_PF + kk + l / 32];
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
                }
            }
        }
        // read the tile's G back by output, 32 columns (two iy rows) at a
        // time: stage the chunk in LDS (operand space, free after the K loop)
        const int sx = 2 * (t >> 1) + px;
#pragma unroll
        for (int j = 0; j < 4; j++) {  // unrolled: acc[i][j] stays in registers
            __syncthreads();  // operand reads (j = 0) / the previous chunk's reads are done
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int e = 0; e < 16; e++)
                    Gc[(w * 64 + i * 32 + (e / 4) * 8 + (l / 32) * 4 + e % 4) * MX_GP + l % 32] = acc[i][j][e];
            __syncthreads();
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int iy = 8 * (t & 1) + 2 * j + a;
                const int sy = 2 * iy + py;
                // the row is on the grid and holds partners of this lane's vy
                const bool yon = iy < MX_N - py && (unsigned)(sy - vy) <= 15u;
                const int gy = yon ? abs(2 * vy - sy) >> 1 : 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int vx = 2 * k + (w >> 2);  // the same in the wave
                    if ((unsigned)(sx - vx) > 15u) continue;
                    const int base = (64 * (abs(2 * vx - sx) >> 1) + 8 * gy) * MX_GP + 16 * a;
                    float gv[8];
#pragma unroll
                    for (int m = 0; m < 8; m++) gv[m] = Gc[base + zoff[m]];
                    double sum = gacc[k];
#pragma unroll
                    for (int m = 0; m < 8; m++) sum += (yon && zon[m]) ? (double)gv[m] : 0.0;
                    gacc[k] = sum;
                }
            }
        }
    }
    float *out = gpart + ((size_t)c * 8 + p) * MX_NB;
#pragma unroll
    for (int k = 0; k < 8; k++) out[threadIdx.x + MX_T * k] = (float)gacc[k];
}

// Operands from the FP64 tables: B[k][j] = b(|k - j|) (hi, lo, fp32) and
// AwT[p][g][n1] = a[in1, gi, jn1, kn1, gj, gk] with g of parity p.
__global__ void k_mx_bmat(const double *__restrict__ b, uint16_t *__restrict__ Bh, uint16_t *__restrict__ Bl,
                          float *__restrict__ Bf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)MX_NB * MX_NB) return;
    const int k = (int)(t / MX_NB), j = (int)(t % MX_NB);
    const int gi = abs(k / 256 - j / 256), gj = abs((k / 16) % 16 - (j / 16) % 16), gk = abs(k % 16 - j % 16);
    const double v = b[(gi * MX_N + gj) * MX_N + gk];
    const uint16_t h = mx_bf16((float)v);
    Bh[t] = h;
    Bl[t] = mx_bf16((float)(v - (double)mx_f32(h)));
    Bf[t] = (float)v;
}

__global__ void k_mx_awt(const double *__restrict__ a, uint16_t *__restrict__ Ah, uint16_t *__restrict__ Al,
                         float *__restrict__ Af)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * MX_G * MX_K) return;
    const int n1 = t % MX_K, gr = (t / MX_K) % MX_G, p = t / (MX_K * MX_G);
    const int gi = 2 * (gr >> 6) + ((p >> 2) & 1), gj = 2 * ((gr >> 3) & 7) + ((p >> 1) & 1),
              gk = 2 * (gr & 7) + (p & 1);
    const int in1 = n1 >> 6, jn1 = (n1 >> 3) & 7, kn1 = n1 & 7;
    const int hv = MX_N / 2, hw = MX_N / 2;
    const long ai = ((long)((in1 * MX_N + gi) * (hv * hw) + jn1 * hw + kn1)) * (MX_N * MX_N) + gj * MX_N + gk;
    const double v = a[ai];
    const uint16_t h = mx_bf16((float)v);
    Ah[t] = h;
    Al[t] = mx_bf16((float)(v - (double)mx_f32(h)));
    Af[t] = (float)v;
}

// FP64 sums: nu from the FP32 loss GEMM, gain from the 8 class partials.
__global__ void k_mx_combine(long count, const float *__restrict__ nuf, int ldn, const float *__restrict__ gpart,
                             double *__restrict__ nu, double *__restrict__ gain)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const long c = t / MX_NB, v = t % MX_NB;
    nu[t] = (double)nuf[c * ldn + v];
    double g = 0.0;
#pragma unroll
    for (int p = 0; p < 8; p++) g += (double)gpart[(c * 8 + p) * MX_NB + v];
    gain[t] = g;
}

int launch_mx_operands(hipStream_t s, const double *b, const double *a, uint16_t *Bh, uint16_t *Bl, float *Bf,
                       uint16_t *Ah, uint16_t *Al, float *Af)
{
    const long nb = (long)MX_NB * MX_NB;
    hipLaunchKernelGGL(k_mx_bmat, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, b, Bh, Bl, Bf);
    GDT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mx_awt, dim3((8 * MX_G * MX_K + 255) / 256), dim3(256), 0, s, a, Ah, Al, Af);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// f (FP64) and the operand region.
static size_t mx_lds_bytes(int mode)
{
    const int m = mode == GDT_GEMM_F32 ? 2 : (mode == GDT_GEMM_BF16_SPLIT ? 1 : 0);
    return 8 * MX_NB + mx_ops_bytes(m);
}

int launch_mx_gain(hipStream_t s, int mode, const uint16_t *Ah, const uint16_t *Al, const float *Af, const double *f,
                   int cells, float *gpart)
{
    static bool attr_set = false;
    if (!attr_set) {
        GDT_HIP(hipFuncSetAttribute((const void *)k_mx_gain<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)mx_lds_bytes(GDT_GEMM_BF16)));
        GDT_HIP(hipFuncSetAttribute((const void *)k_mx_gain<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)mx_lds_bytes(GDT_GEMM_BF16_SPLIT)));
        GDT_HIP(hipFuncSetAttribute((const void *)k_mx_gain<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)mx_lds_bytes(GDT_GEMM_F32)));
        attr_set = true;
    }
    const dim3 grid(8, cells);
    if (mode == GDT_GEMM_BF16)
        hipLaunchKernelGGL(k_mx_gain<0>, grid, dim3(MX_T), mx_lds_bytes(mode), s, Ah, Ah, Af, f, gpart);
    else if (mode == GDT_GEMM_BF16_SPLIT)
        hipLaunchKernelGGL(k_mx_gain<1>, grid, dim3(MX_T), mx_lds_bytes(mode), s, Ah, Al, Af, f, gpart);
    else
        hipLaunchKernelGGL(k_mx_gain<2>, grid, dim3(MX_T), mx_lds_bytes(mode), s, Ah, Al, Af, f, gpart);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_mx_combine(hipStream_t s, int cells, const float *nuf, int ldn, const float *gpart, double *nu,
                      double *gain)
{
    const long count = (long)cells * MX_NB;
    hipLaunchKernelGGL(k_mx_combine, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, count, nuf, ldn, gpart,
                       nu, gain);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}
