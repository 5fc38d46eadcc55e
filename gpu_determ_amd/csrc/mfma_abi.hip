// C ABI of the matrix-core paths (declarations and contracts: include/gdt.h).
#include <algorithm>

#include "gdt_internal.h"

int gdt_gemm_nt(int mode, int M, int N, int K, const void *A_hi, const void *A_lo, const void *B_hi,
                const void *B_lo, float *C, int ldc, int accumulate)
{
    GDT_CHECK(A_hi && B_hi && C && ldc >= N, GDT_E_ARG, "gdt_gemm_nt: bad arguments");
    // skinny shapes split K over a temporary workspace (this entry is
    // synchronous, so it is freed on return)
    DevBuf ws;
    size_t wf = 0;
    if (M > 0 && N > 0 && M % 128 == 0 && N % 128 == 0 && K > 0) {
        const int ks = gemm_k_splits(mode, M, N, K);
        if (ks > 1 && ws.ensure(sizeof(float) * (size_t)ks * M * N) == 0) wf = (size_t)ks * M * N;
    }
    int rc = launch_gemm_nt(nullptr, mode, M, N, K, A_hi, A_lo, B_hi, B_lo, C, ldc, accumulate, ws.as<float>(), wf);
    if (rc == GDT_OK && hipDeviceSynchronize() != hipSuccess) {
        gdt_set_error(GDT_E_HIP, "gdt_gemm_nt: %s", hipGetErrorString(hipGetLastError()));
        rc = GDT_E_HIP;
    }
    ws.release();
    return rc;
}

int gdt_split_f64(const double *x, long n, uint16_t *hi, uint16_t *lo, float *f32)
{
    GDT_CHECK(x && n >= 0, GDT_E_ARG, "gdt_split_f64: bad arguments");
    if (n == 0) return GDT_OK;
    int rc = launch_split_f64(nullptr, x, n, hi, lo, f32);
    if (rc) return rc;
    GDT_HIP(hipDeviceSynchronize());
    return GDT_OK;
}

int gdt_matrix_mfma(gdt_handle h, int mode, const double *f, double *nu, double *gain, double *Q, int cells,
                    unsigned flags)
{
    GDT_CHECK(h, GDT_E_ARG, "null gdt handle");
    GDT_CHECK(f && cells > 0, GDT_E_ARG, "gdt_matrix_mfma: bad arguments");
    GDT_CHECK(mode == GDT_GEMM_BF16 || mode == GDT_GEMM_BF16_SPLIT || mode == GDT_GEMM_F32, GDT_E_ARG,
              "gdt_matrix_mfma: mode %d", mode);
    GDT_CHECK(h->n[0] == 16 && h->n[1] == 16 && h->n[2] == 16 && h->n_pnt == 4096, GDT_E_ARG,
              "gdt_matrix_mfma: needs the whole 16^3 box (grid %dx%dx%d, %d nodes kept)", h->n[0], h->n[1], h->n[2],
              h->n_pnt);
    GDT_CHECK(h->have_tables, GDT_E_STATE, "gdt_matrix_mfma: tables not built");
    const int N = 4096;
    if (!h->mx_ready) {
        const size_t nb = (size_t)N * N, na = (size_t)8 * 512 * 512;
        GDT_CHECK(h->mx_bh.ensure(2 * nb) == 0 && h->mx_bl.ensure(2 * nb) == 0 && h->mx_bf.ensure(4 * nb) == 0 &&
                      h->mx_ah.ensure(2 * na) == 0 && h->mx_al.ensure(2 * na) == 0 && h->mx_af.ensure(4 * na) == 0,
                  GDT_E_NOMEM, "gdt_matrix_mfma: operand allocation failed");
        int rc = launch_mx_operands(h->stream, h->tb.as<double>(), h->ta.as<double>(), h->mx_bh.as<uint16_t>(),
                                    h->mx_bl.as<uint16_t>(), h->mx_bf.as<float>(), h->mx_ah.as<uint16_t>(),
                                    h->mx_al.as<uint16_t>(), h->mx_af.as<float>());
        if (rc) return rc;
        h->mx_ready = true;
    }
    // scratch: F (bf16 hi, lo, fp32; rows padded to 128), loss (fp32), gain
    // class partials (fp32), nu and gain (FP64)
    const int cpad = (cells + 127) / 128 * 128;
    const size_t nF = (size_t)cpad * N;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t oFh = take(2 * nF), oFl = take(2 * nF), oFf = take(4 * nF), oNu = take(4 * nF),
                 oGp = take(4 * (size_t)cells * 8 * N), oN64 = take(8 * (size_t)cells * N),
                 oG64 = take(8 * (size_t)cells * N);
    GDT_CHECK(h->s_mx.ensure(off) == 0, GDT_E_NOMEM, "gdt_matrix_mfma: scratch allocation failed (%zu bytes)", off);
    char *base = h->s_mx.as<char>();
    uint16_t *Fh = (uint16_t *)(base + oFh), *Fl = (uint16_t *)(base + oFl);
    float *Ff = (float *)(base + oFf), *nuf = (float *)(base + oNu), *gp = (float *)(base + oGp);
    double *nu64 = (double *)(base + oN64), *g64 = (double *)(base + oG64);
    if (cpad > cells) {
        const size_t pad = (size_t)(cpad - cells) * N;
        GDT_HIP(hipMemsetAsync(Fh + (size_t)cells * N, 0, 2 * pad, h->stream));
        GDT_HIP(hipMemsetAsync(Fl + (size_t)cells * N, 0, 2 * pad, h->stream));
        GDT_HIP(hipMemsetAsync(Ff + (size_t)cells * N, 0, 4 * pad, h->stream));
    }
    int rc;
    if ((rc = launch_split_f64(h->stream, f, (long)cells * N, Fh, Fl, Ff))) return rc;
    {
        KernelTimer kt(h, "mx_loss_gemm");
        if (mode == GDT_GEMM_F32)
            rc = launch_gemm_nt(h->stream, mode, cpad, N, N, Ff, nullptr, h->mx_bf.p, nullptr, nuf, N, 0);
        else
            rc = launch_gemm_nt(h->stream, mode, cpad, N, N, Fh, Fl, h->mx_bh.p, h->mx_bl.p, nuf, N, 0);
        if (rc) return rc;
    }
    {
        KernelTimer kt(h, "mx_gain");
        if ((rc = launch_mx_gain(h->stream, mode, h->mx_ah.as<uint16_t>(), h->mx_al.as<uint16_t>(),
                                 h->mx_af.as<float>(), f, cells, gp)))
            return rc;
    }
    if ((rc = launch_mx_combine(h->stream, cells, nuf, N, gp, nu64, g64))) return rc;
    if (gain) GDT_HIP(hipMemcpyAsync(gain, g64, 8 * (size_t)cells * N, hipMemcpyDeviceToDevice, h->stream));
    double *Qo = Q ? Q : h->s_Q.ensure(8 * (size_t)cells * N) == 0 ? h->s_Q.as<double>() : nullptr;
    GDT_CHECK(Qo, GDT_E_NOMEM, "gdt_matrix_mfma: scratch allocation failed");
    return launch_correct(h, (flags & GDT_FLAG_CORRECT) ? GDT_METHOD_MATRIX : -1, f, nu64, g64, nu ? nu : nu64, Qo,
                          nullptr, cells, 0.0, nullptr);
}
