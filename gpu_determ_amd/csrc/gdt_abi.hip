// C ABI of libgdt.so (declarations and contracts: include/gdt.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gdt_internal.h"
#include "kmath.h"

static thread_local std::string g_last_error;

void gdt_set_error(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = std::string(buf) + " [status " + std::to_string(code) + "]";
}

static std::atomic<unsigned long> g_devbuf_gen{0};

unsigned long devbuf_generation() { return g_devbuf_gen.load(std::memory_order_acquire); }

int DevBuf::ensure(size_t need)
{
    if (need <= bytes && p) return 0;
    release();
    if (need == 0) need = 8;
    if (hipMalloc(&p, need) != hipSuccess) {
        p = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        return -1;
    }
    bytes = need;
    return 0;
}

void DevBuf::release()
{
    if (p) {
        (void)hipFree(p);
        g_devbuf_gen.fetch_add(1, std::memory_order_acq_rel);
    }
    p = nullptr;
    bytes = 0;
}

int timing_event(gdt_context *h)
{
    if (h->ev_used == (int)h->ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
        h->ev_pool.push_back(e);
    }
    return h->ev_used++;
}

int timing_resolve(gdt_context *h)
{
    if (!h->pending.empty()) {
        GDT_HIP(hipEventSynchronize(h->ev_pool[h->pending.back().b]));
        for (const auto &p : h->pending) {
            float ms = 0.0f;
            GDT_HIP(hipEventElapsedTime(&ms, h->ev_pool[p.a], h->ev_pool[p.b]));
            KernelStat &st = h->stats[p.name];
            st.ms += ms;
            st.count += 1;
        }
        h->pending.clear();
    }
    h->ev_used = 0;
    return GDT_OK;
}

ExtTimer::ExtTimer(gdt_context *h_, const char *name_) : h(h_), name(name_)
{
    if (!h->timing) return;
    if (h->ev_used + 2 > 4096) timing_resolve(h);  // bounded pool: resolve, then reuse
    a = timing_event(h);
    b = a >= 0 ? timing_event(h) : -1;
    if (a >= 0 && b >= 0) {
        start = h->ev_pool[a];
        stop = h->ev_pool[b];
    }
}

ExtTimer::~ExtTimer()
{
    if (start) h->pending.push_back({name, a, b});
}

KernelTimer::KernelTimer(gdt_context *h_, const char *name_) : h(h_), name(name_)
{
    if (!h->timing) return;
    if (h->ev_used + 2 > 4096) timing_resolve(h);  // bounded pool: resolve, then reuse
    a = timing_event(h);
    if (a >= 0) (void)hipEventRecord(h->ev_pool[a], h->stream);
}

KernelTimer::~KernelTimer()
{
    if (!h->timing || a < 0) return;
    const int b = timing_event(h);
    if (b < 0) return;
    (void)hipEventRecord(h->ev_pool[b], h->stream);
    h->pending.push_back({name, a, b});
}

#define H_CHECK(h)                                                   \
    do {                                                             \
        if (!(h)) {                                                  \
            gdt_set_error(GDT_E_ARG, "null gdt handle");             \
            return GDT_E_ARG;                                        \
        }                                                            \
        GDT_HIP(hipSetDevice((h)->device));                          \
    } while (0)

// Tables of kmt_log / kmt_exp, in 80-bit long double arithmetic (64-bit
// significand): every double entry and high + low pair is correct to well
// below one FP64 ulp.
void kmath_tables(double *T)
{
    // high parts with the low 20 significand bits cleared (33 bits), so that
    // their products with exponents and multiples of ln2/64 are exact
    auto high33 = [](long double v) {
        double d = (double)v;
        uint64_t u;
        std::memcpy(&u, &d, 8);
        u &= ~((uint64_t(1) << 20) - 1);
        std::memcpy(&d, &u, 8);
        return d;
    };
    for (int j = 0; j < KMT_LOG_N; j++) {
        const double c = (double)(1.0L / (1.0L + ((long double)j + 0.5L) / 128.0L));
        const long double v = -logl((long double)c);
        T[KMT_C + j] = c;
        T[KMT_LHI + j] = (double)v;
        T[KMT_LLO + j] = (double)(v - (long double)(double)v);
    }
    for (int j = 0; j < KMT_EXP_N; j++) {
        const long double v = exp2l((long double)j / 64.0L);
        T[KMT_SHI + j] = (double)v;
        T[KMT_SLO + j] = (double)(v - (long double)(double)v);
    }
    const long double ln2 = logl(2.0L);
    T[KMT_K] = high33(ln2);
    T[KMT_K + 1] = (double)(ln2 - (long double)T[KMT_K]);
    T[KMT_K + 2] = (double)(64.0L / ln2);
    T[KMT_K + 3] = high33(ln2 / 64.0L);
    T[KMT_K + 4] = (double)(ln2 / 64.0L - (long double)T[KMT_K + 3]);
    for (int j = KMT_K + 5; j < KMT_SIZE; j++) T[j] = 0.0;
}

extern "C" {

const char *gdt_last_error(void) { return g_last_error.c_str(); }

int gdt_version(void) { return 10000; }

int gdt_device_count(int *count)
{
    GDT_CHECK(count, GDT_E_ARG, "count is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *count = c;
    return GDT_OK;
}

int gdt_malloc(void **dptr, size_t bytes)
{
    GDT_CHECK(dptr, GDT_E_ARG, "dptr is null");
    GDT_HIP(hipMalloc(dptr, bytes ? bytes : 8));
    return GDT_OK;
}

int gdt_free(void *dptr)
{
    if (dptr) GDT_HIP(hipFree(dptr));
    return GDT_OK;
}

int gdt_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    GDT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GDT_OK;
}

int gdt_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    GDT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GDT_OK;
}

int gdt_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
    GDT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return GDT_OK;
}

int gdt_memset(void *dptr, int value, size_t bytes)
{
    GDT_HIP(hipMemset(dptr, value, bytes));
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

int gdt_create(gdt_handle *out, const gdt_grid_desc *gd, int device)
{
    GDT_CHECK(out && gd, GDT_E_ARG, "gdt_create: null argument");
    *out = nullptr;
    for (int a = 0; a < 3; a++) {
        GDT_CHECK(gd->n[a] > 0 && gd->n[a] <= 128, GDT_E_ARG, "gdt_create: n[%d] = %d out of range 1..128", a,
                  gd->n[a]);
        GDT_CHECK(gd->vmin[a] < gd->vmax[a], GDT_E_ARG, "gdt_create: vmin[%d] >= vmax[%d]", a, a);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        gdt_set_error(GDT_E_NODEVICE, "gdt_create: no HIP device visible");
        return GDT_E_NODEVICE;
    }
    GDT_CHECK(device >= 0 && device < ndev, GDT_E_ARG, "gdt_create: device %d of %d", device, ndev);
    GDT_HIP(hipSetDevice(device));
    {
        const int rc0 = noalloc_setup(device);  // scratch address of gdt_evaluate_noalloc on this device
        if (rc0 != GDT_OK) return rc0;
    }

    gdt_context *h = new gdt_context();
    h->device = device;
    for (int a = 0; a < 3; a++) {
        h->n[a] = gd->n[a];
        h->vmin[a] = gd->vmin[a];
        h->vmax[a] = gd->vmax[a];
        h->d[a] = (gd->vmax[a] - gd->vmin[a]) / gd->n[a];
    }
    h->R = gd->R;
    h->d3v = h->d[0] * h->d[1] * h->d[2];
    const int n0 = h->n[0], n1 = h->n[1], n2 = h->n[2];
    h->n_box = n0 * n1 * n2;
    // sphere cut (velocity_grid.cu:6-14: centre from the box middle) and the
    // compact node order (velocity_grid.cu:114-126)
    h->h_box2c.assign(h->n_box, -1);
    int cnt = 0, idx = 0;
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++)
            for (int k = 0; k < n2; k++, idx++) {
                const double x = (gd->vmax[0] - gd->vmin[0]) * (2 * i + 1 - n0) / 2.0 / n0;
                const double y = (gd->vmax[1] - gd->vmin[1]) * (2 * j + 1 - n1) / 2.0 / n1;
                const double z = (gd->vmax[2] - gd->vmin[2]) * (2 * k + 1 - n2) / 2.0 / n2;
                const bool keep = (gd->R <= 0.0) || (x * x + y * y + z * z < gd->R * gd->R);
                if (keep) {
                    h->h_box2c[idx] = cnt++;
                    h->h_cidx.push_back(i);
                    h->h_cidx.push_back(j);
                    h->h_cidx.push_back(k);
                }
            }
    h->n_pnt = cnt;
    if (cnt == 0) {
        delete h;
        gdt_set_error(GDT_E_ARG, "gdt_create: the sphere cut keeps no node");
        return GDT_E_ARG;
    }
    std::vector<int> cmp2box(cnt);
    h->h_xi.resize(3 * (size_t)cnt);
    for (int v = 0; v < cnt; v++) {
        const int i = h->h_cidx[3 * v], j = h->h_cidx[3 * v + 1], k = h->h_cidx[3 * v + 2];
        cmp2box[v] = (i * n1 + j) * n2 + k;
        h->h_xi[3 * v] = gd->vmin[0] + h->d[0] * (2.0 * i + 1.0) / 2.0;
        h->h_xi[3 * v + 1] = gd->vmin[1] + h->d[1] * (2.0 * j + 1.0) / 2.0;
        h->h_xi[3 * v + 2] = gd->vmin[2] + h->d[2] * (2.0 * k + 1.0) / 2.0;
    }
    std::vector<double> kmt(KMT_SIZE);
    kmath_tables(kmt.data());
    int rc = GDT_OK;
    if (h->box2c.ensure(sizeof(int) * h->n_box) || h->cidx.ensure(sizeof(int) * 3 * cnt) ||
        h->xi.ensure(sizeof(double) * 3 * cnt) || h->cmp2box.ensure(sizeof(int) * cnt) ||
        h->kmath.ensure(sizeof(double) * KMT_SIZE)) {
        gdt_set_error(GDT_E_NOMEM, "gdt_create: device allocation failed");
        rc = GDT_E_NOMEM;
    }
    if (rc == GDT_OK && hipStreamCreate(&h->stream) != hipSuccess) {
        gdt_set_error(GDT_E_HIP, "gdt_create: stream creation failed");
        rc = GDT_E_HIP;
    }
    if (rc == GDT_OK &&
        (hipMemcpy(h->box2c.p, h->h_box2c.data(), sizeof(int) * h->n_box, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(h->cidx.p, h->h_cidx.data(), sizeof(int) * 3 * cnt, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(h->xi.p, h->h_xi.data(), sizeof(double) * 3 * cnt, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(h->cmp2box.p, cmp2box.data(), sizeof(int) * cnt, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(h->kmath.p, kmt.data(), sizeof(double) * KMT_SIZE, hipMemcpyHostToDevice) != hipSuccess)) {
        gdt_set_error(GDT_E_HIP, "gdt_create: upload failed");
        rc = GDT_E_HIP;
    }
    if (rc != GDT_OK) {
        gdt_destroy(h);
        return rc;
    }
    *out = h;
    return GDT_OK;
}

int gdt_destroy(gdt_handle h)
{
    if (!h) return GDT_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    DevBuf *bufs[] = {&h->box2c, &h->cidx,   &h->xi,     &h->cmp2box, &h->tb,     &h->ta,    &h->gptr,  &h->gbox,
                      &h->gval,  &h->kidx,   &h->kr,     &h->kw,      &h->knode,  &h->kptr,  &h->kent,  &h->kn4,
                      &h->ke4,    &h->ke1,    &h->kwb,     &h->kgptr,  &h->kgent, &h->kcbase, &h->ktile, &h->kmath,
                      &h->mx_bh, &h->mx_bl,  &h->mx_bf,  &h->mx_ah,   &h->mx_al,  &h->mx_af, &h->s_mx,  &h->s_fbox,
                      &h->s_out1, &h->s_out2, &h->s_vals, &h->s_loss, &h->s_Q,   &h->s_nu,  &h->s_tmp, &h->s_dt,
                      &h->s_mpart, &h->s_part};
    for (DevBuf *b : bufs) b->release();
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return GDT_OK;
}

int gdt_synchronize(gdt_handle h)
{
    H_CHECK(h);
    GDT_HIP(hipStreamSynchronize(h->stream));
    return GDT_OK;
}

int gdt_grid_info(gdt_handle h, int *n_pnt, int *n_box, double d[3], double *d3v)
{
    H_CHECK(h);
    if (n_pnt) *n_pnt = h->n_pnt;
    if (n_box) *n_box = h->n_box;
    if (d)
        for (int a = 0; a < 3; a++) d[a] = h->d[a];
    if (d3v) *d3v = h->d3v;
    return GDT_OK;
}

int gdt_grid_nodes(gdt_handle h, int *cidx, double *xi)
{
    H_CHECK(h);
    if (cidx) std::memcpy(cidx, h->h_cidx.data(), sizeof(int) * 3 * h->n_pnt);
    if (xi) std::memcpy(xi, h->h_xi.data(), sizeof(double) * 3 * h->n_pnt);
    return GDT_OK;
}

int gdt_basis(gdt_handle h, double *dst)
{
    H_CHECK(h);
    GDT_CHECK(dst, GDT_E_ARG, "gdt_basis: dst is null");
    return launch_basis(h, dst);
}

// ---------------------------------------------------------------------------
// analytic matrix tables
// ---------------------------------------------------------------------------

int gdt_build_tables(gdt_handle h, int mode)
{
    H_CHECK(h);
    GDT_CHECK(mode == GDT_TABLES_FP64 || mode == GDT_TABLES_REFERENCE, GDT_E_ARG, "gdt_build_tables: mode %d", mode);
    for (int a = 0; a < 3; a++)
        GDT_CHECK(h->n[a] % 2 == 0, GDT_E_ARG, "gdt_build_tables: n[%d] = %d must be even", a, h->n[a]);
    GDT_CHECK(h->n_box <= 32768, GDT_E_ARG, "gdt_build_tables: %d box nodes (max 32768)", h->n_box);
    for (int a = 0; a < 3; a++)
        GDT_CHECK(h->n[a] <= GLIST_AXIS_MAX, GDT_E_ARG, "gdt_build_tables: n[%d] = %d (max %d)", a, h->n[a],
                  GLIST_AXIS_MAX);
    h->mx_ready = false;  // the matrix-core operands are built from the tables
    int rc = launch_build_tables(h, mode);
    if (rc != GDT_OK) return rc;
    if ((rc = gain_prepare(h)) != GDT_OK) return rc;
    h->table_mode = mode;
    h->have_tables = true;
    return GDT_OK;
}

int gdt_tables(gdt_handle h, const double **b, const double **a, long *n_a, long *nnz)
{
    H_CHECK(h);
    GDT_CHECK(h->have_tables, GDT_E_STATE, "gdt_tables: tables not built");
    if (b) *b = h->tb.as<double>();
    if (a) *a = h->ta.as<double>();
    if (n_a) *n_a = h->n_a;
    if (nnz) *nnz = h->n_gent;
    return GDT_OK;
}

// Rebuilds the nonzero-box lists from caller tables (FP64, device).
int gdt_set_tables(gdt_handle h, const double *b, const double *a)
{
    H_CHECK(h);
    GDT_CHECK(b && a, GDT_E_ARG, "gdt_set_tables: null table");
    for (int ax = 0; ax < 3; ax++)
        GDT_CHECK(h->n[ax] % 2 == 0 && h->n[ax] <= GLIST_AXIS_MAX, GDT_E_ARG,
                  "gdt_set_tables: n[%d] = %d must be even and at most %d", ax, h->n[ax], GLIST_AXIS_MAX);
    const int nb = h->n_box;
    const long n_a = (long)nb * nb / 8;
    h->mx_ready = false;  // the matrix-core operands are built from the tables
    GDT_CHECK(h->tb.ensure(sizeof(double) * nb) == 0 && h->ta.ensure(sizeof(double) * n_a) == 0, GDT_E_NOMEM,
              "gdt_set_tables: allocation failed");
    GDT_HIP(hipMemcpyAsync(h->tb.p, b, sizeof(double) * nb, hipMemcpyDeviceToDevice, h->stream));
    GDT_HIP(hipMemcpyAsync(h->ta.p, a, sizeof(double) * n_a, hipMemcpyDeviceToDevice, h->stream));
    // derive the lists on the host (import path; not on the hot path)
    std::vector<double> ha(n_a);
    GDT_HIP(hipMemcpyAsync(ha.data(), a, sizeof(double) * n_a, hipMemcpyDeviceToHost, h->stream));
    GDT_HIP(hipStreamSynchronize(h->stream));
    const int nu = h->n[0], nv = h->n[1], nw = h->n[2], nyz = nv * nw, hu = nu / 2, hv = nv / 2, hw = nw / 2;
    std::vector<int> gp(nb + 1), gb;
    std::vector<double> gv;
    for (int g = 0; g < nb; g++) {
        gp[g] = (int)gb.size();
        const int gk = g % nw, gj = (g / nw) % nv, gi = g / nyz;
        for (int in1 = 0; in1 < hu; in1++)
            for (int jn1 = 0; jn1 < hv; jn1++)
                for (int kn1 = 0; kn1 < hw; kn1++) {
                    const double v =
                        ha[((long)((in1 * nu + gi) * (hv * hw) + jn1 * hw + kn1)) * nyz + gj * nw + gk];
                    if (v != 0.0) {
                        gb.push_back(glist_pack(in1 << 1 | (gi & 1), jn1 << 1 | (gj & 1), kn1 << 1 | (gk & 1)));
                        gv.push_back(v);
                    }
                }
    }
    gp[nb] = (int)gb.size();
    const size_t ne = std::max<size_t>(gb.size(), 1);
    GDT_CHECK(h->gptr.ensure(sizeof(int) * (nb + 1)) == 0 && h->gbox.ensure(sizeof(int) * ne) == 0 &&
                  h->gval.ensure(sizeof(double) * ne) == 0,
              GDT_E_NOMEM, "gdt_set_tables: list allocation failed");
    GDT_HIP(hipMemcpy(h->gptr.p, gp.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
    if (!gb.empty()) {
        GDT_HIP(hipMemcpy(h->gbox.p, gb.data(), sizeof(int) * gb.size(), hipMemcpyHostToDevice));
        GDT_HIP(hipMemcpy(h->gval.p, gv.data(), sizeof(double) * gv.size(), hipMemcpyHostToDevice));
    }
    h->n_a = n_a;
    h->n_gent = (long)gb.size();
    const int rc = gain_prepare(h);
    if (rc != GDT_OK) return rc;
    h->table_mode = -1;
    h->have_tables = true;
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// Korobov stencil
// ---------------------------------------------------------------------------

uint32_t gdt_korobov_pick(uint32_t M, uint64_t seed, int ncand, double *p2)
{
    if (M < 4 || (M & (M - 1)) || ncand < 1) return 0;
    return host_korobov_pick(M, seed, ncand, p2);
}

// Gather tiles of the stencil from the host copy of the group row pointers:
// per group, contiguous destination ranges of at most KOR_TILE_DST
// destinations and about 1/NT of the group's entries; NT is the largest count
// (<= h->ktiles_max) for which every group needs at most h->ktiles_max tiles
// (G * ktiles_max blocks per cell).  Tiles hold whole destinations, so the
// per-destination sums do not depend on the tiling; the moment partials (one
// per tile) do, in their last bits.
static int kor_make_tiles(gdt_context *h)
{
    const int G = KOR_NGROUPS, np = h->n_pnt, tmax = h->ktiles_max;
    const std::vector<int32_t> &gp = h->h_gp;
    std::vector<int4> tiles;
    int ntile = 0, maxe = 0;
    for (int NT = tmax; NT >= 1; NT--) {
        std::vector<std::vector<int4>> per(G);
        ntile = 0;
        maxe = 0;
        for (int g = 0; g < G; g++) {
            const int32_t *pg = &gp[(size_t)g * (np + 1)];
            const long quota = std::max(1L, ((long)pg[np] - pg[0] + NT - 1) / NT);
            int d = 0;
            while (d < np) {
                int e = d;
                while (e < np && e - d < KOR_TILE_DST && (e == d || (long)pg[e] - pg[d] < quota)) e++;
                per[g].push_back(make_int4(d, e, pg[d], pg[e]));
                maxe = std::max(maxe, pg[e] - pg[d]);
                d = e;
            }
            ntile = std::max(ntile, (int)per[g].size());
        }
        if (ntile <= tmax || NT == 1) {
            tiles.assign((size_t)G * ntile, make_int4(0, 0, 0, 0));
            for (int g = 0; g < G; g++)
                for (size_t k = 0; k < per[g].size(); k++) tiles[(size_t)g * ntile + k] = per[g][k];
            break;
        }
    }
    GDT_CHECK(h->ktile.ensure(sizeof(int4) * tiles.size()) == 0, GDT_E_NOMEM, "gather tile allocation failed");
    GDT_HIP(hipMemcpy(h->ktile.p, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
    h->kntile = ntile;
    h->kcap = std::max(64, std::min(maxe, KOR_TILE_CAP));
    h->kor_gen++;
    return GDT_OK;
}

int gdt_korobov_build(gdt_handle h, uint32_t M, uint64_t seed, uint32_t a, double s0)
{
    H_CHECK(h);
    GDT_CHECK(M >= 1024 && M <= (1u << 26) && (M & (M - 1)) == 0, GDT_E_ARG,
              "gdt_korobov_build: M = %u must be a power of two in [2^10, 2^26]", M);
    GDT_CHECK(h->n[0] == h->n[1] && h->n[1] == h->n[2], GDT_E_ARG, "gdt_korobov_build: grid must be cubic");
    GDT_CHECK(std::fabs(h->d[0] - h->d[1]) <= 1e-14 * h->d[0] && std::fabs(h->d[0] - h->d[2]) <= 1e-14 * h->d[0],
              GDT_E_ARG, "gdt_korobov_build: spacing must be equal on all axes");
    GDT_CHECK(s0 > 0.0, GDT_E_ARG, "gdt_korobov_build: s0 must be positive");
    GDT_CHECK(a == 0 || (a & 1u), GDT_E_ARG, "gdt_korobov_build: generator must be odd");
    if (a == 0) a = host_korobov_pick(M, seed, 8, nullptr);
    const double d = h->d[0];
    // node weight K sqrt(D2): pi s0 n_box^2 d^4 / M (weak form of the symmetric
    // direct-minus-inverse node term; the loss carries twice this weight)
    const double K = 3.141592653589793 * s0 * (double)h->n_box * (double)h->n_box * d * d * d * d / (double)M;

    DevBuf t_idx, t_r, t_w, t_v, t_i;
    GDT_CHECK(t_idx.ensure(sizeof(int32_t) * 6 * (size_t)M) == 0 && t_r.ensure(sizeof(double) * M) == 0 &&
                  t_w.ensure(sizeof(double) * M) == 0 && t_v.ensure(M) == 0 &&
                  t_i.ensure(sizeof(int32_t) * 3 * (size_t)M) == 0,
              GDT_E_NOMEM, "gdt_korobov_build: temporary allocation failed");
    int rc = launch_korobov_nodes(h, M, a, K, t_idx.as<int32_t>(), t_r.as<double>(), t_w.as<double>(),
                                  t_i.as<int32_t>(), t_v.as<uint8_t>());
    if (rc != GDT_OK) return rc;
    std::vector<uint8_t> hv(M);
    std::vector<int32_t> hidx(6 * (size_t)M), hint(3 * (size_t)M);
    std::vector<double> hr(M), hw(M);
    GDT_HIP(hipMemcpyAsync(hv.data(), t_v.p, M, hipMemcpyDeviceToHost, h->stream));
    GDT_HIP(hipMemcpyAsync(hidx.data(), t_idx.p, sizeof(int32_t) * 6 * (size_t)M, hipMemcpyDeviceToHost, h->stream));
    GDT_HIP(hipMemcpyAsync(hint.data(), t_i.p, sizeof(int32_t) * 3 * (size_t)M, hipMemcpyDeviceToHost, h->stream));
    GDT_HIP(hipMemcpyAsync(hr.data(), t_r.p, sizeof(double) * M, hipMemcpyDeviceToHost, h->stream));
    GDT_HIP(hipMemcpyAsync(hw.data(), t_w.p, sizeof(double) * M, hipMemcpyDeviceToHost, h->stream));
    GDT_HIP(hipStreamSynchronize(h->stream));
    t_idx.release();
    t_r.release();
    t_w.release();
    t_v.release();
    t_i.release();

    // keep the valid nodes in lattice order
    std::vector<int32_t> cidx6, cint;
    std::vector<double> cr, cw;
    std::vector<uint32_t> cnode;
    for (uint32_t t = 0; t < M; t++)
        if (hv[t]) {
            cidx6.insert(cidx6.end(), hidx.begin() + 6 * (size_t)t, hidx.begin() + 6 * (size_t)t + 6);
            cint.insert(cint.end(), hint.begin() + 3 * (size_t)t, hint.begin() + 3 * (size_t)t + 3);
            cr.push_back(hr[t]);
            cw.push_back(hw[t]);
            cnode.push_back(t);
        }
    const long nn = (long)cr.size();
    GDT_CHECK(nn > 0, GDT_E_ARG, "gdt_korobov_build: no lattice node survives the domain tests");
    GDT_CHECK(nn < (1L << 29), GDT_E_ARG, "gdt_korobov_build: too many nodes");

    // destination-sorted CSR by a stable counting sort: within a segment the
    // entries follow (node, slot) order
    const int np = h->n_pnt;
    std::vector<int32_t> ptr(np + 1, 0);
    auto slots = [&](long t) { return cr[t] == 0.0 ? 4 : 6; };
    for (long t = 0; t < nn; t++)
        for (int s = 0; s < slots(t); s++) ptr[cidx6[6 * t + s] + 1]++;
    for (int v = 0; v < np; v++) ptr[v + 1] += ptr[v];
    const long ne = ptr[np];
    std::vector<uint32_t> ent(ne);
    std::vector<int32_t> fillp(ptr.begin(), ptr.end() - 1);
    for (long t = 0; t < nn; t++)
        for (int s = 0; s < slots(t); s++) {
            const uint32_t role = s < 2 ? 0u : (s < 4 ? 1u : 2u);
            ent[fillp[cidx6[6 * t + s]]++] = ((uint32_t)t << 2) | role;
        }

    GDT_CHECK(h->kidx.ensure(sizeof(int32_t) * 6 * nn) == 0 && h->kr.ensure(sizeof(double) * nn) == 0 &&
                  h->kw.ensure(sizeof(double) * nn) == 0 && h->knode.ensure(sizeof(uint32_t) * nn) == 0 &&
                  h->kptr.ensure(sizeof(int32_t) * (np + 1)) == 0 && h->kent.ensure(sizeof(uint32_t) * ne) == 0,
              GDT_E_NOMEM, "gdt_korobov_build: stencil allocation failed");
    GDT_HIP(hipMemcpy(h->kidx.p, cidx6.data(), sizeof(int32_t) * 6 * nn, hipMemcpyHostToDevice));
    GDT_HIP(hipMemcpy(h->kr.p, cr.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    GDT_HIP(hipMemcpy(h->kw.p, cw.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
    GDT_HIP(hipMemcpy(h->knode.p, cnode.data(), sizeof(uint32_t) * nn, hipMemcpyHostToDevice));
    GDT_HIP(hipMemcpy(h->kptr.p, ptr.data(), sizeof(int32_t) * (np + 1), hipMemcpyHostToDevice));
    GDT_HIP(hipMemcpy(h->kent.p, ent.data(), sizeof(uint32_t) * ne, hipMemcpyHostToDevice));

    h->nn = nn;  // read by the wave partition below
    h->nent = ne;
    // packed 16-byte node records (the f of a cell in LDS): 15-bit node
    // indices, num / 4 and den / 4 in 11 bits each (the energies of a cubic
    // grid's node pairs are all congruent mod 4, so both are multiples of 4,
    // of one sign: r = |num| / |den| >= 0), D2 in 12 bits
    bool stream_ok = np <= KOR_NPTS_MAX;
    for (long t = 0; t < nn && stream_ok; t++) {
        const int32_t num = cint[3 * t], den = cint[3 * t + 1], d2 = cint[3 * t + 2];
        stream_ok = num % 4 == 0 && den % 4 == 0 && std::abs(num) / 4 <= 2047 && std::abs(den) / 4 <= 2047 &&
                    ((num >= 0 && den >= 0) || (num <= 0 && den <= 0)) && d2 >= 0 && d2 < 4096;
    }
    GDT_CHECK(stream_ok, GDT_E_ARG,
              "gdt_korobov_build: grid of %d nodes is beyond the packed-record limits (<= 19,200 nodes)", np);
    h->have_stream = false;
    h->wb_blocks = 0;
    h->h_kptr = ptr;
    {
        // (a | b << 15 | D2 >> 10 << 30, l | m << 15, ls | ms << 15,
        //  |num| / 4 | |den| / 4 << 11 | (D2 & 1023) << 22)
        auto pack = [&](long t, uint32_t *o4) {
            const int32_t *o = &cidx6[6 * (size_t)t];
            const uint32_t d2 = (uint32_t)cint[3 * t + 2];
            const uint32_t num4 = (uint32_t)(std::abs(cint[3 * t]) / 4), den4 = (uint32_t)(std::abs(cint[3 * t + 1]) / 4);
            o4[0] = (uint32_t)o[0] | ((uint32_t)o[1] << 15) | ((d2 >> 10) << 30);
            o4[1] = (uint32_t)o[2] | ((uint32_t)o[3] << 15);
            o4[2] = (uint32_t)o[4] | ((uint32_t)o[5] << 15);
            o4[3] = num4 | (den4 << 11) | ((d2 & 1023u) << 22);
        };
        // Node order of the packed records (node values and gather): sorted
        // by the lambda destination, then by the lambda+s destination,
        // lattice order among equals.  The entries a destination receives as
        // lambda then come from one contiguous run of node values (coalesced
        // loads in the gather), and those it receives as lambda+s from one
        // contiguous sub-run of each neighbour's run (step 26.5 -> 26.3 us
        // against lambda alone).
        std::vector<uint32_t> order(nn);
        for (long t = 0; t < nn; t++) order[t] = (uint32_t)t;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            const int32_t lx = cidx6[6 * (size_t)x + 2], ly = cidx6[6 * (size_t)y + 2];
            return lx != ly ? lx < ly : cidx6[6 * (size_t)x + 4] < cidx6[6 * (size_t)y + 4];
        });
        std::vector<uint32_t> n4(4 * (size_t)nn);
        for (long p = 0; p < nn; p++) pack(order[p], &n4[4 * p]);
        std::vector<int32_t> dest_of(ne);
        for (int v = 0; v < np; v++)
            for (int e = ptr[v]; e < ptr[v + 1]; e++) dest_of[e] = v;
        std::vector<uint32_t> e4(4 * (size_t)ne), e1(ne);
        for (long e = 0; e < ne; e++) {
            const uint32_t t = ent[e] >> 2, role = ent[e] & 3u;
            pack(t, &e4[4 * e]);
            e1[e] = (uint32_t)cint[3 * t + 2] | (role << 12) | ((uint32_t)dest_of[e] << 14);
        }
        GDT_CHECK(h->kn4.ensure(sizeof(uint32_t) * 4 * nn) == 0 && h->ke4.ensure(sizeof(uint32_t) * 4 * ne) == 0 &&
                      h->ke1.ensure(sizeof(uint32_t) * ne) == 0,
                  GDT_E_NOMEM, "gdt_korobov_build: record allocation failed");
        GDT_HIP(hipMemcpy(h->kn4.p, n4.data(), sizeof(uint32_t) * 4 * nn, hipMemcpyHostToDevice));
        GDT_HIP(hipMemcpy(h->ke4.p, e4.data(), sizeof(uint32_t) * 4 * ne, hipMemcpyHostToDevice));
        GDT_HIP(hipMemcpy(h->ke1.p, e1.data(), sizeof(uint32_t) * ne, hipMemcpyHostToDevice));
        // per node-group CSRs of the gather over the sorted node positions p:
        // group g is [kor_group_begin(g), kor_group_begin(g + 1)) (the node
        // kernel's XCD-aligned ranges follow the same rule); within a group
        // the entries are sorted by destination, then by (p, slot).  Slot 2
        // (lambda) is not an entry of its own: the node kernel forms the
        // prefix sums of the lambda terms over each chunk of KOR_CHUNK
        // positions, and a lambda piece -- a run of one lambda destination
        // within a chunk -- is read as two values: +prefix at its last
        // position and, unless it starts the chunk, -prefix before its first
        // position.  Each such position carries an entry in slot 2 per value
        // (first position's before last position's): value nn + (value slots
        // before it), role 1, r = 0.
        const int G = KOR_NGROUPS;
        std::vector<int32_t> gp((size_t)G * (np + 1), 0);
        std::vector<int> pgrp(nn);
        for (int g = 0; g < G; g++)
            for (long p = kor_group_begin(g, nn); p < kor_group_begin(g + 1, nn); p++) pgrp[p] = g;
        auto grp = [&](long p) { return pgrp[p]; };
        // value slots: a piece's first position (not a chunk's first) and its
        // last position; the chunks' first slots and the slots' count
        auto lam_at = [&](long p) { return cidx6[6 * (size_t)order[p] + 2]; };
        std::vector<uint8_t> vslot(nn);  // bit 0: first position, bit 1: last position
        const long nchunk = (nn + KOR_CHUNK - 1) / KOR_CHUNK;
        std::vector<int32_t> cbase(nchunk + 1);
        long nslot = 0;
        for (long p = 0; p < nn; p++) {
            if (p % KOR_CHUNK == 0) cbase[p / KOR_CHUNK] = (int32_t)nslot;
            const bool head = p % KOR_CHUNK != 0 && lam_at(p - 1) != lam_at(p);
            const bool tail = p + 1 == nn || (p + 1) % KOR_CHUNK == 0 || lam_at(p + 1) != lam_at(p);
            vslot[p] = (uint8_t)((head ? 1 : 0) | (tail ? 2 : 0));
            nslot += (head ? 1 : 0) + (tail ? 1 : 0);
        }
        cbase[nchunk] = (int32_t)nslot;
        GDT_CHECK(nn + nslot < (1L << 30), GDT_E_ARG, "gdt_korobov_build: %ld node values exceed the entry format",
                  nn + nslot);
        auto nent_at = [&](long p, int sl) { return sl != 2 ? 1 : (vslot[p] & 1) + (vslot[p] >> 1); };
        for (long p = 0; p < nn; p++) {
            const long t = order[p];
            for (int sl = 0; sl < slots(t); sl++) gp[(size_t)grp(p) * (np + 1) + cidx6[6 * t + sl] + 1] += nent_at(p, sl);
        }
        long run = 0;
        for (int g = 0; g < G; g++) {
            int32_t *pg = &gp[(size_t)g * (np + 1)];
            pg[0] = (int32_t)run;
            for (int v = 0; v < np; v++) {
                run += pg[v + 1];
                pg[v + 1] = (int32_t)run;
            }
        }
        // entry records (p << 2 | role, the last word of the node's record;
        // a lambda value slot: (nn + slot) << 2 | 1, word 0, so r = 0)
        const long neg = gp[(size_t)G * (np + 1) - 1];
        std::vector<uint32_t> gent(2 * (size_t)std::max(neg, 1L));
        std::vector<int32_t> fill(gp);
        long slot = 0;
        for (long p = 0; p < nn; p++) {
            const long t = order[p];
            const int g = grp(p);
            for (int sl = 0; sl < slots(t); sl++) {
                const uint32_t role = sl < 2 ? 0u : (sl < 4 ? 1u : 2u);
                int32_t *fp = &fill[(size_t)g * (np + 1) + cidx6[6 * t + sl]];
                if (sl != 2) {
                    const int32_t pos = (*fp)++;
                    gent[2 * (size_t)pos] = ((uint32_t)p << 2) | role;
                    gent[2 * (size_t)pos + 1] = n4[4 * (size_t)p + 3];
                    continue;
                }
                for (int b = 0; b < 2; b++)
                    if (vslot[p] & (1 << b)) {
                        const int32_t pos = (*fp)++;
                        gent[2 * (size_t)pos] = ((uint32_t)(nn + slot++) << 2) | 1u;
                        gent[2 * (size_t)pos + 1] = 0u;
                    }
            }
        }
        GDT_CHECK(slot == nslot, GDT_E_STATE, "gdt_korobov_build: lambda value slot count");
        // the cell-tile path's own node order: sorted by the alpha
        // destination (then lambda, then lambda+s, lattice order among
        // equals), so that every destination's alpha entries are one
        // contiguous run of positions q.  The tile node kernel walks a
        // destination's run, stores each node's D row and sums the run's
        // alpha terms (-D, and 2 w f_beta for the loss) itself; the tile CSR
        // holds the other entries (beta, lambda, mu, lambda+s, mu-s).
        {
            std::vector<uint32_t> tord(nn);
            for (long t = 0; t < nn; t++) tord[t] = (uint32_t)t;
            std::stable_sort(tord.begin(), tord.end(), [&](uint32_t x, uint32_t y) {
                const int32_t *ox = &cidx6[6 * (size_t)x], *oy = &cidx6[6 * (size_t)y];
                if (ox[0] != oy[0]) return ox[0] < oy[0];
                if (ox[2] != oy[2]) return ox[2] < oy[2];
                return ox[4] < oy[4];
            });
            // alpha runs: aptr[v] .. aptr[v + 1] are the positions whose alpha is v
            std::vector<int32_t> aptr(np + 1, 0);
            for (long q = 0; q < nn; q++) aptr[cidx6[6 * (size_t)tord[q]] + 1]++;
            for (int v = 0; v < np; v++) aptr[v + 1] += aptr[v];
            std::vector<uint32_t> t4(4 * (size_t)nn);
            for (long q = 0; q < nn; q++) pack(tord[q], &t4[4 * q]);
            // entry slots of the tile CSR: 1 (beta) and, where r != 0, 2..5;
            // where r = 0, 1..3
            std::vector<int32_t> cp(np + 1, 0);
            for (long q = 0; q < nn; q++) {
                const long t = tord[q];
                for (int sl = 1; sl < slots(t); sl++) cp[cidx6[6 * t + sl] + 1]++;
            }
            for (int v = 0; v < np; v++) cp[v + 1] += cp[v];
            const long nte = cp[np];
            // entry record (q << 2 | role, g, value as FP64 bits):
            //   role 0 (beta):  g = the partner node alpha, value = 2 w (the
            //                   loss is f_d sum 2 w f_g);
            //   role 1 / 2:     g = the destination itself, value = 1 - r / r,
            //                   r = (num / 4) / (den / 4) -- the quotient the
            //                   group gather forms
            // Two orders of the same entries within each destination's segment:
            //   cent   (q, slot) order, read by the FP64 gather;
            //   cent2  the beta entries first (in q order), then the others in
            //          (q, slot) order; cmid[v] is where the others begin, so
            //          the FP32 gather loads a partner's f row in the first
            //          run only.
            std::vector<uint32_t> cent(4 * (size_t)nte), cent2(4 * (size_t)nte);
            std::vector<int32_t> cmid(np);
            auto put = [&](std::vector<uint32_t> &dst_ent, std::vector<int32_t> &fillp, long q, int sl) {
                const long t = tord[q];
                const uint32_t y = t4[4 * (size_t)q + 3];
                const double num4 = (double)(y & 0x7FFu), den4 = (double)((y >> 11) & 0x7FFu);
                const double r = num4 == 0.0 ? 0.0 : num4 / den4;
                const uint32_t role = sl < 2 ? 0u : (sl < 4 ? 1u : 2u);
                const int32_t dst = cidx6[6 * t + sl];
                const double val = role == 0u ? 2.0 * cw[t] : (role == 1u ? 1.0 - r : r);
                const int32_t g = role == 0u ? cidx6[6 * t] : dst;
                uint64_t cb;
                std::memcpy(&cb, &val, sizeof cb);
                const int32_t pos = fillp[dst]++;
                dst_ent[4 * (size_t)pos] = ((uint32_t)q << 2) | role;
                dst_ent[4 * (size_t)pos + 1] = (uint32_t)g;
                dst_ent[4 * (size_t)pos + 2] = (uint32_t)cb;
                dst_ent[4 * (size_t)pos + 3] = (uint32_t)(cb >> 32);
            };
            {
                std::vector<int32_t> cf(cp.begin(), cp.end() - 1);
                for (long q = 0; q < nn; q++)
                    for (int sl = 1; sl < slots(tord[q]); sl++) put(cent, cf, q, sl);
            }
            {
                std::vector<int32_t> cf(cp.begin(), cp.end() - 1);
                for (long q = 0; q < nn; q++) put(cent2, cf, q, 1);
                for (int v = 0; v < np; v++) cmid[v] = cf[v];
                for (long q = 0; q < nn; q++)
                    for (int sl = 2; sl < slots(tord[q]); sl++) put(cent2, cf, q, sl);
            }
            // per tile position: r and w of the node (FP64, and rounded to
            // FP32 for the mixed node terms) -- r the quotient num / den the
            // node kernels form, w = K sqrt(D2) as the lattice kernel formed it
            std::vector<double> rw(2 * (size_t)nn);
            std::vector<float> rwf(2 * (size_t)nn);
            for (long q = 0; q < nn; q++) {
                const long t = tord[q];
                const uint32_t y = t4[4 * (size_t)q + 3];
                const double num4 = (double)(y & 0x7FFu), den4 = (double)((y >> 11) & 0x7FFu);
                rw[2 * q] = num4 == 0.0 ? 0.0 : num4 / den4;
                rw[2 * q + 1] = cw[t];
                rwf[2 * q] = (float)rw[2 * q];
                rwf[2 * q + 1] = (float)rw[2 * q + 1];
            }
            // destinations longest first: by tile-CSR entries (gather) and by
            // alpha-run length (node kernel)
            std::vector<int32_t> dperm(np), aperm(np);
            for (int v = 0; v < np; v++) dperm[v] = aperm[v] = v;
            std::stable_sort(dperm.begin(), dperm.end(),
                             [&](int32_t x, int32_t y) { return cp[x + 1] - cp[x] > cp[y + 1] - cp[y]; });
            std::stable_sort(aperm.begin(), aperm.end(),
                             [&](int32_t x, int32_t y) { return aptr[x + 1] - aptr[x] > aptr[y + 1] - aptr[y]; });
            // per XCD, one octant of the box (index halves), in Morton order:
            // the destinations a XCD gathers at one time are a compact part of
            // the grid, so a node's row read for lambda is still in that
            // XCD's L2 when the neighbour lambda + s reads it (and mu, mu - s)
            std::vector<int32_t> xperm(np);
            std::vector<uint64_t> xkey(np);
            for (int v = 0; v < np; v++) {
                const int i = h->h_cidx[3 * v], j = h->h_cidx[3 * v + 1], k = h->h_cidx[3 * v + 2];
                const uint64_t oct = (uint64_t)((i >= h->n[0] / 2) * 4 + (j >= h->n[1] / 2) * 2 + (k >= h->n[2] / 2));
                uint64_t mc = 0;
                for (int bit = 0; bit < 15; bit++)
                    mc |= (((uint64_t)(i >> bit) & 1u) << (3 * bit + 2)) | (((uint64_t)(j >> bit) & 1u) << (3 * bit + 1)) |
                          (((uint64_t)(k >> bit) & 1u) << (3 * bit));
                xkey[v] = (oct << 48) | mc;
                xperm[v] = v;
            }
            std::stable_sort(xperm.begin(), xperm.end(), [&](int32_t x, int32_t y) { return xkey[x] < xkey[y]; });
            for (int x = 0; x <= 8; x++) h->kxoff[x] = 0;
            for (int v = 0; v < np; v++) h->kxoff[(xkey[v] >> 48) + 1]++;
            for (int x = 0; x < 8; x++) h->kxoff[x + 1] += h->kxoff[x];
            GDT_CHECK(h->krw.ensure(sizeof(double) * 2 * nn) == 0 && h->krwf.ensure(sizeof(float) * 2 * nn) == 0 &&
                          h->kt4.ensure(sizeof(uint32_t) * 4 * nn) == 0 &&
                          h->kaptr.ensure(sizeof(int32_t) * (np + 1)) == 0 &&
                          h->kaperm.ensure(sizeof(int32_t) * np) == 0 &&
                          h->kcptr.ensure(sizeof(int32_t) * (np + 1)) == 0 &&
                          h->kcmid.ensure(sizeof(int32_t) * np) == 0 &&
                          h->kcent.ensure(sizeof(uint32_t) * 4 * std::max(nte, 1L)) == 0 &&
                          h->kcent2.ensure(sizeof(uint32_t) * 4 * std::max(nte, 1L)) == 0 &&
                          h->kdperm.ensure(sizeof(int32_t) * np) == 0 &&
                          h->kxperm.ensure(sizeof(int32_t) * np) == 0,
                      GDT_E_NOMEM, "gdt_korobov_build: cell-tile stencil allocation failed");
            GDT_HIP(hipMemcpy(h->krw.p, rw.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice));
            GDT_HIP(hipMemcpy(h->krwf.p, rwf.data(), sizeof(float) * 2 * nn, hipMemcpyHostToDevice));
            GDT_HIP(hipMemcpy(h->kt4.p, t4.data(), sizeof(uint32_t) * 4 * nn, hipMemcpyHostToDevice));
            GDT_HIP(hipMemcpy(h->kaptr.p, aptr.data(), sizeof(int32_t) * (np + 1), hipMemcpyHostToDevice));
            GDT_HIP(hipMemcpy(h->kaperm.p, aperm.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
            GDT_HIP(hipMemcpy(h->kcptr.p, cp.data(), sizeof(int32_t) * (np + 1), hipMemcpyHostToDevice));
            GDT_HIP(hipMemcpy(h->kcmid.p, cmid.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
            if (nte > 0) {
                GDT_HIP(hipMemcpy(h->kcent.p, cent.data(), sizeof(uint32_t) * 4 * nte, hipMemcpyHostToDevice));
                GDT_HIP(hipMemcpy(h->kcent2.p, cent2.data(), sizeof(uint32_t) * 4 * nte, hipMemcpyHostToDevice));
            }
            GDT_HIP(hipMemcpy(h->kdperm.p, dperm.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
            GDT_HIP(hipMemcpy(h->kxperm.p, xperm.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice));
            h->ntent = nte;
        }
        h->h_gp.swap(gp);
        GDT_CHECK(h->kgptr.ensure(sizeof(int32_t) * h->h_gp.size()) == 0 &&
                      h->kgent.ensure(sizeof(uint32_t) * gent.size()) == 0 &&
                      h->kcbase.ensure(sizeof(int32_t) * cbase.size()) == 0,
                  GDT_E_NOMEM, "gdt_korobov_build: group CSR allocation failed");
        GDT_HIP(hipMemcpy(h->kgptr.p, h->h_gp.data(), sizeof(int32_t) * h->h_gp.size(), hipMemcpyHostToDevice));
        GDT_HIP(hipMemcpy(h->kgent.p, gent.data(), sizeof(uint32_t) * gent.size(), hipMemcpyHostToDevice));
        GDT_HIP(hipMemcpy(h->kcbase.p, cbase.data(), sizeof(int32_t) * cbase.size(), hipMemcpyHostToDevice));
        h->nv = nn + nslot;
        int rc1 = kor_make_tiles(h);
        if (rc1) return rc1;
        int rc2 = kor_setup(h, 1);
        if (rc2) return rc2;
        h->have_stream = true;
    }
    h->kM = M;
    h->ka = a;
    h->ks0 = s0;
    h->kK = K;
    h->nn = nn;
    h->nent = ne;
    h->have_kor = true;
    if (h->graph_exec) {
        (void)hipGraphExecDestroy(h->graph_exec);
        h->graph_exec = nullptr;
    }
    return GDT_OK;
}

int gdt_korobov_info(gdt_handle h, uint32_t *M, uint32_t *a, long *nodes, long *entries, double *K)
{
    H_CHECK(h);
    GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_korobov_info: stencil not built");
    if (M) *M = h->kM;
    if (a) *a = h->ka;
    if (nodes) *nodes = h->nn;
    if (entries) *entries = h->nent;
    if (K) *K = h->kK;
    return GDT_OK;
}

int gdt_korobov_export(gdt_handle h, int32_t *idx6, double *r, double *w, uint32_t *node)
{
    H_CHECK(h);
    GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_korobov_export: stencil not built");
    if (idx6) GDT_HIP(hipMemcpy(idx6, h->kidx.p, sizeof(int32_t) * 6 * h->nn, hipMemcpyDeviceToHost));
    if (r) GDT_HIP(hipMemcpy(r, h->kr.p, sizeof(double) * h->nn, hipMemcpyDeviceToHost));
    if (w) GDT_HIP(hipMemcpy(w, h->kw.p, sizeof(double) * h->nn, hipMemcpyDeviceToHost));
    if (node) GDT_HIP(hipMemcpy(node, h->knode.p, sizeof(uint32_t) * h->nn, hipMemcpyDeviceToHost));
    return GDT_OK;
}

int gdt_korobov_export_csr(gdt_handle h, int32_t *ptr, uint32_t *ent)
{
    H_CHECK(h);
    GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_korobov_export_csr: stencil not built");
    if (ptr) GDT_HIP(hipMemcpy(ptr, h->kptr.p, sizeof(int32_t) * (h->n_pnt + 1), hipMemcpyDeviceToHost));
    if (ent) GDT_HIP(hipMemcpy(ent, h->kent.p, sizeof(uint32_t) * h->nent, hipMemcpyDeviceToHost));
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// collision operator
// ---------------------------------------------------------------------------

static int to_cell_major(gdt_context *h, const double *src, int cells, int layout, DevBuf &tmp, const double **out)
{
    if (layout == GDT_LAYOUT_CELL_MAJOR) {
        *out = src;
        return GDT_OK;
    }
    GDT_CHECK(layout == GDT_LAYOUT_NODE_MAJOR, GDT_E_ARG, "unknown layout %d", layout);
    GDT_CHECK(tmp.ensure(sizeof(double) * (size_t)cells * h->n_pnt) == 0, GDT_E_NOMEM, "layout scratch");
    int rc = launch_transpose(h, src, tmp.as<double>(), h->n_pnt, cells);
    *out = tmp.as<double>();
    return rc;
}

// Matrix-method collide into cell-major compact outputs.
static int collide_matrix_cm(gdt_context *h, const double *f_cm, double *nu, double *gain, int cells)
{
    const size_t nbox = (size_t)cells * h->n_box;
    GDT_CHECK(h->s_fbox.ensure(sizeof(double) * nbox) == 0 && h->s_out1.ensure(sizeof(double) * nbox) == 0,
              GDT_E_NOMEM, "collide scratch allocation failed");
    int rc = launch_cmp_to_box(h, f_cm, h->s_fbox.as<double>(), cells, GDT_LAYOUT_CELL_MAJOR);
    if (rc) return rc;
    if (nu) {
        if ((rc = launch_loss(h, h->s_fbox.as<double>(), h->s_out1.as<double>(), cells))) return rc;
        if ((rc = launch_box_to_cmp(h, h->s_out1.as<double>(), nu, cells, GDT_LAYOUT_CELL_MAJOR))) return rc;
    }
    if (gain) {
        if ((rc = launch_gain(h, h->s_fbox.as<double>(), h->s_out1.as<double>(), cells))) return rc;
        if ((rc = launch_box_to_cmp(h, h->s_out1.as<double>(), gain, cells, GDT_LAYOUT_CELL_MAJOR))) return rc;
    }
    return GDT_OK;
}

static bool use_fused(unsigned flags) { return (flags & GDT_FLAG_FUSED) != 0; }
static bool use_tiles(unsigned flags) { return (flags & (GDT_FLAG_CELL_TILES | GDT_FLAG_MIXED)) != 0; }

// Scratch of the Korobov path for `cells` cells: moment sums, node values,
// group partials, and the fused path's wave partition (outside any capture).
static int kor_prepare(gdt_context *h, int cells)
{
    GDT_CHECK_GRID_Y(cells, "Korobov: cells");
    const int nb = std::max(kor_blocks(h, cells), kor_gather_parts(h));
    const size_t mp = sizeof(double) * 20 * (size_t)cells * nb;
    const size_t dw = sizeof(double) * 2 * (size_t)h->nv * cells;
    const size_t pt = sizeof(double) * 2 * (size_t)KOR_NGROUPS * h->n_pnt * cells;
    GDT_CHECK(h->s_mpart.ensure(mp) == 0 && h->s_vals.ensure(dw) == 0 && h->s_part.ensure(pt) == 0, GDT_E_NOMEM,
              "Korobov scratch allocation failed (%zu bytes of node values)", dw);
    return kor_setup(h, cells);
}

// Group partials of loss and Q (and the moment sums in s_mpart) of the
// Korobov operator for `cells` cell-major cells.
struct KorPartials {
    int G = 1, nparts = 1;
    double *L = nullptr, *Q = nullptr;
};

static int kor_partials(gdt_context *h, const double *f_cm, int cells, unsigned flags, KorPartials &kp)
{
    int rc = kor_prepare(h, cells);
    if (rc) return rc;
    if (use_fused(flags)) {
        kp.G = 1;
        kp.nparts = kor_blocks(h, cells);
        kp.L = h->s_part.as<double>();
        kp.Q = kp.L + (size_t)cells * h->n_pnt;
        return launch_kor_fused(h, f_cm, kp.L, kp.Q, h->s_mpart.as<double>(), cells);
    }
    kp.G = KOR_NGROUPS;
    kp.nparts = kor_gather_parts(h);
    kp.L = h->s_part.as<double>();
    kp.Q = kp.L + (size_t)KOR_NGROUPS * cells * h->n_pnt;
    if ((rc = launch_kor_node(h, f_cm, h->s_vals.as<double>(), cells))) return rc;
    return launch_kor_gather(h, h->s_vals.as<double>(), kp.L, kp.Q, h->s_mpart.as<double>(), cells);
}

// Korobov collide into cell-major loss / Q (uncorrected).
static int collide_korobov_cm(gdt_context *h, const double *f_cm, double *loss, double *Q, int cells,
                              unsigned flags)
{
    KorPartials kp;
    int rc = kor_partials(h, f_cm, cells, flags, kp);
    if (rc) return rc;
    return launch_kor_apply(h, false, kp.G, kp.nparts, h->s_mpart.as<double>(), f_cm, kp.L, kp.Q, Q, loss, nullptr,
                            nullptr, cells, 0.0, nullptr);
}

int gdt_collide_matrix(gdt_handle h, const double *f, double *nu, double *gain, int cells, int layout)
{
    H_CHECK(h);
    GDT_CHECK(h->have_tables, GDT_E_STATE, "gdt_collide_matrix: tables not built");
    GDT_CHECK(f && cells > 0, GDT_E_ARG, "gdt_collide_matrix: bad arguments");
    const double *fcm;
    int rc = to_cell_major(h, f, cells, layout, h->s_tmp, &fcm);
    if (rc) return rc;
    if (layout == GDT_LAYOUT_CELL_MAJOR) return collide_matrix_cm(h, fcm, nu, gain, cells);
    const size_t nc = (size_t)cells * h->n_pnt;
    GDT_CHECK(h->s_nu.ensure(sizeof(double) * nc) == 0 && h->s_out2.ensure(sizeof(double) * nc) == 0, GDT_E_NOMEM,
              "collide scratch allocation failed");
    rc = collide_matrix_cm(h, fcm, nu ? h->s_nu.as<double>() : nullptr, gain ? h->s_out2.as<double>() : nullptr, cells);
    if (rc) return rc;
    if (nu && (rc = launch_transpose(h, h->s_nu.as<double>(), nu, cells, h->n_pnt))) return rc;
    if (gain && (rc = launch_transpose(h, h->s_out2.as<double>(), gain, cells, h->n_pnt))) return rc;
    return GDT_OK;
}

int gdt_collide_korobov(gdt_handle h, const double *f, double *loss, double *Q, int cells, int layout)
{
    H_CHECK(h);
    GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_collide_korobov: stencil not built");
    GDT_CHECK(f && Q && cells > 0, GDT_E_ARG, "gdt_collide_korobov: bad arguments");
    const double *fcm;
    int rc = to_cell_major(h, f, cells, layout, h->s_tmp, &fcm);
    if (rc) return rc;
    if (layout == GDT_LAYOUT_CELL_MAJOR) return collide_korobov_cm(h, fcm, loss, Q, cells, 0);
    const size_t nc = (size_t)cells * h->n_pnt;
    GDT_CHECK(h->s_loss.ensure(sizeof(double) * nc) == 0 && h->s_Q.ensure(sizeof(double) * nc) == 0, GDT_E_NOMEM,
              "collide scratch allocation failed");
    rc = collide_korobov_cm(h, fcm, h->s_loss.as<double>(), h->s_Q.as<double>(), cells, 0);
    if (rc) return rc;
    if (loss && (rc = launch_transpose(h, h->s_loss.as<double>(), loss, cells, h->n_pnt))) return rc;
    return launch_transpose(h, h->s_Q.as<double>(), Q, cells, h->n_pnt);
}

int gdt_correct(gdt_handle h, int method, const double *f, double *nu, const double *gain, double *Q, double *X,
                int cells, int layout)
{
    H_CHECK(h);
    GDT_CHECK(f && nu && gain && Q && cells > 0, GDT_E_ARG, "gdt_correct: bad arguments");
    GDT_CHECK(layout == GDT_LAYOUT_CELL_MAJOR, GDT_E_ARG, "gdt_correct: cell-major layout only");
    GDT_CHECK(method == GDT_METHOD_MATRIX || method == GDT_METHOD_KOROBOV, GDT_E_ARG, "gdt_correct: method %d",
              method);
    if (method == GDT_METHOD_KOROBOV) return launch_correct(h, method, f, nu, gain, nullptr, Q, X, cells, 0.0, nullptr);
    return launch_correct(h, method, f, nu, gain, nu, Q, X, cells, 0.0, nullptr);
}

int gdt_evaluate(gdt_handle h, int method, const double *f, double *nu, double *gain, double *Q, int cells, int layout,
                 unsigned flags)
{
    H_CHECK(h);
    GDT_CHECK(f && cells > 0, GDT_E_ARG, "gdt_evaluate: bad arguments");
    const bool timing = (flags & GDT_FLAG_TIME) != 0;
    h->timing = timing;
    const size_t nc = (size_t)cells * h->n_pnt;
    int rc = GDT_OK;
    const double *fcm;
    if ((rc = to_cell_major(h, f, cells, layout, h->s_tmp, &fcm))) goto done;
    GDT_CHECK(h->s_nu.ensure(sizeof(double) * nc) == 0 && h->s_out2.ensure(sizeof(double) * nc) == 0 &&
                  h->s_Q.ensure(sizeof(double) * nc) == 0 && h->s_loss.ensure(sizeof(double) * nc) == 0,
              GDT_E_NOMEM, "gdt_evaluate: scratch allocation failed");
    {
        double *nu_c = h->s_nu.as<double>(), *g_c = h->s_out2.as<double>(), *q_c = h->s_Q.as<double>();
        if (method == GDT_METHOD_MATRIX) {
            GDT_CHECK(h->have_tables, GDT_E_STATE, "gdt_evaluate: tables not built");
            if ((rc = collide_matrix_cm(h, fcm, nu_c, g_c, cells))) goto done;
            if (flags & GDT_FLAG_CORRECT) {
                if ((rc = launch_correct(h, method, fcm, nu_c, g_c, nu_c, q_c, nullptr, cells, 0.0, nullptr))) goto done;
            } else {
                if ((rc = launch_correct(h, -1, fcm, nu_c, g_c, nu_c, q_c, nullptr, cells, 0.0, nullptr))) goto done;
            }
        } else if (method == GDT_METHOD_KOROBOV) {
            GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_evaluate: stencil not built");
            double *L = h->s_loss.as<double>();
            if (use_tiles(flags)) {
                const bool mixed = (flags & GDT_FLAG_MIXED) != 0;
                if ((rc = ct_prepare(h, cells, mixed))) goto done;
                if ((rc = ct_run(h, mixed, fcm, cells, (flags & GDT_FLAG_CORRECT) != 0, q_c, L, nu_c, 0.0,
                                 nullptr)))
                    goto done;
            } else {
                KorPartials kp;
                if ((rc = kor_partials(h, fcm, cells, flags, kp))) goto done;
                // q_c = corrected Q, nu_c = corrected loss frequency, L = loss
                if ((rc = launch_kor_apply(h, (flags & GDT_FLAG_CORRECT) != 0, kp.G, kp.nparts,
                                           h->s_mpart.as<double>(), fcm, kp.L, kp.Q, q_c, L, nu_c, nullptr, cells,
                                           0.0, nullptr)))
                    goto done;
            }
            // gain = Q + f nu_c
            if ((rc = launch_correct(h, -3, fcm, nu_c, q_c, nullptr, g_c, nullptr, cells, 0.0, nullptr))) goto done;
        } else {
            gdt_set_error(GDT_E_ARG, "gdt_evaluate: method %d", method);
            rc = GDT_E_ARG;
            goto done;
        }
        double *outs[3] = {nu, gain, Q};
        double *srcs[3] = {nu_c, g_c, q_c};
        for (int q = 0; q < 3; q++) {
            if (!outs[q]) continue;
            if (layout == GDT_LAYOUT_CELL_MAJOR)
                GDT_HIP(hipMemcpyAsync(outs[q], srcs[q], sizeof(double) * nc, hipMemcpyDeviceToDevice, h->stream));
            else if ((rc = launch_transpose(h, srcs[q], outs[q], cells, h->n_pnt)))
                goto done;
        }
    }
done:
    h->timing = false;
    return rc;
}

int gdt_relax(gdt_handle h, double *f, const double *nu, const double *gain, int cells, double dt, double courant,
              double *time, unsigned flags)
{
    H_CHECK(h);
    GDT_CHECK(f && nu && gain && cells > 0, GDT_E_ARG, "gdt_relax: bad arguments");
    const long count = (long)cells * h->n_pnt;  // element-wise: layout-independent
    if (flags & GDT_FLAG_COURANT) {
        GDT_CHECK(h->s_dt.ensure(sizeof(double)) == 0, GDT_E_NOMEM, "gdt_relax: scratch");
        int rc = launch_courant_dt(h, nu, count, courant, h->s_dt.as<double>());
        if (rc) return rc;
        return launch_euler(h, f, nu, gain, count, 0.0, h->s_dt.as<double>(), time);
    }
    return launch_euler(h, f, nu, gain, count, dt, nullptr, time);
}

int gdt_device_synchronize(void)
{
    GDT_HIP(hipDeviceSynchronize());
    return GDT_OK;
}

int gdt_convert_f32_to_f64(const float *src, double *dst, size_t count)
{
    GDT_CHECK(src && dst, GDT_E_ARG, "gdt_convert_f32_to_f64: null pointer");
    return launch_convert(src, dst, count, nullptr, nullptr);
}

int gdt_convert_f64_to_f32(const double *src, float *dst, size_t count)
{
    GDT_CHECK(src && dst, GDT_E_ARG, "gdt_convert_f64_to_f32: null pointer");
    return launch_convert(nullptr, nullptr, count, src, dst);
}

// One explicit step of cell-major f (no host synchronisation).
static int one_step(gdt_context *h, int method, double *f, int cells, double dt, unsigned flags)
{
    int rc;
    const bool corr = (flags & GDT_FLAG_CORRECT) != 0;
    if (method == GDT_METHOD_KOROBOV && use_tiles(flags))
        return ct_run(h, (flags & GDT_FLAG_MIXED) != 0, f, cells, corr, nullptr, nullptr, nullptr, dt, f);
    if (method == GDT_METHOD_KOROBOV) {
        KorPartials kp;
        if ((rc = kor_partials(h, f, cells, flags, kp))) return rc;
        return launch_kor_apply(h, corr, kp.G, kp.nparts, h->s_mpart.as<double>(), f, kp.L, kp.Q, nullptr,
                                nullptr, nullptr, nullptr, cells, dt, f);
    }
    double *nu = h->s_nu.as<double>(), *g = h->s_out2.as<double>(), *Q = h->s_Q.as<double>();
    if ((rc = collide_matrix_cm(h, f, nu, g, cells))) return rc;
    return launch_correct(h, corr ? GDT_METHOD_MATRIX : -1, f, nu, g, nu, Q, nullptr, cells, dt, f);
}

int gdt_run_steps(gdt_handle h, int method, double *f, int cells, double dt, int steps, unsigned flags)
{
    H_CHECK(h);
    GDT_CHECK(f && cells > 0 && steps >= 0, GDT_E_ARG, "gdt_run_steps: bad arguments");
    GDT_CHECK(method == GDT_METHOD_MATRIX || method == GDT_METHOD_KOROBOV, GDT_E_ARG, "gdt_run_steps: method");
    if (method == GDT_METHOD_KOROBOV) GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_run_steps: stencil not built");
    if (method == GDT_METHOD_MATRIX) GDT_CHECK(h->have_tables, GDT_E_STATE, "gdt_run_steps: tables not built");
    const size_t nc = (size_t)cells * h->n_pnt;
    GDT_CHECK(h->s_nu.ensure(sizeof(double) * nc) == 0 && h->s_out2.ensure(sizeof(double) * nc) == 0 &&
                  h->s_Q.ensure(sizeof(double) * nc) == 0 && h->s_loss.ensure(sizeof(double) * nc) == 0,
              GDT_E_NOMEM, "gdt_run_steps: scratch allocation failed");
    if (method == GDT_METHOD_MATRIX) {
        const size_t nbox = (size_t)cells * h->n_box;
        GDT_CHECK(h->s_fbox.ensure(sizeof(double) * nbox) == 0 && h->s_out1.ensure(sizeof(double) * nbox) == 0,
                  GDT_E_NOMEM, "gdt_run_steps: scratch allocation failed");
    } else {
        int rc0 = use_tiles(flags) ? ct_prepare(h, cells, (flags & GDT_FLAG_MIXED) != 0) : kor_prepare(h, cells);
        if (rc0) return rc0;
    }
    const bool graph = (flags & GDT_FLAG_GRAPH) && !(flags & GDT_FLAG_TIME);
    h->timing = (flags & GDT_FLAG_TIME) != 0;
    int rc = GDT_OK;
    if (!graph) {
        for (int s = 0; s < steps && rc == GDT_OK; s++) rc = one_step(h, method, f, cells, dt, flags);
        h->timing = false;
        return rc;
    }
    h->timing = false;
    if (steps == 0) return GDT_OK;
    // one graph holds `chunk` consecutive steps (a divisor of steps when one
    // <= 64 exists), so replays are back to back
    int chunk = 1;
    for (int c = std::min(steps, 64); c >= 1; c--)
        if (steps % c == 0) {
            chunk = c;
            break;
        }
    if (chunk < 8 && steps > 64) chunk = 64;
    const unsigned key = flags & (GDT_FLAG_CORRECT | GDT_FLAG_FUSED | GDT_FLAG_CELL_TILES | GDT_FLAG_MIXED);
    // the captured kernels hold raw pointers to scratch and table buffers:
    // any DevBuf freed since the capture (a larger call elsewhere, a table
    // rebuild) makes the graph stale
    const unsigned long gen = devbuf_generation();
    if (!h->graph_exec || h->g_method != method || h->g_cells != cells || h->g_dt != dt || h->g_flags != key ||
        h->g_f != f || h->g_chunk != chunk || h->g_gen != gen) {
        if (h->graph_exec) {
            (void)hipGraphExecDestroy(h->graph_exec);
            h->graph_exec = nullptr;
        }
        hipGraph_t g = nullptr;
        GDT_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int s2 = 0; s2 < chunk && rc == GDT_OK; s2++) rc = one_step(h, method, f, cells, dt, flags);
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (rc != GDT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        GDT_HIP(e);
        e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GDT_HIP(e);
        h->g_method = method;
        h->g_cells = cells;
        h->g_dt = dt;
        h->g_flags = key;
        h->g_f = f;
        h->g_chunk = chunk;
        h->g_gen = devbuf_generation();
    }
    const int full = steps / chunk, rem = steps % chunk;
    for (int s = 0; s < full; s++) GDT_HIP(hipGraphLaunch(h->graph_exec, h->stream));
    for (int s = 0; s < rem && rc == GDT_OK; s++) rc = one_step(h, method, f, cells, dt, flags);
    return rc;
}

int gdt_kernel_stats(gdt_handle h, const char *name, double *total_ms, long *count)
{
    H_CHECK(h);
    GDT_CHECK(name, GDT_E_ARG, "gdt_kernel_stats: name is null");
    int rc = timing_resolve(h);
    if (rc) return rc;
    auto it = h->stats.find(name);
    if (total_ms) *total_ms = it == h->stats.end() ? 0.0 : it->second.ms;
    if (count) *count = it == h->stats.end() ? 0 : it->second.count;
    return GDT_OK;
}

int gdt_kernel_stats_reset(gdt_handle h)
{
    H_CHECK(h);
    int rc = timing_resolve(h);
    h->stats.clear();
    return rc;
}

int gdt_transpose(gdt_handle h, const double *src, double *dst, int rows, int cols)
{
    H_CHECK(h);
    GDT_CHECK(src && dst && rows > 0 && cols > 0, GDT_E_ARG, "gdt_transpose: bad arguments");
    return launch_transpose(h, src, dst, rows, cols);
}

}  // extern "C"

int gdt_init_mixture(gdt_handle h, const double *params, int comps, long cells, double *f)
{
    H_CHECK(h);
    GDT_CHECK(params && f && comps > 0 && cells > 0, GDT_E_ARG, "gdt_init_mixture: bad arguments");
    return launch_init_mixture(h, params, comps, cells, f);
}

int gdt_kmath_tables(double *T)
{
    GDT_CHECK(T, GDT_E_ARG, "gdt_kmath_tables: null pointer");
    kmath_tables(T);
    return GDT_OK;
}

int gdt_kmath_pow(gdt_handle h, const double *x, const double *r, double *out, long n)
{
    GDT_CHECK(h && x && r && out && n >= 0, GDT_E_ARG, "gdt_kmath_pow: bad arguments");
    if (n == 0) return GDT_OK;
    GDT_HIP(hipSetDevice(h->device));
    int rc = launch_kmath_pow(h, x, r, out, n);
    if (rc) return rc;
    GDT_HIP(hipStreamSynchronize(h->stream));
    return GDT_OK;
}

int gdt_korobov_tiles(gdt_handle h, int *groups, int *tiles, int *cap)
{
    H_CHECK(h);
    GDT_CHECK(h->have_kor, GDT_E_STATE, "gdt_korobov_tiles: stencil not built");
    if (groups) *groups = KOR_NGROUPS;
    if (tiles) *tiles = h->kntile;
    if (cap) *cap = h->kcap;
    return GDT_OK;
}

int gdt_korobov_gather_info(gdt_handle h, long *entries, long *slots)
{
    H_CHECK(h);
    GDT_CHECK(h->have_kor && h->have_stream, GDT_E_STATE, "gdt_korobov_gather_info: stencil not built");
    if (entries) *entries = h->h_gp.empty() ? 0 : h->h_gp.back();
    if (slots) *slots = h->nv - h->nn;
    return GDT_OK;
}

int gdt_korobov_cell_tiles(gdt_handle h, int chunk_tiles, int node_blocks, int gather_blocks, int wide)
{
    H_CHECK(h);
    GDT_HIP(hipSetDevice(h->device));
    GDT_HIP(hipStreamSynchronize(h->stream));  // no evaluation in flight uses the old shape
    if (h->graph_exec) {                        // captured launches carry the old grids
        (void)hipGraphExecDestroy(h->graph_exec);
        h->graph_exec = nullptr;
    }
    return ct_config(h, chunk_tiles, node_blocks, gather_blocks, wide);
}

int gdt_korobov_set_tiles(gdt_handle h, int tiles_max)
{
    H_CHECK(h);
    GDT_CHECK(tiles_max >= 1 && tiles_max <= KOR_TILES_LIMIT, GDT_E_ARG,
              "gdt_korobov_set_tiles: %d tiles per group (1..%d)", tiles_max, KOR_TILES_LIMIT);
    h->ktiles_max = tiles_max;
    if (!h->have_kor) return GDT_OK;  // applies to the next build
    GDT_HIP(hipSetDevice(h->device));
    GDT_HIP(hipStreamSynchronize(h->stream));  // no step in flight reads the old tiles
    int rc = kor_make_tiles(h);
    if (rc) return rc;
    if (h->graph_exec) {  // captured launches carry the old grid
        (void)hipGraphExecDestroy(h->graph_exec);
        h->graph_exec = nullptr;
    }
    return kor_setup(h, 1);
}
