// The reference's host API on top of libgdt (compute_matrix.h:13,
// velocity.h:41, collision_integral.h:11-24).  FP32 device arrays in the
// reference's [velocity][cell] layout are converted into the FP64 core and
// back; the collision operator itself runs in libgdt's kernels.
//
// Errors follow the reference: a failing call prints its message to stderr
// and the function returns without throwing (the reference prints
// cudaGetErrorString and continues); only load_options throws.
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/gdt.h"
#include "collision_integral.h"
#include "compute_matrix.h"
#include "velocity.h"

namespace {

gdt_handle g_ctx = nullptr;                      // context of the last grid given to init_*
std::map<const float*, gdt_handle> g_tables;     // a-table pointer -> context that built it

bool ok(int rc, const char* what)
{
    if (rc == GDT_OK) return true;
    std::fprintf(stderr, "%s failed: %s\n", what, gdt_last_error());
    return false;
}

void fail(const char* msg) { std::fprintf(stderr, "%s\n", msg); }

#define CHECK(x, what)              \
    do {                            \
        if (!ok((x), (what))) return; \
    } while (0)

// Lazily grown FP64 device scratch (the reference keeps static scratch too,
// collision_integral.cu:568-577).
struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
    double* get(size_t count)  // NULL (message printed) on failure
    {
        const size_t need = count * sizeof(double);
        if (need > bytes) {
            gdt_free(p);
            p = nullptr;
            bytes = 0;
            if (!ok(gdt_malloc(&p, need), "gdt_malloc")) return nullptr;
            bytes = need;
        }
        return static_cast<double*>(p);
    }
};
Scratch s_f, s_nu, s_gain, s_Q, s_time;

float* alloc_f32(size_t count)  // NULL (message printed) on failure
{
    void* p = nullptr;
    if (!ok(gdt_malloc(&p, count * sizeof(float)), "gdt_malloc")) return nullptr;
    return static_cast<float*>(p);
}

// Nodes of the handle's grid; -1 (message printed) unless the grid keeps the
// whole box, the layout of the reference's arrays.
int grid_points(gdt_handle h)
{
    int n_pnt = 0, n_box = 0;
    if (!ok(gdt_grid_info(h, &n_pnt, &n_box, nullptr, nullptr), "gdt_grid_info")) return -1;
    if (n_pnt != n_box) {
        fail("the drop-in API works on the whole box (R must keep every node); use the C ABI for sphere-cut grids");
        return -1;
    }
    return n_pnt;
}

}  // namespace

// velocity.cu:53-65: 14 moment-basis planes over the box.
void init_correction_array(float** correction_array, const DeviceVelocityGrid& vgrid, const Options& opts)
{
    g_ctx = vgrid.ctx;
    const size_t n = (size_t)opts.nxyz * 14;
    double* tmp = s_f.get(n);
    *correction_array = nullptr;
    if (!tmp) return;
    CHECK(gdt_basis(vgrid.ctx, tmp), "gdt_basis");
    CHECK(gdt_synchronize(vgrid.ctx), "gdt_synchronize");
    if (!(*correction_array = alloc_f32(n))) return;
    CHECK(gdt_convert_f64_to_f32(tmp, *correction_array, n), "convert");
}

// compute_matrix.cu:564-600: b and a tables (and, for the Korobov method,
// the lattice stencil).
void init_matrices(const DeviceVelocityGrid& vgrid, float** b, float** a, const Options& opts)
{
    g_ctx = vgrid.ctx;
    *b = *a = nullptr;
    CHECK(gdt_build_tables(vgrid.ctx, opts.tables ? GDT_TABLES_REFERENCE : GDT_TABLES_FP64), "gdt_build_tables");
    CHECK(gdt_synchronize(vgrid.ctx), "gdt_synchronize");
    const double *db = nullptr, *da = nullptr;
    long n_a = 0, nnz = 0;
    CHECK(gdt_tables(vgrid.ctx, &db, &da, &n_a, &nnz), "gdt_tables");
    if (!(*b = alloc_f32(opts.nxyz)) || !(*a = alloc_f32(n_a))) return;
    CHECK(gdt_convert_f64_to_f32(db, *b, opts.nxyz), "convert");
    CHECK(gdt_convert_f64_to_f32(da, *a, n_a), "convert");
    if (opts.method == 1) CHECK(gdt_korobov_build(vgrid.ctx, opts.korobov_M, opts.seed, 0, opts.s0), "gdt_korobov_build");
    g_tables[*a] = vgrid.ctx;
    std::printf("ALLOCATED %ld BYTES, FOR INVERSE COLISIONS MATRIX (%ld nonzero)\n", (long)(n_a * sizeof(float)), nnz);
}

// collision_integral.cu:550-773: loss frequency (corrected) and gain of every
// cell.  Tables not built by init_matrices are imported from b and a.
void evolve_colisions(float* f, float* direct_integral, float* inverse_integral, float* b, float* a,
                      float* correction_array, const Options& opts)
{
    (void)correction_array;  // the basis is recomputed in FP64 from the grid
    gdt_handle h = g_ctx;
    auto it = g_tables.find(a);
    if (it != g_tables.end()) {
        h = it->second;
    } else {
        if (!h) return fail("evolve_colisions: no grid (call init_correction_array / init_matrices first)");
        const size_t nb = opts.nxyz, na = (size_t)opts.nxyz * opts.nxyz / 8;
        double* db = s_nu.get(nb);
        double* da = s_gain.get(na);
        if (!db || !da) return;
        CHECK(gdt_convert_f32_to_f64(b, db, nb), "convert");
        CHECK(gdt_convert_f32_to_f64(a, da, na), "convert");
        CHECK(gdt_set_tables(h, db, da), "gdt_set_tables");
        g_tables[a] = h;
    }
    const int N = grid_points(h);
    if (N < 0) return;
    const size_t count = (size_t)N * opts.npx;
    double* f64 = s_f.get(count);
    double* nu = s_nu.get(count);
    double* gain = s_gain.get(count);
    double* Q = s_Q.get(count);
    if (!f64 || !nu || !gain || !Q) return;
    CHECK(gdt_convert_f32_to_f64(f, f64, count), "convert");
    const unsigned flags = opts.correction ? GDT_FLAG_CORRECT : 0u;
    CHECK(gdt_evaluate(h, opts.method == 1 ? GDT_METHOD_KOROBOV : GDT_METHOD_MATRIX, f64, nu, gain, Q, opts.npx,
                       GDT_LAYOUT_NODE_MAJOR, flags),
          "gdt_evaluate");
    CHECK(gdt_synchronize(h), "gdt_synchronize");
    CHECK(gdt_convert_f64_to_f32(nu, direct_integral, count), "convert");
    CHECK(gdt_convert_f64_to_f32(gain, inverse_integral, count), "convert");
}

// collision_integral.cu:776-880: dt = min courant / nu, f <- f (1 - nu dt) +
// gain dt, *time += dt.
void relax_f(float* f, float* frequency, float* inverse_integral, float courant, float* time, const Options& opts)
{
    gdt_handle h = g_ctx;
    if (!h) return fail("relax_f: no grid (call init_correction_array / init_matrices first)");
    const int N = grid_points(h);
    if (N < 0) return;
    const size_t count = (size_t)N * opts.npx;
    double* f64 = s_f.get(count);
    double* nu = s_nu.get(count);
    double* gain = s_gain.get(count);
    double* t64 = s_time.get(1);
    if (!f64 || !nu || !gain || !t64) return;
    CHECK(gdt_convert_f32_to_f64(f, f64, count), "convert");
    CHECK(gdt_convert_f32_to_f64(frequency, nu, count), "convert");
    CHECK(gdt_convert_f32_to_f64(inverse_integral, gain, count), "convert");
    CHECK(gdt_convert_f32_to_f64(time, t64, 1), "convert");
    CHECK(gdt_relax(h, f64, nu, gain, opts.npx, 0.0, courant, t64, GDT_FLAG_COURANT), "gdt_relax");
    CHECK(gdt_synchronize(h), "gdt_synchronize");
    CHECK(gdt_convert_f64_to_f32(f64, f, count), "convert");
    CHECK(gdt_convert_f64_to_f32(t64, time, 1), "convert");
}
