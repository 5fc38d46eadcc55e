// Uneven batches of the Korobov method (gdt_batch_* in include/gdt.h): many
// cells on several velocity grids, one corrected evaluation (or explicit
// step) of every cell per run step, as three work-list kernels per segment of
// cells, or (GDT_FLAG_CELL_TILES / GDT_FLAG_MIXED) each class's cells in cell
// tiles; the whole step captured as one hipGraph.
//
// A segment holds up to seg_cells cells (of any grids, in class order); its
// scratch (node values, group partials, moment sums) is reused by the next
// segment.  The tuner times candidate (segment size, nodes per thread of the
// node kernel) pairs on the caller's cells and keeps the fastest; neither knob
// changes any value (node values do not depend on the node blocking, and the
// gather tiles / apply blocks are the class's own), so every cell's result is
// bitwise the result of a non-batched run of that cell.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "gdt_internal.h"

struct gdt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<gdt_context *> cls;
    std::vector<unsigned> cls_gen;  // the classes' gather layouts the work lists were built for
    std::vector<double *> f, Q;
    std::vector<int> cells;
    int lds_node = 0, lds_gather = 0, cu = 256;

    // configuration
    int seg_cells = 0, node_q = 0;
    struct Segment {
        std::vector<KorClassDev> hdesc;  // scratch fields hold offsets until rebased
        DevBuf desc, node_items, gather_items, apply_items;
        int n_node = 0, n_gather = 0, n_apply = 0;
        int b_node = 0, b_gather = 0, b_apply = 0;
    };
    std::vector<Segment> segs;
    DevBuf scratch;

    // graph of one run step
    hipGraphExec_t graph = nullptr;
    double g_dt = 0.0;
    unsigned g_flags = 0;
    unsigned long g_gen = 0;
    bool g_valid = false;
};

#define B_CHECK(b)                                                  \
    do {                                                            \
        if (!(b)) {                                                 \
            gdt_set_error(GDT_E_ARG, "null gdt batch handle");      \
            return GDT_E_ARG;                                       \
        }                                                           \
    } while (0)

static void drop_graph(gdt_batch *b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->g_valid = false;
}

static void drop_segments(gdt_batch *b)
{
    for (auto &sg : b->segs) {
        sg.desc.release();
        sg.node_items.release();
        sg.gather_items.release();
        sg.apply_items.release();
    }
    b->segs.clear();
}

// Segments, work lists and scratch for (seg_cells, node_q).
static int batch_configure(gdt_batch *b, int seg_cells, int node_q)
{
    GDT_CHECK(seg_cells > 0 && node_q >= 2, GDT_E_ARG, "gdt_batch: segment cells %d, nodes per thread %d", seg_cells,
              node_q);
    drop_graph(b);
    const int ncls = (int)b->cls.size();
    // cells in class order, cut into segments
    std::vector<std::pair<int, int>> order;
    for (int s = 0; s < ncls; s++)
        for (int c = 0; c < b->cells[s]; c++) order.push_back({s, c});
    const int total = (int)order.size();
    const int nseg = (total + seg_cells - 1) / seg_cells;
    drop_segments(b);
    b->segs.resize(nseg);
    size_t scratch_max = 0;
    std::vector<int> node_blocks(ncls);
    for (int s = 0; s < ncls; s++) {
        const long per = (long)node_q * 1024;
        node_blocks[s] = (int)std::max(1L, (b->cls[s]->nn + per - 1) / per);
    }
    for (int j = 0; j < nseg; j++) {
        auto &sg = b->segs[j];
        const int lo = j * seg_cells, hi = std::min(total, lo + seg_cells);
        // slots per class in this segment
        std::vector<int> nslot(ncls, 0);
        for (int i = lo; i < hi; i++) nslot[order[i].first]++;
        // scratch layout: per class V, Lp, Qp, mp regions (byte offsets, 256-aligned)
        std::vector<size_t> oV(ncls), oL(ncls), oQ(ncls), oM(ncls);
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off += (bytes + 255) & ~(size_t)255;
            return o;
        };
        for (int s = 0; s < ncls; s++) {
            const gdt_context *h = b->cls[s];
            const size_t n = (size_t)nslot[s];
            oV[s] = take(n * h->nv * sizeof(double2));
            oL[s] = take(n * KOR_NGROUPS * h->n_pnt * sizeof(double));
            oQ[s] = take(n * KOR_NGROUPS * h->n_pnt * sizeof(double));
            oM[s] = take(n * 20 * KOR_NGROUPS * h->kntile * sizeof(double));
        }
        scratch_max = std::max(scratch_max, off);
        // work lists
        std::vector<int4> ni, gi, ai;
        std::vector<int> slot(ncls, 0);
        for (int i = lo; i < hi; i++) {
            const int s = order[i].first, c = order[i].second;
            const gdt_context *h = b->cls[s];
            const int sl = slot[s]++;
            for (int k = 0; k < node_blocks[s]; k++) ni.push_back(make_int4(s, c, k, sl));
            for (int k = 0; k < KOR_NGROUPS * h->kntile; k++) gi.push_back(make_int4(s, c, k, sl));
            for (int k = 0; k < (h->n_pnt + KOR_APPLY_THREADS - 1) / KOR_APPLY_THREADS; k++)
                ai.push_back(make_int4(s, c, k, sl));
        }
        sg.n_node = (int)ni.size();
        sg.n_gather = (int)gi.size();
        sg.n_apply = (int)ai.size();
        // grids: one block per work item (the gather's rounded up to a
        // multiple of the group count, so item i stays on XCD i % G).  The
        // kernels keep the grid-stride form; a persistent grid of a few blocks
        // per CU, walking the list in a static order, balanced the uneven item
        // costs worse than the hardware's block dispatch (8,192 cells of 24^3:
        // 101-110 k against 133 k evals/s)
        sg.b_node = std::max(1, sg.n_node);
        sg.b_gather = std::max(KOR_NGROUPS, (sg.n_gather + KOR_NGROUPS - 1) / KOR_NGROUPS * KOR_NGROUPS);
        sg.b_apply = std::max(1, sg.n_apply);
        GDT_CHECK(sg.node_items.ensure(sizeof(int4) * ni.size()) == 0 &&
                      sg.gather_items.ensure(sizeof(int4) * gi.size()) == 0 &&
                      sg.apply_items.ensure(sizeof(int4) * ai.size()) == 0,
                  GDT_E_NOMEM, "gdt_batch: work-list allocation failed");
        GDT_HIP(hipMemcpy(sg.node_items.p, ni.data(), sizeof(int4) * ni.size(), hipMemcpyHostToDevice));
        GDT_HIP(hipMemcpy(sg.gather_items.p, gi.data(), sizeof(int4) * gi.size(), hipMemcpyHostToDevice));
        GDT_HIP(hipMemcpy(sg.apply_items.p, ai.data(), sizeof(int4) * ai.size(), hipMemcpyHostToDevice));
        // class descriptors; scratch fields hold byte offsets until rebased
        std::vector<KorClassDev> &d = sg.hdesc;
        d.assign(ncls, KorClassDev());
        for (int s = 0; s < ncls; s++) {
            const gdt_context *h = b->cls[s];
            KorClassDev &C = d[s];
            C.npts = h->n_pnt;
            C.ntile = h->kntile;
            C.cap = h->kcap;
            C.nparts = KOR_NGROUPS * h->kntile;
            C.nn = h->nn;
            C.K = h->kK;
            C.nrec = h->kn4.as<uint4>();
            C.tiles = h->ktile.as<int4>();
            C.gptr = h->kgptr.as<int32_t>();
            C.ent = h->kgent.as<uint2>();
            C.kmt = h->kmath.as<double>();
            C.xi = h->xi.as<double>();
            C.f = b->f[s];
            C.Qout = b->Q.empty() ? nullptr : b->Q[s];
            C.V = (double2 *)(uintptr_t)oV[s];
            C.Lp = (double *)(uintptr_t)oL[s];
            C.Qp = (double *)(uintptr_t)oQ[s];
            C.mp = (double *)(uintptr_t)oM[s];
            C.node_blocks = node_blocks[s];
            C.pad = 0;
            C.nv = h->nv;
            C.cbase = h->kcbase.as<int32_t>();
        }
    }
    GDT_CHECK(b->scratch.ensure(scratch_max) == 0, GDT_E_NOMEM, "gdt_batch: scratch allocation failed (%zu bytes)",
              scratch_max);
    // rebase the scratch offsets onto the scratch buffer and upload
    const uintptr_t base = (uintptr_t)b->scratch.p;
    for (auto &sg : b->segs) {
        std::vector<KorClassDev> d = sg.hdesc;
        for (auto &C : d) {
            C.V = (double2 *)(base + (uintptr_t)C.V);
            C.Lp = (double *)(base + (uintptr_t)C.Lp);
            C.Qp = (double *)(base + (uintptr_t)C.Qp);
            C.mp = (double *)(base + (uintptr_t)C.mp);
        }
        GDT_CHECK(sg.desc.ensure(sizeof(KorClassDev) * ncls) == 0, GDT_E_NOMEM, "gdt_batch: descriptor allocation failed");
        GDT_HIP(hipMemcpy(sg.desc.p, d.data(), sizeof(KorClassDev) * ncls, hipMemcpyHostToDevice));
    }
    b->seg_cells = seg_cells;
    b->node_q = node_q;
    // dynamic LDS of the launches, from the classes' current layouts (a
    // re-tiling changes the gather's entries per pass)
    b->lds_node = b->lds_gather = 0;
    b->cls_gen.resize(ncls);
    for (int s = 0; s < ncls; s++) {
        b->lds_node = std::max(b->lds_node, kor_node_lds(b->cls[s]));
        b->lds_gather = std::max(b->lds_gather, (int)sizeof(double2) * b->cls[s]->kcap);
        b->cls_gen[s] = b->cls[s]->kor_gen;
    }
    return GDT_OK;
}

// Launches of a class context redirected onto the batch stream for the
// scope (the cell-tile launchers use the context's stream).
struct ClassStream {
    gdt_context *h;
    hipStream_t old;
    ClassStream(gdt_context *h_, hipStream_t s) : h(h_), old(h_->stream) { h->stream = s; }
    ~ClassStream() { h->stream = old; }
};

static bool batch_tiles(unsigned flags) { return (flags & (GDT_FLAG_CELL_TILES | GDT_FLAG_MIXED)) != 0u; }

// Cell-tile scratch of every class for all its cells (outside any capture).
static int batch_tiles_prepare(gdt_batch *b, unsigned flags)
{
    for (size_t s = 0; s < b->cls.size(); s++)
        if (b->cells[s] > 0)
            if (int rc = ct_prepare(b->cls[s], b->cells[s], (flags & GDT_FLAG_MIXED) != 0u)) return rc;
    return GDT_OK;
}

// One run step over every segment on the batch stream: the work lists, or,
// with GDT_FLAG_CELL_TILES / GDT_FLAG_MIXED, every class's cells in cell
// tiles (each class's stencil read once per tile of 64 cells).
static int batch_step(gdt_batch *b, double dt, unsigned flags)
{
    const bool corr = (flags & GDT_FLAG_CORRECT) != 0;
    if (batch_tiles(flags)) {
        for (size_t s = 0; s < b->cls.size(); s++) {
            if (b->cells[s] == 0) continue;
            ClassStream cs(b->cls[s], b->stream);
            int rc = ct_run(b->cls[s], (flags & GDT_FLAG_MIXED) != 0u, b->f[s], b->cells[s], corr,
                            b->Q.empty() ? nullptr : b->Q[s], nullptr, nullptr, dt, dt != 0.0 ? b->f[s] : nullptr);
            if (rc) return rc;
        }
        return GDT_OK;
    }
    for (auto &sg : b->segs) {
        const KorClassDev *d = sg.desc.as<KorClassDev>();
        int rc;
        if ((rc = launch_batch_node(b->stream, d, sg.node_items.as<int4>(), sg.n_node, sg.b_node, b->lds_node)))
            return rc;
        if ((rc = launch_batch_gather(b->stream, d, sg.gather_items.as<int4>(), sg.n_gather, sg.b_gather,
                                      b->lds_gather)))
            return rc;
        if ((rc = launch_batch_apply(b->stream, corr, d, sg.apply_items.as<int4>(), sg.n_apply, sg.b_apply, dt)))
            return rc;
    }
    return GDT_OK;
}

int gdt_batch_create(gdt_batch_handle *out, int nclass, const gdt_handle *classes, double *const *f,
                     const int *cells)
{
    GDT_CHECK(out && nclass > 0 && classes && f && cells, GDT_E_ARG, "gdt_batch_create: bad arguments");
    *out = nullptr;
    gdt_batch *b = new gdt_batch();
    for (int s = 0; s < nclass; s++) {
        gdt_context *h = classes[s];
        if (!h || !h->have_kor || !h->have_stream || !f[s] || cells[s] < 0) {
            delete b;
            gdt_set_error(GDT_E_ARG, "gdt_batch_create: class %d has no Korobov stencil, f or cell count", s);
            return GDT_E_ARG;
        }
        if (s > 0 && h->device != classes[0]->device) {
            delete b;
            gdt_set_error(GDT_E_ARG, "gdt_batch_create: classes on different devices");
            return GDT_E_ARG;
        }
        b->cls.push_back(h);
        b->f.push_back(f[s]);
        b->cells.push_back(cells[s]);
    }
    b->device = classes[0]->device;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        (void)hipGetLastError();
        gdt_batch_destroy(b);
        gdt_set_error(GDT_E_HIP, "gdt_batch_create: stream/event creation failed");
        return GDT_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cu = prop.multiProcessorCount;
    int rc = batch_set_lds();
    if (rc) {
        gdt_batch_destroy(b);
        return rc;
    }
    *out = b;
    return GDT_OK;
}

int gdt_batch_destroy(gdt_batch_handle b)
{
    if (!b) return GDT_OK;
    drop_graph(b);
    drop_segments(b);
    b->scratch.release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return GDT_OK;
}

int gdt_batch_set_output(gdt_batch_handle b, double *const *Q)
{
    B_CHECK(b);
    b->Q.clear();
    if (Q)
        for (size_t s = 0; s < b->cls.size(); s++) b->Q.push_back(Q[s]);
    if (b->seg_cells) return batch_configure(b, b->seg_cells, b->node_q);
    return GDT_OK;
}

// A class re-tiled or rebuilt after the work lists were made would leave them
// pointing at its old layout.
static int batch_check_classes(const gdt_batch *b)
{
    for (size_t s = 0; s < b->cls.size() && s < b->cls_gen.size(); s++)
        GDT_CHECK(b->cls[s]->kor_gen == b->cls_gen[s], GDT_E_STATE,
                  "gdt_batch: class %zu changed its stencil or tiles after the batch was configured "
                  "(call gdt_batch_config again)", s);
    return GDT_OK;
}

int gdt_batch_config(gdt_batch_handle b, int seg_cells, int node_q)
{
    B_CHECK(b);
    return batch_configure(b, seg_cells, node_q);
}

int gdt_batch_tune(gdt_batch_handle b, unsigned flags, int *seg_cells, int *node_q, double *best_ms)
{
    B_CHECK(b);
    int total = 0;
    for (int c : b->cells) total += c;
    GDT_CHECK(total > 0, GDT_E_ARG, "gdt_batch_tune: no cells");
    // segment candidates up to 1024 cells (the scratch grows with the segment:
    // 1024 cells of 32^3 at 2^21 nodes hold ~5 GB of node values and partials)
    std::vector<int> segs;
    for (int k : {64, 256, 1024})
        if (k < total) segs.push_back(k);
    if (total <= 1024) segs.push_back(total);
    // the classes' gather tile limit first, at the largest segment and 32
    // nodes per thread (the tile count sets the moment partials per cell and
    // the per-tile work: fewer, larger tiles pay off when many cells share the
    // GPU), then (segment, nodes per thread) at the chosen limit
    auto time_step = [&](float &t) -> int {
        t = 1e30f;
        for (int rep = 0; rep < 2; rep++) {  // evaluation only (dt = 0): f is not changed
            GDT_HIP(hipEventRecord(b->ev0, b->stream));
            if (int rc = batch_step(b, 0.0, flags)) return rc;
            GDT_HIP(hipEventRecord(b->ev1, b->stream));
            GDT_HIP(hipEventSynchronize(b->ev1));
            float ms = 0.0f;
            GDT_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            t = std::min(t, ms);
        }
        return GDT_OK;
    };
    if (batch_tiles(flags)) {
        // cell tiles: the classes' launch shape -- tiles whose node terms are
        // in flight, node and gather grids -- set on every class as
        // gdt_korobov_cell_tiles does; the tile width stays (it sets the
        // order of the moment sums), so results do not change
        drop_graph(b);
        float best = 1e30f;
        int bc = 4, bn = 4096, bg = 4096;
        for (int chunk : {4, 8, 16})
            for (int nb : {2048, 4096})
                for (int gb : {2048, 4096}) {
                    for (gdt_context *c : b->cls)
                        if (int rc = gdt_korobov_cell_tiles(c, chunk, nb, gb, c->ct_wide)) return rc;
                    if (int rc = batch_tiles_prepare(b, flags)) return rc;
                    float t;
                    if (int rc = time_step(t)) return rc;
                    if (t < best) {
                        best = t;
                        bc = chunk;
                        bn = nb;
                        bg = gb;
                    }
                }
        for (gdt_context *c : b->cls)
            if (int rc = gdt_korobov_cell_tiles(c, bc, bn, bg, c->ct_wide)) return rc;
        if (seg_cells) *seg_cells = bc * ct_cells_per_tile(b->cls[0], (flags & GDT_FLAG_MIXED) != 0u);
        if (node_q) *node_q = bg;
        if (best_ms) *best_ms = best;
        return GDT_OK;
    }
    float best_t = 1e30f;
    int bt = KOR_TILES_MAX;
    for (int tl : {32, 64, 128}) {
        for (gdt_context *c : b->cls)
            if (int rc = gdt_korobov_set_tiles(c, tl)) return rc;
        if (int rc = batch_configure(b, segs.back(), 32)) return rc;
        float t;
        if (int rc = time_step(t)) return rc;
        if (t < best_t) {
            best_t = t;
            bt = tl;
        }
    }
    for (gdt_context *c : b->cls)
        if (int rc = gdt_korobov_set_tiles(c, bt)) return rc;
    float best = 1e30f;
    int bs = segs.back(), bq = 2;
    for (int k : segs)
        for (int q : {2, 8, 32, 128}) {
            int rc = batch_configure(b, k, q);
            if (rc) return rc;
            float t;
            if ((rc = time_step(t))) return rc;
            if (t < best) {
                best = t;
                bs = k;
                bq = q;
            }
        }
    int rc = batch_configure(b, bs, bq);
    if (rc) return rc;
    if (seg_cells) *seg_cells = bs;
    if (node_q) *node_q = bq;
    if (best_ms) *best_ms = best;
    return GDT_OK;
}

int gdt_batch_run(gdt_batch_handle b, double dt, int steps, unsigned flags)
{
    B_CHECK(b);
    GDT_CHECK(steps >= 0, GDT_E_ARG, "gdt_batch_run: steps %d", steps);
    if (batch_tiles(flags)) {
        if (int rc = batch_tiles_prepare(b, flags)) return rc;
    } else {
        if (!b->seg_cells) {
            int rc = batch_configure(b, 256, 2);
            if (rc) return rc;
        }
        if (int rc = batch_check_classes(b)) return rc;
    }
    if (!(flags & GDT_FLAG_GRAPH)) {
        for (int s = 0; s < steps; s++) {
            int rc = batch_step(b, dt, flags);
            if (rc) return rc;
        }
        return GDT_OK;
    }
    // the captured launches hold the scratch pointers of the work lists or of
    // the classes' cell tiles: recapture after any buffer was freed
    const unsigned long gen = devbuf_generation();
    if (!b->g_valid || b->g_dt != dt || b->g_flags != flags || b->g_gen != gen) {
        drop_graph(b);
        hipGraph_t g = nullptr;
        GDT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_step(b, dt, flags);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        GDT_HIP(e);
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GDT_HIP(e);
        b->g_dt = dt;
        b->g_flags = flags;
        b->g_gen = devbuf_generation();
        b->g_valid = true;
    }
    for (int s = 0; s < steps; s++) GDT_HIP(hipGraphLaunch(b->graph, b->stream));
    return GDT_OK;
}

int gdt_batch_synchronize(gdt_batch_handle b)
{
    B_CHECK(b);
    GDT_HIP(hipStreamSynchronize(b->stream));
    return GDT_OK;
}

int gdt_batch_info(gdt_batch_handle b, int *seg_cells, int *node_q, int *segments)
{
    B_CHECK(b);
    if (seg_cells) *seg_cells = b->seg_cells;
    if (node_q) *node_q = b->node_q;
    if (segments) *segments = (int)b->segs.size();
    return GDT_OK;
}
