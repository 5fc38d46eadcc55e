// Allocation-free Korobov evaluation: gdt_evaluate_noalloc.
//
// The same kernels as gdt_evaluate's Korobov method (node values, grouped
// gather, apply), with their scratch -- node values, group partials of Q and
// loss, moment partials -- carved out of `scratch`, a file-scope device array
// of NOALLOC_SCRATCH_BYTES per GPU, instead of context buffers.  The entry
// point calls no hipMalloc / hipFree: callers that cannot allocate (code
// under stream capture, serving loops) pass only their inputs and outputs.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address comes from hipGetSymbolAddress, called once per device
// (noalloc_setup, from gdt_create with that device current) and kept per
// device.  Every call on a GPU uses that GPU's one array: calls that overlap
// on one GPU (two streams, two threads) race on it.  No part of it needs to
// start at zero: each kernel writes every value it reads back.
#include <mutex>

#include "gdt_internal.h"

#define NOALLOC_SCRATCH_BYTES (64u << 20)  // 64 MiB per device
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

int noalloc_setup(int device)
{
    GDT_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, GDT_E_ARG, "noalloc: device %d beyond %d", device,
              NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return GDT_OK;
    void *p = nullptr;
    GDT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the caller made `device` current
    GDT_CHECK(p != nullptr, GDT_E_HIP, "noalloc: no device address for the scratch array on device %d", device);
    g_scratch_ptr[device] = p;
    return GDT_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" int gdt_evaluate_noalloc(gdt_handle h, const double *f, double *nu, double *gain, double *Q, int cells,
                                    unsigned flags)
{
    GDT_CHECK(h, GDT_E_ARG, "null gdt handle");
    GDT_CHECK(f && Q && cells > 0, GDT_E_ARG, "gdt_evaluate_noalloc: bad arguments");
    GDT_CHECK(!gain || nu, GDT_E_ARG, "gdt_evaluate_noalloc: gain (Q + f nu) needs the nu output");
    GDT_CHECK((flags & ~GDT_FLAG_CORRECT) == 0u, GDT_E_ARG,
              "gdt_evaluate_noalloc: flags %#x (only GDT_FLAG_CORRECT is accepted)", flags);
    GDT_CHECK(h->have_kor && h->have_stream, GDT_E_STATE, "gdt_evaluate_noalloc: stencil not built");
    GDT_CHECK(h->lds_attr >= kor_node_lds(h), GDT_E_STATE, "gdt_evaluate_noalloc: kernels not prepared");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = h->device >= 0 && h->device < NOALLOC_MAX_DEVICES ? g_scratch_ptr[h->device] : nullptr;
    }
    GDT_CHECK(base, GDT_E_STATE, "gdt_evaluate_noalloc: no scratch address for device %d", h->device);
    // node values (D, 2wP) per node and the lambda value slots, per cell; group
    // partials of Q and loss; moment sums per gather tile
    const int nparts = kor_gather_parts(h);
    const size_t vb = align256(sizeof(double) * 2 * (size_t)h->nv * cells);
    const size_t pb = align256(sizeof(double) * 2 * (size_t)KOR_NGROUPS * h->n_pnt * cells);
    const size_t mb = align256(sizeof(double) * 20 * (size_t)nparts * cells);
    const size_t per_cell = align256(sizeof(double) * 2 * (size_t)h->nv) +
                            align256(sizeof(double) * 2 * (size_t)KOR_NGROUPS * h->n_pnt) +
                            align256(sizeof(double) * 20 * (size_t)nparts);
    GDT_CHECK(vb + pb + mb <= NOALLOC_SCRATCH_BYTES, GDT_E_ARG,
              "gdt_evaluate_noalloc: %d cells need %zu bytes of scratch; the static scratch holds %u bytes "
              "(about %zu cells of this stencil)",
              cells, vb + pb + mb, NOALLOC_SCRATCH_BYTES, (size_t)NOALLOC_SCRATCH_BYTES / per_cell);
    GDT_HIP(hipSetDevice(h->device));
    char *s = static_cast<char *>(base);
    double *V = reinterpret_cast<double *>(s);
    double *Lp = reinterpret_cast<double *>(s + vb);
    double *Qp = Lp + (size_t)KOR_NGROUPS * cells * h->n_pnt;
    double *mp = reinterpret_cast<double *>(s + vb + pb);
    h->timing = false;  // timed launches would create events
    int rc = launch_kor_node(h, f, V, cells);
    if (rc) return rc;
    if ((rc = launch_kor_gather(h, V, Lp, Qp, mp, cells))) return rc;
    if ((rc = launch_kor_apply(h, (flags & GDT_FLAG_CORRECT) != 0u, KOR_NGROUPS, nparts, mp, f, Lp, Qp, Q, nullptr,
                               nu, nullptr, cells, 0.0, nullptr)))
        return rc;
    if (gain) return launch_correct(h, -3, f, nu, Q, nullptr, gain, nullptr, cells, 0.0, nullptr);
    return GDT_OK;
}
