// VelocityGrid / DeviceVelocityGrid of the drop-in (velocity_grid.cu:6-126):
// host node centres and sphere-cut index lists in FP32 as the reference
// builds them; the device mirror holds the same arrays in device memory and
// the libgdt context of the grid.
#include <cassert>
#include <cstdio>
#include <stdexcept>

#include "../../../include/gdt.h"
#include "velocity_grid.hpp"

namespace {

// velocity_grid.cu:6-14: centre measured from the middle of the box
bool in_domain(Vector3i n, Vector3f lo, Vector3f hi, int i, int j, int k, float R)
{
    const float u = (hi.x - lo.x) * (2 * i + 1 - n.x) / 2. / n.x;
    const float v = (hi.y - lo.y) * (2 * j + 1 - n.y) / 2. / n.y;
    const float w = (hi.z - lo.z) * (2 * k + 1 - n.z) / 2. / n.z;
    return R <= 0 || u * u + v * v + w * w < R * R;
}

template <class T>
T* to_device(const T* host, int count)
{
    void* p = nullptr;
    if (gdt_malloc(&p, sizeof(T) * (count > 0 ? count : 1)) != GDT_OK ||
        gdt_memcpy_htod(p, host, sizeof(T) * count) != GDT_OK)
        throw std::runtime_error(std::string("device grid upload failed: ") + gdt_last_error());
    return static_cast<T*>(p);
}

}  // namespace

VelocityGrid::VelocityGrid(Vector3i n_points, Vector3f v_min, Vector3f v_max, float r)
{
    init(n_points, v_min, v_max, r);
}

DeviceVelocityGrid& VelocityGrid::device() const
{
    static DeviceVelocityGrid g(*this);
    return g;
}

DeviceVelocityGrid::DeviceVelocityGrid(const VelocityGrid& g) : VelocityGridBase(g)
{
    u_index = to_device(g.u_index, n_pnt);
    v_index = to_device(g.v_index, n_pnt);
    w_index = to_device(g.w_index, n_pnt);
    u = to_device(g.u, n_u);
    v = to_device(g.v, n_v);
    w = to_device(g.w, n_w);
    gdt_grid_desc d;
    d.n[0] = g.n_u;
    d.n[1] = g.n_v;
    d.n[2] = g.n_w;
    d.vmin[0] = g.min_u;
    d.vmin[1] = g.min_v;
    d.vmin[2] = g.min_w;
    d.vmax[0] = g.max_u;
    d.vmax[1] = g.max_v;
    d.vmax[2] = g.max_w;
    d.R = g.R;
    gdt_handle h = nullptr;
    if (gdt_create(&h, &d, 0) != GDT_OK)
        throw std::runtime_error(std::string("gdt_create failed: ") + gdt_last_error());
    ctx = h;
}

VelocityGrid::~VelocityGrid()
{
    if (!u) return;
    delete[] u;
    delete[] v;
    delete[] w;
    delete[] u_index;
    delete[] v_index;
    delete[] w_index;
    device().free();
}

void DeviceVelocityGrid::free()
{
    gdt_free(u);
    gdt_free(v);
    gdt_free(w);
    gdt_free(u_index);
    gdt_free(v_index);
    gdt_free(w_index);
    u = v = w = nullptr;
    u_index = v_index = w_index = nullptr;
    if (ctx) gdt_destroy(ctx);
    ctx = nullptr;
}

// velocity_grid.cu:73-126
void VelocityGrid::init(Vector3i n_points, Vector3f v_min, Vector3f v_max, float Rc)
{
    assert(n_points.x > 0 && n_points.y > 0 && n_points.z > 0);
    assert(v_min.x < v_max.x && v_min.y < v_max.y && v_min.z < v_max.z);
    n_pnt = 0;
    for (int i = 0; i < n_points.x; i++)
        for (int j = 0; j < n_points.y; j++)
            for (int k = 0; k < n_points.z; k++) n_pnt += in_domain(n_points, v_min, v_max, i, j, k, Rc);
    n_u = n_points.x;
    n_v = n_points.y;
    n_w = n_points.z;
    n_uvw = n_u * n_v * n_w;
    R = Rc;
    u = new float[n_u];
    v = new float[n_v];
    w = new float[n_w];
    u_index = new int[n_pnt];
    v_index = new int[n_pnt];
    w_index = new int[n_pnt];
    min_u = v_min.x;
    min_v = v_min.y;
    min_w = v_min.z;
    max_u = v_max.x;
    max_v = v_max.y;
    max_w = v_max.z;
    d_u = (max_u - min_u) / n_u;
    d_v = (max_v - min_v) / n_v;
    d_w = (max_w - min_w) / n_w;
    d3v = d_u * d_v * d_w;
    for (int i = 0; i < n_u; i++) u[i] = min_u + d_u * (2. * i + 1.) / 2.;
    for (int i = 0; i < n_v; i++) v[i] = min_v + d_v * (2. * i + 1.) / 2.;
    for (int i = 0; i < n_w; i++) w[i] = min_w + d_w * (2. * i + 1.) / 2.;
    int idx = 0;
    for (int i = 0; i < n_u; i++)
        for (int j = 0; j < n_v; j++)
            for (int k = 0; k < n_w; k++)
                if (in_domain(n_points, v_min, v_max, i, j, k, Rc)) {
                    u_index[idx] = i;
                    v_index[idx] = j;
                    w_index[idx] = k;
                    idx++;
                }
}
