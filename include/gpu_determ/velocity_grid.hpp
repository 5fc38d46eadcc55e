// Drop-in for the reference's velocity_grid.hpp (velocity_grid.hpp:6-54):
// the same members and methods.  DeviceVelocityGrid's arrays live in device
// memory; it also carries the libgdt context of the grid (`ctx`), which the
// collision functions use.
#ifndef GPU_DETERM_VELOCITY_GRID_HPP
#define GPU_DETERM_VELOCITY_GRID_HPP

#include "common.hpp"

struct gdt_context;

class VelocityGridBase
{
public:
    virtual ~VelocityGridBase() {}

    int n_pnt = 0;

    int n_u = 0;
    int n_w = 0;
    int n_v = 0;

    float* u = 0;
    float* v = 0;
    float* w = 0;

    int* u_index = 0;
    int* v_index = 0;
    int* w_index = 0;

    int n_uvw = 0;

    float min_u, min_v, min_w;
    float max_u, max_v, max_w;
    float d_u, d_v, d_w;
    float d3v;

    float R = 0.0f;  // not in the reference: the cut radius init() was given
};

class DeviceVelocityGrid;

class VelocityGrid : public VelocityGridBase
{
public:
    VelocityGrid() {}
    VelocityGrid(Vector3i n_points, Vector3f v_min, Vector3f v_max, float r);
    virtual ~VelocityGrid();

    void init(Vector3i n_points, Vector3f v_min, Vector3f v_max, float R);
    DeviceVelocityGrid& device() const;
};

class DeviceVelocityGrid : public VelocityGridBase
{
    friend class VelocityGrid;
    DeviceVelocityGrid(const VelocityGrid&);

public:
    // can't use a destructor: copies are passed around by value
    void free();
    gdt_context* ctx = nullptr;
};

#endif
