// Drop-in for the reference's common.hpp (common.hpp:26-35).
#ifndef GPU_DETERM_COMMON_HPP
#define GPU_DETERM_COMMON_HPP

struct IntParams
{
    int n_pnt;
    int n_u;
    int n_v;
    int n_w;
    int n_uvw;
};

template <typename T>
struct Vector3D
{
    Vector3D() : x(), y(), z() {}
    Vector3D(const T& x_, const T& y_, const T& z_) : x(x_), y(y_), z(z_) {}
    T x, y, z;
};

using Vector3i = Vector3D<int>;
using Vector3f = Vector3D<float>;

#endif
