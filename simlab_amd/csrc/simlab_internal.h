// Internal declarations shared by the host sources of libsimlab_cfd.so.
#pragma once

#include "../../include/simlab_cfd.h"

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

namespace simlab {

// D3Q19 velocity set in the reference's order (shader.rs:22-42):
// 0 rest; 1-6 the axes (+x,-x,+y,-y,+z,-z); 7-18 diagonals in +/- pairs.
constexpr int CV[19][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                           {0, 0, 1},   {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                           {-1, 0, -1}, {0, 1, 1},   {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                           {1, 0, -1},  {-1, 0, 1},  {0, 1, -1}, {0, -1, 1}};
constexpr int opp(int i) { return i == 0 ? 0 : (i % 2 == 1 ? i + 1 : i - 1); }
constexpr double weight(int i) { return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

// lattice.rs:172-189 (mlattice.rs:146-163): host rounding of a wall q to
// half (add bit 12, flush subnormals to signed zero)
inline uint16_t f16_bits_host(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    uint16_t sign = (uint16_t)((b >> 16) & 0x8000);
    int exp = (int)((b >> 23) & 0xff) - 127 + 15;
    uint32_t man = b & 0x7fffff;
    if (exp <= 0) return sign;
    if (exp >= 31) return sign | 0x7c00;
    uint32_t h = sign | ((uint32_t)exp << 10) | (man >> 13);
    if (man & 0x1000) h += 1;
    return (uint16_t)h;
}

struct Surface {
    std::vector<std::array<double, 3>> verts;
    std::vector<std::array<uint32_t, 3>> tris;
};

struct StaticVoxels {
    std::vector<uint32_t> solid;
    std::vector<simlab_wall_cell> walls;
    size_t lo[3] = {0, 0, 0};
    size_t hi[3] = {0, 0, 0};
};

struct SpinVoxels {
    double occ_lo[3] = {0, 0, 0};
    double occ_h = 1.0;
    size_t occ_dims[3] = {1, 1, 1};
    std::vector<uint32_t> occ_bits;
    std::vector<float> tris;   // ntri x 12
    double radius = 0.0;
    double axial_min = 0.0, axial_max = 0.0;
};

bool total_cmp_less(double x, double y);
bool seg_tri(const double o[3], const double d[3], const double *t0, const double *t1, const double *t2, double &tt);
std::vector<uint8_t> inside_grid_public(const Surface &s, const double origin[3], double h, const size_t dims[3]);
StaticVoxels voxelize_static(const Surface &s, const size_t dims[3]);
// the same voxelization with the inside test and the link search on the GPU
// (the current device); identical output
StaticVoxels voxelize_static_gpu(const Surface &s, const size_t dims[3]);
// the same voxelization with the D3Q27 links (voxel.rs:354-359)
StaticVoxels voxelize_static27(const Surface &s, const size_t dims[3]);
Surface subdivide(const Surface &s, double max_edge);
SpinVoxels voxelize_spin(const Surface &s, const double axis[3], double h);

// error reporting across the C ABI
void set_error(const std::string &msg);
inline int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}
template <class F> int guarded(F &&f) {
    try {
        set_error("");
        return f();
    } catch (const std::bad_alloc &) {
        return fail(SIMLAB_ERR_OOM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(SIMLAB_ERR_GPU, e.what());
    } catch (...) {
        return fail(SIMLAB_ERR_GPU, "unknown error");
    }
}

}  // namespace simlab

extern "C" {
typedef struct simlab_spin_voxels simlab_spin_voxels;
}
