// Device side of the GPU static voxelizer (voxel_gpu.hip), called by the
// host orchestration in voxel.cpp. Plain pointers to device memory.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace simlab {

// most crossings one scan line keeps; a line with more sets *overflow
constexpr int GVOX_MAX_LINE_HITS = 64;

// Scan lines along axis a of a grid of gdims points: one thread per line
// (j along b = (a + 1) % 3, k along c = (a + 2) % 3). gv: grid-frame vertex
// coordinates (nverts x 3 doubles); tris: ntris x 3 vertex indices;
// row_start / row_list: the triangles whose c-range meets row k (CSR).
// ORs the vote bits of every grid point into votes.
hipError_t gvox_parity(const double *gv, const uint32_t *tris, const uint32_t *row_start, const uint32_t *row_list,
                       int a, const uint64_t gdims[3], uint32_t *votes, uint32_t *overflow, hipStream_t s);

// inside[l] = majority of the valid votes of point l
hipError_t gvox_majority(const uint32_t *votes, uint64_t n, uint8_t *inside, hipStream_t s);

// A wall cell as the link kernel finds it: grid point l, link mask, q per
// direction and, for links through a sheet thinner than a cell, the q of
// the reciprocal link at the neighbour (-1 where none).
struct GvoxCell {
    uint64_t l;
    uint32_t mask;
    uint32_t pad;
    float q[19];
    float thin_q[19];
};

struct GvoxLinkArgs {
    const double *verts;         // nverts x 3, lattice frame
    const uint32_t *tris;        // ntris x 3
    const uint8_t *inside;       // per grid point
    const uint32_t *bin_start;   // CSR of triangles per unit bin
    const uint32_t *bin_list;
    double bin_lo[3];
    uint64_t bin_dims[3];
    uint64_t gd[3];              // examined grid (points)
    int64_t slo[3];              // its origin in lattice cells
    uint64_t dims[3];            // lattice
    GvoxCell *out;               // capacity entries
    uint64_t capacity;
    unsigned long long *count;   // wall cells found (may exceed capacity)
};

// one thread per grid point: the D3Q19 wall links of the fluid points
hipError_t gvox_links(const GvoxLinkArgs &a, hipStream_t s);

}  // namespace simlab
