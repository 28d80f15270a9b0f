// Interface between the refined D3Q27 host lattice (mlattice.cpp) and its
// gfx950 kernels (mkernels.hip). Plain structs passed by value as kernel
// arguments.
#pragma once

#include "../../include/simlab_cfd.h"

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace simlab {
namespace m27 {

constexpr uint32_t Q = 27, REST = 13;
// words of a wall record as `records()` gives them (mshader.rs:33-37)
constexpr uint32_t RS = 24, R_Q = 2, R_FT = 15;
constexpr uint32_t MW = 12;              // values per moments node (mshader.rs:38)
constexpr uint32_t HIST_CAP = 1024;      // finest steps per force read-back (mlattice.rs:222)
constexpr uint32_t RED_CHUNK = 4096;     // wall records per reduction block (mshader.rs:40)
// cell types, low 3 bits of a flags word (mshader.rs:44-48); the record
// index + 1 (or a face's type) in the bits above
constexpr uint32_t T_FLUID = 0, T_SOLID = 1, T_FACE = 2, T_GHOST = 3, T_RESTRICT = 4;

// Populations: two sets (A, B) per level, each with the cells in groups of
// 64 consecutive indices and a group's 27 directions one after the other:
// slot (i, n) of a set is element slot(i, n); set s starts at s * set_elems.
constexpr uint32_t GROUP = 64;
__host__ __device__ inline size_t slot(uint32_t i, uint32_t n) {
    return (size_t)(n / GROUP) * (Q * GROUP) + (size_t)i * GROUP + (n % GROUP);
}
__host__ __device__ inline size_t set_elems(uint64_t ncell) {
    return (size_t)Q * ((ncell + GROUP - 1) / GROUP) * GROUP;
}

// Device view of one level (the constants of mshader.rs `LevelShader`).
struct DevLevel {
    uint32_t nx, ny, nz;
    uint32_t n;
    int32_t storage;                 // simlab_storage: FP32, FP16 or FP64
    uint32_t per[3];
    void *pops;                      // 2 * set_elems(n) elements
    uint32_t *flags;
    const uint32_t *lists;           // face cells, then ghosts, then restricted cells
    uint32_t nface, nghost, nrestrict;
    double tau0, smag2, sponge_tau, sigma;
    double sp_lo[3], sp_hi[3];
    double u_inf[3], u_init[3], grav[3], center[3];
    int32_t absorb, total_p, n_sponge, forcing;
    int32_t has_child, has_parent;
    int32_t ch_lo[3];
    uint32_t ch_n[3];
    uint32_t ch_nodes;
    uint32_t pn[3];                  // this level's nodes seen from its parent
    uint32_t pnn;
    void *mom_child;                 // [2][ch_nodes][MW]: moments on the child box surface
    void *mom_up;                    // [pnn][MW]: moments of the even nodes, for the parent
    const void *mom_parent;          // the parent's mom_child
    const void *mom_child_up;        // the child's mom_up
    // wall records (finest level), SoA: field j of record r at j * nrec + r
    uint32_t nrec;
    const uint32_t *rmask;           // [nrec]
    const uint32_t *rq;              // [13][nrec] f16 pairs by moving-direction slot
    void *rft;                       // [6][nrec] F, T in the arithmetic type
    // slices owned by this level: (pixel, cell) pairs, accumulators shared by all levels
    uint32_t npix;
    const uint32_t *spix;
    float *sacc;                     // [pixels][5]
    float4 *dump;                    // [n] (rho, u) for macros(), may be null
};

// The per-dispatch values of mlattice.rs `StepParams`.
struct StepArgs {
    uint32_t src;                    // the population set the step reads
    uint32_t mw;                     // moments slot the step writes
    uint32_t mold, mnew;             // ghosts: the parent's moments slots
    float blend;                     // ghosts: time interpolation weight
    uint32_t seed;                   // FP16 stochastic rounding seed
    double omega[3];                 // body rotation rate (rad/step)
    double rot[3][3];                // body axes (columns) in lattice coordinates
};

hipError_t m_init(const DevLevel &L, hipStream_t st);
hipError_t m_step(const DevLevel &L, const StepArgs &a, hipStream_t st);
hipError_t m_face(const DevLevel &L, const StepArgs &a, hipStream_t st);
hipError_t m_ghost(const DevLevel &L, const StepArgs &a, hipStream_t st);
hipError_t m_restrict(const DevLevel &L, const StepArgs &a, hipStream_t st);
// records' F and T of one finest step summed in FP64: partials[block][6]
hipError_t m_reduce1(const DevLevel &L, double *partials, hipStream_t st);
// per step of a chunk: hist[k][0..6] = sum over blocks of partials[k][block][6]
hipError_t m_reduce2(const double *partials, uint32_t nblk, uint32_t nsteps, double *hist, hipStream_t st);
hipError_t m_dump(const DevLevel &L, uint32_t src, hipStream_t st);
hipError_t m_sample_verts(const DevLevel &L, const StepArgs &a, float *probes, uint32_t nvert, double voff,
                          hipStream_t st);
hipError_t m_sample_slices(const DevLevel &L, uint32_t src, hipStream_t st);
// the source set decoded to f - w, reference order [27][n] doubles
hipError_t m_get_pops(const DevLevel &L, uint32_t src, double *out, hipStream_t st);

}  // namespace m27
}  // namespace simlab
