// Interface between the host lattice (lattice.cpp) and the gfx950 kernels
// (kernels.hip). Plain structs passed by value as kernel arguments.
#pragma once

#include "../../include/simlab_cfd.h"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>

namespace simlab {

constexpr uint32_t KMAX = 512;   // steps per device-to-host force copy (lattice.rs:134)
constexpr uint32_t HIST_STRIDE = 8;   // doubles per step in the force history
constexpr uint32_t SPECIAL_BLOCK = 256;
// flags bit 31: a wall cell whose cut links are all halfway (q = 1/2) and
// handled without a record; bits 2-19 then hold its link mask (direction i
// at bit i + 1). Otherwise bits 2-30 hold the record index + 1.
constexpr uint32_t FLAG_HALFWAY = 1u << 31;
// dilation of the near-part grid of a spinning part, in lattice spacings
// (a link hit needs a triangle within 1.415 * sqrt(3) = 2.45 cells, a solid
// neighbour or a refill the body within sqrt(2)); a part moves less than a
// cell per step at any stable tip speed
constexpr int NEAR_CELLS = 5;

// Populations: cells (index n = (z * ny + y) * nx + x) in groups of
// POP_GROUP consecutive indices; a group keeps its 19 directions one after
// the other, so slot (i, n) is element pop_index(i, n) of `pops`. The
// reference's order (shader.rs:147-152, slot i * ncell + n) is what the
// population read/write entry points exchange with the caller.
constexpr uint32_t POP_GROUP = 64;
__host__ __device__ inline size_t pop_index(uint32_t i, uint32_t n) {
    return (size_t)(n / POP_GROUP) * (19 * POP_GROUP) + (size_t)i * POP_GROUP + (n % POP_GROUP);
}
// elements of the population buffer of a lattice of ncell cells
__host__ __device__ inline size_t pop_elems(uint64_t ncell) {
    return (size_t)19 * ((ncell + POP_GROUP - 1) / POP_GROUP) * POP_GROUP;
}

// Device view of one lattice. Wall records are SoA: field j of record r is
// element j * nrec + r.
struct DevLattice {
    uint32_t nx, ny, nz;
    uint32_t ncell;
    uint32_t per[3];
    uint32_t vel_face[6];        // per face: 1 for a velocity face
    int32_t storage;             // simlab_storage

    void *pops;
    uint32_t *flags;             // bits 0-1 type (0 fluid, 1 solid, 2 velocity face, 3 open face), bits 2.. record+1
    // bit g set: the POP_GROUP cells of group g (cells 64 g .. 64 g + 63)
    // are plain fluid cells (flags 0) whose flags never change, so a wave of
    // them needs no flag load (null: not used)
    const uint32_t *plain;
    // 1: bulk blocks are taken in XCD order (the GPU deals workgroups to its
    // 8 XCDs in turn; the swizzle gives each XCD a contiguous run of tiles)
    uint32_t xcd_swizzle;
    // 1: rows start on a group boundary (nx a multiple of POP_GROUP) and the
    // 16-bit population buffer holds under 2^31 elements, so the bulk
    // kernel's parity-1 step addresses every neighbour slot as one of three
    // per-lane offsets plus a wave-uniform shift
    uint32_t rows_aligned;

    uint32_t nrec;
    uint32_t *rmask;             // [nrec]
    uint32_t *rcell;             // [nrec]
    uint32_t *rq;                // [9][nrec] f16 pairs (dir 2k+1 low, 2k+2 high)
    void *rfpost;                // [18][nrec] post-collision f~ of dirs 1..18 (real)
    void *rmac;                  // [4][nrec] rho, ux, uy, uz (real)
    void *rft;                   // [6][nrec] F, T (real)

    uint32_t nface;
    uint32_t *faces;


    double *hist;                // [KMAX][HIST_STRIDE]
    uint32_t nhw;                // halfway wall cells (no record)
    uint32_t *hw_cells;          // [nhw], sorted
    uint32_t hw_blocks;          // k_special blocks for them
    double *partials;            // [KMAX][wall_blocks + hw_blocks][8] per-block F, T of each step of a chunk
    // x-face sources that feed only their x-face cell leave their post-collision
    // state here ([side][nz][ny] entries of 12 floats, tagged with the
    // step's mark) for k_xfaces after the step
    float *xsrc;
    uint32_t xdefer;

    // physics (lattice units)
    double tau0, smag2, sponge_tau, sigma;
    double u_inf[3], u_init[3], grav[3], center[3];
    double sp_lo[3], sp_hi[3];
    double sp_ilo[3], sp_ihi[3];  // 1 / width (0 where no sponge)
    int32_t abs_u, total_p, wall_model, moving, forcing, sponge;

    // launch shape of the stream-collide kernel
    uint32_t bx;                 // threads along x per block (multiple of 64)
    uint32_t by;                 // x-rows per block
    uint32_t xblocks;            // ceil(nx / bx)
    uint32_t yblocks;            // ceil(ny / by)
    uint32_t wall_blocks;        // 256-thread blocks that process the records (>= 1)

    // moving geometry (swept box + body-frame occupancy)
    int32_t box_lo[3];
    uint32_t box_n[3];
    float occ_lo[3];
    float occ_invh;
    int32_t occ_n[3];
    const uint32_t *occ;
    uint32_t ntri;
    const float4 *tris;          // ntri x 3 (v0, e1, e2)
    // link hits of a spinning part's triangles, [2][nrec * 9][2]: per half
    // (StepArgs::hbuf, the step's parity), box cell and odd direction the
    // nearest and farthest hit's f32 bits (0x7f7fffff and 0 when none)
    uint32_t *qscr;
    // body-frame cells of one lattice spacing (from near_lo), two bit planes
    // of near_words words each (x fastest): plane 0 marks the cells within
    // NEAR_CELLS of the part's surface or inside it (a box cell outside them
    // has no link, hit or refill this step), plane 1 those within
    // NEAR_CELLS + 1 (box cells outside it are not within NEAR_CELLS at the
    // next step either, so their records need not be kept: they step as
    // plain fluid cells in the bulk kernel)
    float near_lo[3];
    int32_t near_n[3];
    uint32_t near_words;
    const uint32_t *near;

    // probes
    uint32_t nvert;
    const float4 *verts;         // nvert x 2 (position, normal)
    float vertex_offset;
    float *vacc;                 // nvert x 2
    uint32_t npix;
    const int4 *slicedef;
    float *sacc;                 // npix x 5
};

// Per-step uniform (lattice.rs:123-132): body rotation now / before (columns
// = body axes), angular velocity, stochastic-rounding seed, history slot.
struct StepArgs {
    float r[3][4];
    float q[3][4];
    float omega[4];
    uint32_t seed;
    uint32_t slot;
    // this step's tag, unique in the process (never 0): an x-face source's
    // deferred state carries it, so that only this step's states are taken
    uint32_t mark_lo, mark_hi;
    // a spinning part: the rotation the lattice's previous voxelization used
    // (vq_valid = 0 when there is none, or it is not known)
    float vq[3][4];
    uint32_t vq_valid;
    // a spinning part: the half of the hit scratch that holds this step's
    // hits (the step's parity), and whether k_voxelize clears the other
    // half's entries of the cells near the part at rotation vq (the hits of
    // the step before, voxelized earlier in the same call)
    uint32_t hbuf, hclear;
};

// A batch of lattices stepped together (device arrays prepared by the host):
// lats[nl] device views, boff_bulk[nl + 1] / boff_sp[nl + 1] first block of
// each lattice in the bulk / special launch, `threads` (SPECIAL_BLOCK) per
// bulk block (each lattice's bx * by divides it). fused: one kernel per step
// on s_special (the special blocks, then the bulk blocks at boff_bulk after
// them; boff_bulk must follow boff_sp in the same array).
struct BatchLaunch {
    const DevLattice *lats;
    const uint32_t *boff_bulk, *boff_sp;
    uint32_t nl, nbulk, nsp, threads;
    int32_t storage;
    bool force, sponge;
    bool fused;
    bool rows;   // every member has row-aligned addressing (DevLattice::rows_aligned)
};
// special kernel of the batch on s_special, bulk kernel on s_bulk (or the
// one-kernel step on s_special)
hipError_t launch_batch_step(const BatchLaunch &b, int parity, const StepArgs &a, hipStream_t s_special, hipStream_t s_bulk);

// The launchers of one population storage (kernels.hip compiled once per
// storage); the functions below dispatch on the lattice's storage.
struct StorageLaunchers {
    hipError_t (*init)(const DevLattice &, hipStream_t);
    hipError_t (*faces)(const DevLattice &, int, uint32_t, hipStream_t);
    hipError_t (*xfaces)(const DevLattice &, int, const StepArgs &, hipStream_t);
    hipError_t (*stream_collide)(const DevLattice &, int, const StepArgs &, hipStream_t, hipStream_t, hipEvent_t,
                                 hipEvent_t);
    hipError_t (*step_fused)(const DevLattice &, int, const StepArgs &, hipStream_t, hipEvent_t, hipEvent_t);
    hipError_t (*macros)(const DevLattice &, int, float4 *, hipStream_t);
    hipError_t (*mass_partial)(const DevLattice &, int, double *, uint32_t, hipStream_t);
    hipError_t (*samplers)(const DevLattice &, int, const StepArgs &, hipStream_t);
    hipError_t (*pbb_step)(const DevLattice &, int, const StepArgs &, const float *, hipStream_t);
    hipError_t (*adjoint_step)(const DevLattice &, const DevLattice &, int, const float *, double *, double, int,
                               hipStream_t);
    hipError_t (*batch_step)(const BatchLaunch &, int, const StepArgs &, hipStream_t, hipStream_t);
    hipError_t (*voxelize)(const DevLattice &, int, const StepArgs &, hipStream_t);
    hipError_t (*special)(const DevLattice &, int, const StepArgs &, hipStream_t);
    hipError_t (*bulk_z)(const DevLattice &, int, const StepArgs &, hipStream_t, uint32_t, uint32_t, uint32_t, uint32_t,
                         hipEvent_t, hipEvent_t);
};
extern const StorageLaunchers launchers_fp32, launchers_fp16, launchers_bf16, launchers_fp64;

hipError_t launch_init(const DevLattice &L, hipStream_t s);
// standalone face update before a step of this parity (seed of that step)
hipError_t launch_faces(const DevLattice &L, int parity, uint32_t seed, hipStream_t s);
// the x-face cells from the sources' states the step (of this parity and
// tag) left in xsrc
hipError_t launch_xfaces(const DevLattice &L, int parity, const StepArgs &a, hipStream_t s);
// k_special on s_special, k_bulk on s_bulk (independent within a step).
// With ev0/ev1 the bulk kernel's dispatch records its own start and end in
// them (hipExtLaunchKernel: no extra packets between the kernels)
hipError_t launch_stream_collide(const DevLattice &L, int parity, const StepArgs &a, hipStream_t s_special, hipStream_t s_bulk,
                                 hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// the same as one kernel on one stream (the tile's bx * by must divide
// SPECIAL_BLOCK); ev0/ev1 as above, for that kernel
hipError_t launch_step_fused(const DevLattice &L, int parity, const StepArgs &a, hipStream_t s, hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr);
// the special kernel alone, and the bulk kernel over nplanes z-planes:
// z = zbase + k, and z += zskip_len from zskip_at on (k = 0 .. nplanes-1)
hipError_t launch_special(const DevLattice &L, int parity, const StepArgs &a, hipStream_t s);
hipError_t launch_bulk_planes(const DevLattice &L, int parity, const StepArgs &a, hipStream_t s, uint32_t zbase,
                              uint32_t nplanes, uint32_t zskip_at, uint32_t zskip_len, hipEvent_t ev0 = nullptr,
                              hipEvent_t ev1 = nullptr);
hipError_t launch_macros(const DevLattice &L, int parity, float4 *out, hipStream_t s);
hipError_t launch_mass(const DevLattice &L, int parity, double *partials, uint32_t nblocks, double *out, hipStream_t s);
// re-voxelization of a spinning part for a step: triangle hits of the
// step's rotation into its half of the hit scratch; the box cells' records
// from them (and the clearing of the other half's entries the step before
// left, when a.hclear); the clearing of a step's own entries (the last step
// of a call, so that the scratch is clear between calls)
hipError_t launch_scatter(const DevLattice &L, const StepArgs &a, hipStream_t s);
hipError_t launch_voxelize(const DevLattice &L, int parity, const StepArgs &a, hipStream_t s);
hipError_t launch_qclear_near(const DevLattice &L, const StepArgs &a, hipStream_t s);
// resets all of the link scratch of a spinning part (setup)
hipError_t launch_qclear(const DevLattice &L, hipStream_t s);
hipError_t launch_samplers(const DevLattice &L, int parity, const StepArgs &a, hipStream_t s);
// force history of the first nsteps steps of a chunk from the partials
hipError_t launch_reduce_hist(const DevLattice &L, uint32_t nsteps, hipStream_t s);
// the same over an explicit partials array ([nsteps][nparts][8] -> hist)
hipError_t launch_reduce_partials(const double *partials, uint32_t nparts, uint32_t nsteps, double *hist, hipStream_t s);
// partial bounce-back: one forward step (FP32/FP64 storage), and the adjoint
// of one step (A.pops: adjoint populations in the storage's type; grad: d/d phi)
hipError_t launch_pbb_step(const DevLattice &L, int parity, const StepArgs &a, const float *phi, hipStream_t s);
hipError_t launch_adjoint_step(const DevLattice &L, const DevLattice &A, int parity, const float *phi, double *grad,
                               double wj, int axis, hipStream_t s);

}  // namespace simlab
