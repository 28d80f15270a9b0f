/*
 * simlab_cfd.h — C ABI of the MI355X (gfx950) D3Q19 wind-tunnel lattice and
 * the refined D3Q27 solver.
 *
 * Drop-in boundary for the uniform-grid lattice of SimLab's `crates/cfd`
 * (crates/cfd/src/lattice.rs:220-841) and its refined multi-level lattice
 * (crates/cfd/src/mlattice.rs:139-948). The Rust `Lattice` and
 * `MultiLattice` types forward each of their methods to one function here;
 * the Rust side of that binding is shown in INTEGRATION.md. Everything is in
 * lattice units (cell size 1, time step 1, reference density 1), as in the
 * reference.
 *
 * Conventions
 *  - Every function returns SIMLAB_OK (0) or an error code; the message of
 *    the last error on the calling thread is `simlab_cfd_last_error()`. The
 *    messages of setup errors are the reference's `bail!` strings.
 *  - No C++ exception crosses this boundary. Inputs are copied at `_new`;
 *    outputs go to caller-allocated arrays whose sizes `_info` reports.
 *  - One HIP stream per lattice handle. Distinct handles may be used from
 *    distinct threads; calls on one handle must be serialised by the caller
 *    (as `&mut self` does in Rust).
 */
#ifndef SIMLAB_CFD_H
#define SIMLAB_CFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIMLAB_CFD_ABI_VERSION 1

/* Status codes. */
enum {
    SIMLAB_OK = 0,
    SIMLAB_ERR_INVALID = 1,   /* bad setup / arguments (reference `bail!`)  */
    SIMLAB_ERR_GPU = 2,       /* HIP runtime failure, no device             */
    SIMLAB_ERR_OOM = 3,       /* "not enough GPU memory for this grid"      */
    SIMLAB_ERR_STATE = 4      /* call not valid in this state (e.g. macros
                                 without debug_dump)                         */
};

/* Population storage. lattice.rs:11-32 has Fp32 and Fp16 (scaled IEEE half,
 * stochastic rounding); Bf16 and Fp64 are additions of this library: Bf16
 * stores the same deviations from the weights as bfloat16 with stochastic
 * rounding, Fp64 is the double-precision oracle storage (FP64 arithmetic). */
typedef enum {
    SIMLAB_STORAGE_FP32 = 0,
    SIMLAB_STORAGE_FP16 = 1,
    SIMLAB_STORAGE_BF16 = 2,
    SIMLAB_STORAGE_FP64 = 3
} simlab_storage;

/* Face boundary condition, lattice.rs:34-43 (`FaceBc`). */
typedef enum {
    SIMLAB_FACE_VELOCITY = 0,
    SIMLAB_FACE_OPEN = 1,
    SIMLAB_FACE_PERIODIC = 2
} simlab_face_bc;

/* `LatticeSetup`, lattice.rs:45-76; field meaning and units as there. */
typedef struct simlab_lattice_setup {
    uint32_t dims[3];
    int32_t faces[6];            /* -x, +x, -y, +y, -z, +z (simlab_face_bc) */
    double nu;                   /* lattice kinematic viscosity             */
    double smagorinsky;          /* C_s (0 = no turbulence model)           */
    double u_inf[3];
    double u_init[3];
    double body_force[3];        /* uniform acceleration                    */
    uint32_t sponge[6];          /* viscosity sponge width per face, cells  */
    double sponge_tau;
    double sponge_sigma;
    int32_t absorb_velocity;
    int32_t total_pressure_inlet;
    int32_t storage;             /* simlab_storage                          */
    int32_t debug_dump;          /* keep the buffer `macros()` reads        */
    double center[3];            /* moment reference point / spin centre    */
    int32_t wall_model;          /* Spalding wall model in wall cells       */
    int32_t reserved0;
} simlab_lattice_setup;

/* `WallCell`, voxel.rs:328-335: a wall-adjacent fluid cell, the bit mask of
 * its cut links (bit i = direction i) and the Bouzidi q per direction. */
typedef struct simlab_wall_cell {
    uint32_t cell;
    uint32_t mask;
    float q[27];
} simlab_wall_cell;

enum { SIMLAB_GEOM_NONE = 0, SIMLAB_GEOM_STATIC = 1, SIMLAB_GEOM_SPIN = 2 };

/* `Geometry`, lattice.rs:87-92. Static: `StaticVoxels` (voxel.rs:337-344):
 * solid cell indices and wall cells (sorted by cell). Spin: `SpinVoxels`
 * (voxel.rs:492-505) plus the swept box in lattice cells. */
typedef struct simlab_geometry {
    int32_t kind;                /* SIMLAB_GEOM_*                           */
    int32_t reserved0;
    uint64_t nsolid;
    const uint32_t *solid;
    uint64_t nwall;
    const simlab_wall_cell *walls;
    /* spin */
    double occ_lo[3];
    double occ_h;
    uint32_t occ_dims[3];
    uint32_t reserved1;
    uint64_t nocc_words;
    const uint32_t *occ_bits;
    uint64_t ntri;
    const float *tris;           /* ntri x 12: v0.xyz,0, e1.xyz,0, e2.xyz,0 */
    int32_t box_lo[3];
    uint32_t box_n[3];
} simlab_geometry;

/* `SliceDef`, shader.rs:77-84. */
typedef struct simlab_slice_def {
    int32_t origin[3];
    int32_t u[3];
    int32_t v[3];
    uint32_t nu;
    uint32_t nv;
} simlab_slice_def;

/* `Probes`, lattice.rs:94-102. verts: nvert x 6 doubles (body-frame
 * position relative to the centre, then the unit normal). */
typedef struct simlab_probes {
    uint64_t nvert;
    const double *verts;
    uint64_t nslice;
    const simlab_slice_def *slices;
    double vertex_offset;
} simlab_probes;

/* `StepIn`, lattice.rs:104-121. rot: row-major 3x3 whose columns are the
 * body axes in lattice coordinates; omega in rad per step. */
typedef struct simlab_step_in {
    double rot[3][3];
    double rot_prev[3][3];
    double omega[3];
    int32_t sample;
    int32_t reserved0;
} simlab_step_in;

/* Public fields of `Lattice` (lattice.rs:148-170). */
typedef struct simlab_lattice_info {
    uint64_t steps_done;
    uint64_t cells;
    uint64_t memory_bytes;
    uint32_t dims[3];
    uint32_t nrec;
    uint32_t nface;
    uint32_t npix;
    uint32_t nvert;
    int32_t storage;             /* effective storage                       */
    int32_t moving;
    double tau0;
    uint64_t bulk_cells;         /* cells the bulk stream-collide kernel
                                    updates                                */
    uint64_t halfway_cells;      /* wall cells with only halfway (q = 1/2)
                                    links, bounced back without a record;
                                    nrec counts the other wall cells       */
} simlab_lattice_info;

typedef struct simlab_lattice simlab_lattice;            /* opaque */
typedef struct simlab_batch simlab_batch;                /* opaque */
typedef struct simlab_static_voxels simlab_static_voxels; /* opaque */

/* Message of the last failed call on this thread ("" if none). */
const char *simlab_cfd_last_error(void);

/* `gpu_info` (gpu.rs:56-63): "name · HIP · N.N GB max buffer" (the device
 * name, or AMD and its architecture where the driver gives none; the free
 * device memory, the largest buffer one allocation can have). */
int simlab_cfd_gpu_info(char *buf, size_t len);

/* `Lattice::max_cells` (lattice.rs:223-227): most cells a lattice with this
 * storage can hold by the buffer layout alone (memory budgets are the
 * caller's). */
uint64_t simlab_cfd_max_cells(int32_t storage);

/* `Lattice::new` (lattice.rs:229-673). probes may be NULL. */
int simlab_cfd_lattice_new(const simlab_lattice_setup *setup,
                           const simlab_geometry *geom,
                           const simlab_probes *probes,
                           simlab_lattice **out);

/* Drop a lattice and its device buffers (Rust `Drop`). NULL is a no-op. */
void simlab_cfd_lattice_free(simlab_lattice *lat);

/* Public fields of the lattice. */
int simlab_cfd_lattice_info(const simlab_lattice *lat, simlab_lattice_info *out);

/* `Lattice::run` (lattice.rs:699-753): runs nsteps steps and writes the
 * force and moment on the geometry after each step (nsteps x 6 doubles:
 * Fx Fy Fz Tx Ty Tz, lattice units, moment about `center`). The sum over
 * the wall records is accumulated in FP64 on the device. ft_out may be NULL. */
int simlab_cfd_lattice_run(simlab_lattice *lat, const simlab_step_in *steps,
                           size_t nsteps, double *ft_out);

/* `run` with nsteps copies of `StepIn::still(false)` (lattice.rs:118-120),
 * the form every test and benchmark of the reference uses. */
int simlab_cfd_lattice_run_still(simlab_lattice *lat, size_t nsteps,
                                 double *ft_out);

/* `run` that allocates nothing. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation). The same steps
 * and results as simlab_cfd_lattice_run, but the kernels' per-block force
 * partials and the per-chunk force history live in a 16 MiB device array of
 * the library (one per GPU, held until the process ends) instead of the
 * lattice's own buffers. A chunk then holds at most
 * 16 MiB / ((8 * wall blocks + 8) * 8 B) steps (wall blocks = the lattice's
 * wall-record cells / 256 plus its halfway wall cells / 256, each rounded
 * up); a lattice whose single step needs more than 16 MiB fails with
 * SIMLAB_ERR_INVALID. Every call on one GPU shares that array and leaves
 * the last call's contents in it: calls that overlap on one GPU (from
 * several threads) race. Timing is not recorded. */
int simlab_cfd_lattice_run_noalloc(simlab_lattice *lat, const simlab_step_in *steps,
                                   size_t nsteps, double *ft_out);

/* `Lattice::macros` (lattice.rs:818-835): (rho, ux, uy, uz) of every cell,
 * NaN where solid or on a face; needs debug_dump. out: cells x 4 floats. */
int simlab_cfd_lattice_macros(simlab_lattice *lat, float *out);

/* `Lattice::records` (lattice.rs:837-841): max(nrec,1) x 40 words. */
int simlab_cfd_lattice_records(simlab_lattice *lat, uint32_t *out);

/* `Lattice::reset_samples` (lattice.rs:785-791). */
int simlab_cfd_lattice_reset_samples(simlab_lattice *lat);

/* `Lattice::vertex_means` (lattice.rs:793-801): nvert floats. */
int simlab_cfd_lattice_vertex_means(simlab_lattice *lat, float *out);

/* `Lattice::slice_means` (lattice.rs:803-816): npix x 4 floats. */
int simlab_cfd_lattice_slice_means(simlab_lattice *lat, float *out);

/* Total fluid mass sum(rho) over the cells that collide (fluid, not face,
 * not solid), accumulated in FP64 on the device. New: the reference has no
 * mass check; it is the global conservation check of the FP16/BF16 storages. */
int simlab_cfd_lattice_mass(simlab_lattice *lat, double *mass);

/* Raw population slots in the reference's order (slot of direction i and
 * cell n at i * cells + n, whatever the device layout; 19 x cells elements
 * of the storage type: float, uint16 half/bf16 bits, or double), for tests.
 * `bytes` must equal 19 * cells * element size. */
int simlab_cfd_lattice_get_pops(simlab_lattice *lat, void *out, size_t bytes);
int simlab_cfd_lattice_set_pops(simlab_lattice *lat, const void *in, size_t bytes);

/* Checkpoint and restore (new: the reference has none). A checkpoint holds
 * the whole evolving state of a lattice: populations, flags, wall records,
 * sampler sums and steps_done; restoring it into a lattice built with the
 * same setup and geometry continues bit for bit as the original would.
 * checkpoint_size gives the bytes the caller provides. */
int simlab_cfd_lattice_checkpoint_size(const simlab_lattice *lat, size_t *bytes);
int simlab_cfd_lattice_checkpoint(simlab_lattice *lat, void *out, size_t bytes);
int simlab_cfd_lattice_restore(simlab_lattice *lat, const void *in, size_t bytes);

/* Device timing of the last `run` call: total milliseconds of its steps
 * (HIP events around each chunk of steps on the lattice's stream) and the
 * summed duration and count of the stream-collide launches (the one-kernel
 * step, or the bulk kernel of the two-stream step), from HIP events the
 * kernel's own dispatch records (hipExtLaunchKernel: no packets between the
 * kernels, so the figure is the kernel's own duration). Timing must have
 * been enabled. */
int simlab_cfd_lattice_set_timing(simlab_lattice *lat, int enable);
int simlab_cfd_lattice_timing(const simlab_lattice *lat, double *step_ms,
                              double *main_ms, uint64_t *main_launches);

/* Launch-shape autotuner (new; no counterpart in the reference: wgpu fixes
 * the workgroup size). Times `steps` steps (default 20) of the lattice for
 * each candidate tile of the stream-collide kernel (bx x-cells, a multiple
 * of 64, by rows, 64-1024 threads) and each launch (one kernel per step,
 * where the tile divides 256 threads, or two kernels on two streams), keeps
 * the fastest, and restores the lattice's state (populations, records,
 * steps_done) after every trial. With use_cache the choice made earlier in
 * this process for a lattice of the same size, storage and wall presence is
 * taken without timing. tile_out (may be NULL) gets {bx, by}; step_us_out
 * the best step time (0 if cached). */
int simlab_cfd_lattice_autotune(simlab_lattice *lat, uint32_t steps, int use_cache, uint32_t *tile_out,
                                double *step_us_out);
/* Sets the tile of the stream-collide kernel explicitly. */
int simlab_cfd_lattice_set_tile(simlab_lattice *lat, uint32_t bx, uint32_t by);
/* The tile in use: tile_out = {bx, by}. */
int simlab_cfd_lattice_get_tile(const simlab_lattice *lat, uint32_t *tile_out);

/* Batches (new): several independent lattices (one storage, the same body
 * force and sponge settings, static geometry, the same steps_done) stepped
 * together, one launch per kernel and step for the whole batch (256-thread
 * blocks; a lattice's tile is kept if it divides 256). Lattices of more than
 * 4 Mi cells step with their own launches inside the call (their launch
 * overhead is small beside their step). The lattices stay owned by the
 * caller and must outlive the batch. run_still: nsteps still steps of every
 * lattice; ft_out (may be NULL) is lattices x nsteps x 6 doubles, as `run`
 * gives per lattice. */
int simlab_cfd_batch_new(simlab_lattice *const *lats, size_t n, simlab_batch **out);
int simlab_cfd_batch_run_still(simlab_batch *b, size_t nsteps, double *ft_out);
void simlab_cfd_batch_free(simlab_batch *b);

/* Partial bounce-back (new; no counterpart in the reference): solid fraction
 * phi in [0, 1] per cell (cells x floats, x fastest; NULL sets 0). A cell
 * blends its post-collision populations with its bounced inputs,
 * f_i <- (1 - phi) f_i + phi f_in,opp(i), and the porous solid takes the
 * force phi * sum_i c_i (f_i - f_in,opp(i)). Needs FP32 or FP64 storage and
 * a lattice without wall geometry, body force, sponge or wall model. */
int simlab_cfd_lattice_set_solid_fraction(simlab_lattice *lat, const float *phi);

/* Drag of the porous solid and its discrete adjoint (new).
 * adjoint_forward: restarts from the initial state, runs nsteps steps with
 * partial bounce-back, keeps the populations every ckpt_every steps, and
 * returns J = mean over steps [avg_from, nsteps) of the force component
 * `axis` (0 x, 1 y, 2 z), summed in FP64; ft_out (nsteps x 6, may be NULL)
 * gets the per-step force (moment entries 0).
 * adjoint_backward: dJ/dphi for every cell (cells x doubles) by the discrete
 * adjoint of those steps (the same arithmetic type as the storage),
 * recomputing each ckpt_every-step segment from its checkpoint. Needs
 * smagorinsky = 0. */
int simlab_cfd_adjoint_forward(simlab_lattice *lat, size_t nsteps, size_t ckpt_every, size_t avg_from, int axis,
                               double *J, double *ft_out);
int simlab_cfd_adjoint_backward(simlab_lattice *lat, double *dJ_dphi);

/* `voxelize_static` (voxel.rs:348-352, 363-490) on the host CPU: solid cells
 * and wall links of a triangle surface given in lattice coordinates (cell
 * centres at integer points). verts: nverts x 3 doubles; tris: ntris x 3. */
int simlab_cfd_voxelize_static(const double *verts, uint64_t nverts,
                               const uint32_t *tris, uint64_t ntris,
                               const uint64_t dims[3],
                               simlab_static_voxels **out);
/* `voxelize_static` with its two expensive parts on the GPU (the current
 * device): the ray-parity inside test (one thread per scan line) and the
 * wall-link search (one thread per examined point). The same double
 * precision operations in the same order as the CPU version, no fused
 * multiply-adds: the output is identical, solid cells, links and q bits.
 * New: the reference voxelizes on the CPU only. A scan line crossing more
 * than 64 surfaces makes the inside test fall back to the CPU for that
 * call. */
int simlab_cfd_voxelize_static_gpu(const double *verts, uint64_t nverts,
                                   const uint32_t *tris, uint64_t ntris,
                                   const uint64_t dims[3],
                                   simlab_static_voxels **out);
/* Views into a voxelization (valid until it is freed); lo/hi: examined box. */
int simlab_cfd_voxels_view(const simlab_static_voxels *v, uint64_t *nsolid,
                           const uint32_t **solid, uint64_t *nwall,
                           const simlab_wall_cell **walls, uint64_t lo[3],
                           uint64_t hi[3]);
void simlab_cfd_voxels_free(simlab_static_voxels *v);

/* `voxelize_spin` (voxel.rs:509-553) on the host CPU: a spinning part given
 * in body coordinates (lattice units, origin at the spin centre): occupancy
 * bits at spacing h (x fastest) and triangles split to edges <= 1.5 cells,
 * stored as (v0, e1, e2) float x 12; `axis` is the spin axis for the radius
 * and axial extent. The views stay valid until the object is freed. */
typedef struct simlab_spin_voxels simlab_spin_voxels;   /* opaque */
int simlab_cfd_voxelize_spin(const double *verts, uint64_t nverts, const uint32_t *tris, uint64_t ntris,
                             const double axis[3], double h, simlab_spin_voxels **out);
int simlab_cfd_spin_voxels_view(const simlab_spin_voxels *v, double occ_lo[3], double *occ_h, uint64_t occ_dims[3],
                                uint64_t *nocc_words, const uint32_t **occ_bits, uint64_t *ntri, const float **tris,
                                double *radius, double axial[2]);
void simlab_cfd_spin_voxels_free(simlab_spin_voxels *v);

/* `voxelize_static27` (voxel.rs:354-359, 363-490): the same voxelization
 * with the D3Q27 links (direction i = (i % 3 - 1, (i / 3) % 3 - 1,
 * i / 9 - 1), d3q27.rs:1-30), for the refined solver. Host CPU. */
int simlab_cfd_voxelize_static27(const double *verts, uint64_t nverts,
                                 const uint32_t *tris, uint64_t ntris,
                                 const uint64_t dims[3],
                                 simlab_static_voxels **out);

/* ------------------------------------------------------------------------
 * The refined D3Q27 cumulant solver (crates/cfd/src/mlattice.rs,
 * mshader.rs, d3q27.rs): nested node-aligned boxes, each level half the
 * cell size of its parent and stepped twice per parent step (the same
 * lattice velocity on every level). The still part lives on the finest
 * level, the domain faces on the coarsest. Lattice units of the finest
 * level unless stated. Storage FP32, FP16 (deviations from the weights,
 * scaled by 2^15, the reference's per-store stochastic rounding) or FP64
 * (the oracle's arithmetic; new). */
#define SIMLAB_MULTI_MAX_LEVELS 8

/* `LevelBox`, mlattice.rs:15-21: lo is the parent node of this level's
 * node 0, n the parent nodes it spans per axis (dims = 2 (n - 1) + 1). */
typedef struct simlab_level_box {
    int32_t lo[3];
    uint32_t n[3];
} simlab_level_box;

/* `MultiSetup`, mlattice.rs:23-48; field meaning and units as there. */
typedef struct simlab_multi_setup {
    uint32_t dims0[3];           /* coarsest level size in its cells        */
    uint32_t nboxes;             /* levels - 1                              */
    simlab_level_box boxes[SIMLAB_MULTI_MAX_LEVELS - 1];
    int32_t faces[6];            /* simlab_face_bc, -x +x -y +y -z +z       */
    double nu0;                  /* viscosity on the coarsest level         */
    double smagorinsky;
    double u_inf[3];
    double u_init[3];
    double body_force0[3];       /* coarsest level; halves per level        */
    uint32_t sponge[6];          /* coarsest level cells                    */
    double sponge_tau;
    double sponge_sigma;
    int32_t absorb_velocity;
    int32_t total_pressure_inlet;
    int32_t storage;             /* FP32, FP16 or FP64                      */
    int32_t debug_dump;
    double center[3];            /* moment reference, finest coordinates    */
} simlab_multi_setup;

/* Public fields of `MultiLattice` (mlattice.rs:139-160) and its levels. */
typedef struct simlab_multi_info {
    uint32_t levels;
    uint32_t sub_steps;          /* finest steps per coarsest step          */
    uint64_t fine_steps;
    uint64_t memory_bytes;
    uint32_t nrec;
    uint32_t nvert;
    uint32_t npix;
    int32_t storage;
    uint32_t dims[SIMLAB_MULTI_MAX_LEVELS][3];
    uint64_t cells[SIMLAB_MULTI_MAX_LEVELS];
    double tau[SIMLAB_MULTI_MAX_LEVELS];   /* relaxation time per level    */
} simlab_multi_info;

typedef struct simlab_multi simlab_multi;   /* opaque */

/* `MultiLattice::new` (mlattice.rs:185-690). geom: SIMLAB_GEOM_NONE or
 * SIMLAB_GEOM_STATIC with D3Q27 wall cells in finest-level cells (from
 * simlab_cfd_voxelize_static27); probes may be NULL (slice origins and
 * steps in finest-level node coordinates). */
int simlab_cfd_multi_new(const simlab_multi_setup *setup, const simlab_geometry *geom,
                         const simlab_probes *probes, simlab_multi **out);
void simlab_cfd_multi_free(simlab_multi *m);
int simlab_cfd_multi_info(const simlab_multi *m, simlab_multi_info *out);
/* `MultiLattice::run` (mlattice.rs:725-784): nsteps finest steps (a multiple
 * of sub_steps); force and moment on the part after each finest step
 * (nsteps x 6 doubles, finest lattice units), summed in FP64. */
int simlab_cfd_multi_run(simlab_multi *m, const simlab_step_in *steps, size_t nsteps, double *ft_out);
int simlab_cfd_multi_run_still(simlab_multi *m, size_t nsteps, double *ft_out);
/* `macros(l)` (mlattice.rs:913-930): (rho, ux, uy, uz) of every cell of
 * level l in that level's units, NaN where not a stepped fluid cell; needs
 * debug_dump. out: cells(l) x 4 floats. */
int simlab_cfd_multi_macros(simlab_multi *m, uint32_t level, float *out);
/* `records` (mlattice.rs:932-934): max(nrec, 1) x 24 words (mask, cell,
 * 13 words of f16 q pairs by moving-direction slot, F, T as f32 bits). */
int simlab_cfd_multi_records(simlab_multi *m, uint32_t *out);
/* `reset_samples`, `vertex_means`, `slice_means` (mlattice.rs:868-911). */
int simlab_cfd_multi_reset_samples(simlab_multi *m);
int simlab_cfd_multi_vertex_means(simlab_multi *m, float *out);
int simlab_cfd_multi_slice_means(simlab_multi *m, float *out);
/* The populations a level's next step reads (the source set), decoded to
 * deviations f - w as doubles, direction i of cell n at i * cells + n. For
 * tests. out: 27 x cells(l) doubles. */
int simlab_cfd_multi_get_pops(simlab_multi *m, uint32_t level, double *out);

/* Measurement only (no counterpart in the reference): the HBM bandwidth the
 * streaming access patterns of one D3Q19 step move with no arithmetic, on
 * the grouped 16-bit layout of a periodic nx x ny x nz grid (cells a multiple
 * of 64): gbs[0] the AA even step in place, gbs[1] the AA odd step in place,
 * gbs[2] a two-lattice pull step, gbs[3] an aligned out-of-place step; GB/s
 * of the 76 B per cell each moves, averaged over reps passes. Allocates two
 * population buffers for the call. */
int simlab_cfd_probe_stream_patterns(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t reps, double gbs[4]);

/* `inside_grid` (voxel.rs:73-95): inside flags (0/1) for the grid points
 * origin + h * (i, j, k), i fastest. out: dims[0]*dims[1]*dims[2] bytes. */
int simlab_cfd_inside_grid(const double *verts, uint64_t nverts,
                           const uint32_t *tris, uint64_t ntris,
                           const double origin[3], double h,
                           const uint64_t dims[3], uint8_t *out);

#ifdef __cplusplus
}
#endif

#endif /* SIMLAB_CFD_H */
