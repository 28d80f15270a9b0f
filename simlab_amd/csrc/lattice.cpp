// Host side of the MI355X lattice: setup validation, face list, flags and
// wall records (crates/cfd/src/lattice.rs:229-526), device buffers, the
// per-step launch sequence (lattice.rs:699-753) and the read-backs
// (lattice.rs:755-841), behind the C ABI of include/simlab_cfd.h.

#include "hip_util.h"
#include "kernels.h"
#include "simlab_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

namespace simlab {

namespace {
thread_local std::string g_err;

size_t elem_bytes(int storage) {
    switch (storage) {
        case SIMLAB_STORAGE_FP32: return 4;
        case SIMLAB_STORAGE_FP64: return 8;
        default: return 2;
    }
}
size_t real_bytes(int storage) { return storage == SIMLAB_STORAGE_FP64 ? 8 : 4; }

float half_to_float(uint16_t h) {
    uint32_t s = (uint32_t)(h & 0x8000) << 16, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    float v;
    if (e == 0) {
        v = std::ldexp((float)m, -24);
        return s ? -v : v;
    }
    uint32_t bits = e == 31 ? (s | 0x7f800000u | (m << 13)) : (s | ((e + 112) << 23) | (m << 13));
    std::memcpy(&v, &bits, 4);
    return v;
}

}  // namespace

void set_error(const std::string &msg) { g_err = msg; }

}  // namespace simlab

using namespace simlab;

struct simlab_lattice {
    int device = 0;
    simlab_lattice_setup setup{};
    int storage = 0;
    uint64_t ncell = 0;
    uint64_t steps_done = 0;
    uint32_t nrec = 0, nface = 0, npix = 0, nvert = 0;
    uint64_t nhalfway = 0;           // wall cells handled without a record
    uint64_t memory_bytes = 0;
    bool moving = false;
    hipStream_t stream = nullptr;    // faces, special kernel, copies
    hipStream_t stream2 = nullptr;   // bulk kernel, concurrent with the special kernel
    hipEvent_t fork = nullptr, join = nullptr;
    // spinning part: the next step's triangle hits are found on stream3
    // while this step's stream-collide kernels run
    hipStream_t stream3 = nullptr;
    hipEvent_t ev_vox = nullptr, ev_scat = nullptr;
    bool scat_ready = false;         // the next step's hits are on stream3 behind ev_scat
    bool spin_overlap = true;
    // the rotation of the last voxelization (columns as in StepArgs), valid
    // while the records it wrote are the lattice's state
    float vox_prev[3][4] = {};
    bool vox_prev_valid = false;
    uint64_t nbulk = 0;              // cells the bulk kernel updates
    DevLattice L{};
    DevBuf xsrc;                     // deferred x-face sources (L.xsrc)
    DevBuf plain;                    // plain-fluid groups (L.plain)
    DevBuf pops, flags, rmask, rcell, rq, rfpost, rmac, rft, faces, hist, hw_cells, partials,
        counter, dump, mass_part, mass_out, occ, tris, qscr, near, verts, vacc, slicedef, sacc;
    bool fused_faces = true;   // face update fused into the previous step
    // partial bounce-back and the drag adjoint
    DevBuf phi, grad, adj, ckpt, seg, pbb_part;
    bool pbb = false;               // solid fraction set
    bool has_solid = false;         // the geometry marks solid cells
    uint64_t adj_steps = 0, adj_every = 0, adj_from = 0;
    int adj_axis = 0;
    bool adj_ready = false;         // checkpoints of the last adjoint_forward are valid
    bool concurrent = true;    // k_special on a second stream beside k_bulk
    int side_special = -1;     // k_special on the side stream: 1 always, 0 never, -1 by its size
    bool fused_launch = true;  // one kernel per step when the tile allows
    double *hist_host = nullptr;   // pinned, KMAX x HIST_STRIDE
    bool timing = false;
    std::vector<hipEvent_t> ev;    // 2 per step of a chunk + 2 around the chunk
    double last_step_ms = 0.0, last_main_ms = 0.0;
    uint64_t last_main_launches = 0;

    ~simlab_lattice() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (stream2) (void)hipStreamSynchronize(stream2);
        if (stream3) (void)hipStreamSynchronize(stream3);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (ev_vox) (void)hipEventDestroy(ev_vox);
        if (ev_scat) (void)hipEventDestroy(ev_scat);
        if (hist_host) (void)hipHostFree(hist_host);
        if (stream3) (void)hipStreamDestroy(stream3);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

const char *validate(const simlab_lattice_setup *s, std::string &msg) {
    // lattice.rs:233-244
    uint64_t n = (uint64_t)s->dims[0] * s->dims[1] * s->dims[2];
    if (n == 0 || n >= 0xffffffffull) {
        msg = "grid of " + std::to_string(n) + " cells is out of range";
        return msg.c_str();
    }
    for (int k = 0; k < 3; ++k)
        if (s->dims[k] < 3) return "grid must be at least 3 cells in every direction";
    for (int a = 0; a < 3; ++a)
        if ((s->faces[2 * a] == SIMLAB_FACE_PERIODIC) != (s->faces[2 * a + 1] == SIMLAB_FACE_PERIODIC))
            return "periodic faces must come in pairs";
    for (int f = 0; f < 6; ++f)
        if (s->faces[f] < 0 || s->faces[f] > 2) return "unknown face boundary condition";
    if (s->storage < 0 || s->storage > 3) return "unknown population storage";
    return nullptr;
}

// Launch shape: whole x-rows per block. bx = nx rounded up to a wave (at most
// 1024 lanes), by rows so a block has about `target` threads.
// Default launch: lattices of up to FUSED_MAX_CELLS cells step with one
// kernel per step (special and bulk blocks together: no cross-stream wait,
// which costs about 15 us a step), larger ones with the two kernels on two
// streams (each at its own register count, so the bulk kernel keeps its
// occupancy). Tiles: two rows of bx cells in the fused kernel, bx = 128, or
// 64 where that leaves fewer idle lanes in a row (FP32 on 256x128x128:
// 128 x 2 is 2 % faster than whole 256-cell rows); in the
// split launch 16-bit storages run best with 128 x 1
// (two waves per block, more blocks per CU), 32/64-bit ones with whole
// rows. Measured on 256x128x128 and 384^3; `simlab_cfd_lattice_autotune`
// searches tiles and both launches per lattice.
constexpr uint64_t FUSED_MAX_CELLS = 16u << 20;
// two-stream step: a special kernel of up to this many blocks runs on the
// side stream (all of them fit on the GPU beside the bulk kernel's)
constexpr uint32_t SIDE_SPECIAL_BLOCKS = 512;

void choose_shape(DevLattice &L, uint32_t target_threads, uint32_t by_override, bool fused) {
    // rows wider than 1024 cells split into equal pieces (a multiple of 64)
    const uint32_t pieces = (L.nx + 1023) / 1024;
    uint32_t bx = (((L.nx + pieces - 1) / pieces + 63) / 64) * 64;
    if (bx > 1024) bx = 1024;
    const bool half = L.storage == SIMLAB_STORAGE_FP16 || L.storage == SIMLAB_STORAGE_BF16;
    // 128 cells wide, or 64 where that leaves fewer idle lanes in a row
    // (a 192-cell row: 64 x 2 tiles step FP16 22 % and FP32 9 % faster than
    // 128 x 2, whose second tile idles half its lanes)
    const uint32_t idle128 = (L.nx + 127) / 128 * 128 - L.nx, idle64 = (L.nx + 63) / 64 * 64 - L.nx;
    const uint32_t narrow = idle64 < idle128 ? 64u : 128u;
    if (!std::getenv("SIMLAB_CFD_BLOCK")) {
        if (fused) {
            bx = std::min<uint32_t>(bx, narrow);   // 64 or 128: divides SPECIAL_BLOCK
            target_threads = 2 * bx;               // two rows per tile
        } else if (half && bx > 128) {
            bx = narrow;
            target_threads = 128;
        }
    }
    uint32_t by = by_override ? by_override : std::max<uint32_t>(1, target_threads / bx);
    by = std::max<uint32_t>(1, std::min<uint32_t>(by, std::min<uint32_t>(1024 / bx, L.ny)));
    L.bx = bx;
    L.by = by;
    L.xblocks = (L.nx + bx - 1) / bx;
    L.yblocks = (L.ny + by - 1) / by;
    L.wall_blocks = std::max<uint32_t>(1, (L.nrec + SPECIAL_BLOCK - 1) / SPECIAL_BLOCK);
}

// between the device's grouped population order and the reference's
// direction-major order (slot i * ncell + n): to_ref copies dev -> ref,
// otherwise ref -> dev
void reorder_pops(uint64_t ncell, size_t eb, char *dev, char *ref, bool to_ref) {
    for (uint32_t i = 0; i < 19; ++i)
        for (uint64_t n = 0; n < ncell; ++n) {
            char *d = dev + pop_index(i, (uint32_t)n) * eb;
            char *r = ref + ((size_t)i * ncell + n) * eb;
            if (to_ref) std::memcpy(r, d, eb);
            else std::memcpy(d, r, eb);
        }
}

// Body-frame grid of unit cells from lo: a cell is marked when it holds a
// solid occupancy sample or meets a triangle's bounding box, then the marks
// are dilated by NEAR_CELLS cells (Chebyshev; first bit plane) and by one
// more (second plane). A point in an unmarked cell of the first plane is
// farther than NEAR_CELLS - 1 from every surface and solid sample.
std::vector<uint32_t> near_grid(const simlab_geometry &g, float lo[3], int32_t n[3]) {
    double flo[3];
    for (int k = 0; k < 3; ++k) {
        flo[k] = g.occ_lo[k] - NEAR_CELLS - 2.0;
        n[k] = (int32_t)std::ceil(g.occ_dims[k] * g.occ_h) + 2 * NEAR_CELLS + 5;
        lo[k] = (float)flo[k];
    }
    const size_t nx = n[0], ny = n[1], nz = n[2];
    std::vector<uint8_t> m(nx * ny * nz, 0);
    auto cell = [&](double v, int k) {
        return std::min<int64_t>(std::max<int64_t>((int64_t)std::floor(v - flo[k]), 0), n[k] - 1);
    };
    const uint64_t ox = g.occ_dims[0], oy = g.occ_dims[1];
    for (uint64_t w = 0; w < g.nocc_words; ++w) {
        uint32_t bits = g.occ_bits[w];
        while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            const uint64_t l = w * 32 + b;
            const uint64_t i = l % ox, j = (l / ox) % oy, kk = l / (ox * oy);
            const double p[3] = {g.occ_lo[0] + i * g.occ_h, g.occ_lo[1] + j * g.occ_h, g.occ_lo[2] + kk * g.occ_h};
            m[((size_t)cell(p[2], 2) * ny + cell(p[1], 1)) * nx + cell(p[0], 0)] = 1;
        }
    }
    for (uint64_t t = 0; t < g.ntri; ++t) {
        const float *r = g.tris + 12 * t;
        int64_t a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            const double v0 = r[k], v1 = r[k] + r[4 + k], v2 = r[k] + r[8 + k];
            a[k] = cell(std::min(std::min(v0, v1), v2), k);
            b[k] = cell(std::max(std::max(v0, v1), v2), k);
        }
        for (int64_t z = a[2]; z <= b[2]; ++z)
            for (int64_t y = a[1]; y <= b[1]; ++y)
                for (int64_t x = a[0]; x <= b[0]; ++x) m[((size_t)z * ny + y) * nx + x] = 1;
    }
    // separable max filter of radius r along x, y, z
    std::vector<uint8_t> t2(m.size());
    const size_t stride[3] = {1, nx, nx * ny};
    auto dilate = [&](int r) {
        for (int ax = 0; ax < 3; ++ax) {
            const int64_t len = n[ax];
            for (size_t c = 0; c < m.size(); ++c) {
                const int64_t pos = (int64_t)((c / stride[ax]) % (size_t)len);
                uint8_t v = 0;
                for (int64_t d = -r; d <= r && !v; ++d) {
                    const int64_t q = pos + d;
                    if (q >= 0 && q < len) v = m[c + (size_t)(q - pos) * stride[ax]];
                }
                t2[c] = v;
            }
            m.swap(t2);
        }
    };
    const size_t words = (m.size() + 31) / 32;
    std::vector<uint32_t> bits(2 * words, 0);
    dilate(NEAR_CELLS);
    for (size_t c = 0; c < m.size(); ++c)
        if (m[c]) bits[c >> 5] |= 1u << (c & 31);
    dilate(1);
    for (size_t c = 0; c < m.size(); ++c)
        if (m[c]) bits[words + (c >> 5)] |= 1u << (c & 31);
    return bits;
}

// an explicit tile of bx x-cells (a multiple of 64) by `by` rows
void set_shape(DevLattice &L, uint32_t bx, uint32_t by) {
    L.bx = bx;
    L.by = by;
    L.xblocks = (L.nx + bx - 1) / bx;
    L.yblocks = (L.ny + by - 1) / by;
}

uint32_t default_block_threads() {
    const char *e = std::getenv("SIMLAB_CFD_BLOCK");
    if (e) {
        long v = std::strtol(e, nullptr, 10);
        if (v >= 64 && v <= 1024) return (uint32_t)v;
    }
    return 256;
}

// step tags, unique in the process (the first is 1; x-face source states
// start zeroed, so no stale state ever carries a tag a later step gets)
std::atomic<uint64_t> g_step_mark{0};

void pack_step(const simlab_step_in &s, uint64_t step, uint32_t slot, StepArgs &a) {
    // lattice.rs:675-686: columns of the rotations as f32; seed per step
    for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < 3; ++r) {
            a.r[k][r] = (float)s.rot[r][k];
            a.q[k][r] = (float)s.rot_prev[r][k];
        }
        a.r[k][3] = 0.0f;
        a.q[k][3] = 0.0f;
        a.omega[k] = (float)s.omega[k];
    }
    a.omega[3] = 0.0f;
    a.seed = (uint32_t)step * 2654435761u;
    a.slot = slot;
    const uint64_t mark = ++g_step_mark;
    a.mark_lo = (uint32_t)mark;
    a.mark_hi = (uint32_t)(mark >> 32);
    std::memset(a.vq, 0, sizeof a.vq);
    a.vq_valid = 0;
    a.hbuf = (uint32_t)(step & 1u);
    a.hclear = 0;
}

int new_lattice(const simlab_lattice_setup *su, const simlab_geometry *g, const simlab_probes *pr, simlab_lattice **out) {
    if (!su || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
    *out = nullptr;
    std::string msg;
    if (const char *m = validate(su, msg)) return fail(SIMLAB_ERR_INVALID, m);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(SIMLAB_ERR_GPU, "no usable GPU: no HIP device found");
    }
    auto lat = std::make_unique<simlab_lattice>();
    hip_check(hipGetDevice(&lat->device), "hipGetDevice");
    lat->setup = *su;
    lat->storage = su->storage;
    const uint32_t nx = su->dims[0], ny = su->dims[1], nz = su->dims[2];
    const uint64_t n = (uint64_t)nx * ny * nz;
    lat->ncell = n;
    const int geom_kind = g ? g->kind : SIMLAB_GEOM_NONE;
    if (geom_kind == SIMLAB_GEOM_SPIN) {
        for (int k = 0; k < 3; ++k)
            if (g->box_lo[k] < 1 || (int64_t)g->box_lo[k] + g->box_n[k] > (int64_t)su->dims[k] - 1)
                return fail(SIMLAB_ERR_INVALID, "the spinning part's swept box leaves the domain");
        if (!(g->occ_h > 0.0) || !g->occ_bits || g->nocc_words * 32 < (uint64_t)g->occ_dims[0] * g->occ_dims[1] * g->occ_dims[2] ||
            (g->ntri && !g->tris))
            return fail(SIMLAB_ERR_INVALID, "spinning geometry: occupancy bits or triangles missing");
    }
    if (geom_kind == SIMLAB_GEOM_STATIC) {
        for (uint64_t r = 0; r < g->nwall; ++r)
            if (g->walls[r].cell >= n) return fail(SIMLAB_ERR_INVALID, "wall cell outside the grid");
        for (uint64_t r = 0; r < g->nsolid; ++r)
            if (g->solid[r] >= n) return fail(SIMLAB_ERR_INVALID, "solid cell outside the grid");
    }

    // face cells of non-periodic axes (lattice.rs:283-324)
    auto face_type = [&](uint32_t x, uint32_t y, uint32_t z) -> int {
        const bool on[6] = {x == 0, x == nx - 1, y == 0, y == ny - 1, z == 0, z == nz - 1};
        int ty = -1;
        for (int f = 0; f < 6; ++f)
            if (on[f] && su->faces[f] != SIMLAB_FACE_PERIODIC) {
                const int t = su->faces[f] == SIMLAB_FACE_VELOCITY ? 2 : 3;
                ty = (ty == 2) ? 2 : t;
            }
        return ty;
    };
    std::vector<uint32_t> faces;
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y) {
            const bool edge = y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
            if (edge) {
                for (uint32_t x = 0; x < nx; ++x)
                    if (face_type(x, y, z) >= 0) faces.push_back((z * ny + y) * nx + x);
            } else {
                for (uint32_t x : {0u, nx - 1})
                    if (face_type(x, y, z) >= 0) faces.push_back((z * ny + y) * nx + x);
            }
        }
    lat->nface = (uint32_t)faces.size();

    // flags: face type, then solid (lattice.rs:489-526)
    std::vector<uint32_t> h_flags(n, 0);
    for (uint32_t f : faces) {
        uint32_t z = f / (nx * ny), r = f - z * nx * ny, y = r / nx, x = r - y * nx;
        h_flags[f] = (uint32_t)face_type(x, y, z);
    }
    if (geom_kind == SIMLAB_GEOM_STATIC)
        for (uint64_t k = 0; k < g->nsolid; ++k)
            if (h_flags[g->solid[k]] == 0) h_flags[g->solid[k]] = 1;
    lat->has_solid = geom_kind == SIMLAB_GEOM_STATIC && g->nsolid > 0;

    // Wall cells whose cut links all sit halfway (q = 1/2) and lead into
    // solid cells need no record: the bulk kernel bounces
    // them back from the slot the cell's own last output went to (static
    // geometry, no wall model, no sponge). Every other wall cell keeps a
    // record (record_words, lattice.rs:202-218), SoA.
    bool hw_ok = geom_kind == SIMLAB_GEOM_STATIC && !su->wall_model;
    for (int k = 0; k < 6; ++k)
        if (su->sponge[k]) hw_ok = false;
    if (const char *e = std::getenv("SIMLAB_CFD_HALFWAY")) hw_ok = hw_ok && e[0] != '0';
    std::vector<uint32_t> keep;
    uint64_t nhalfway = 0;
    if (geom_kind == SIMLAB_GEOM_STATIC) {
        for (uint64_t r = 0; r < g->nwall; ++r) {
            const simlab_wall_cell &w = g->walls[r];
            const uint32_t c = w.cell;
            bool hw = hw_ok && (h_flags[c] & 3u) == 0;
            if (hw) {
                const uint32_t cz = c / (nx * ny), rr = c - cz * nx * ny, cy = rr / nx, cx = rr - cy * nx;
                for (int i = 1; i < 19 && hw; ++i) {
                    if (!(w.mask & (1u << i))) continue;
                    if (f16_bits_host(w.q[i]) != 0x3800) hw = false;   // q != 1/2
                    const uint32_t tx = (uint32_t)(((int)cx + CV[i][0] + (int)nx) % (int)nx);
                    const uint32_t ty = (uint32_t)(((int)cy + CV[i][1] + (int)ny) % (int)ny);
                    const uint32_t tz = (uint32_t)(((int)cz + CV[i][2] + (int)nz) % (int)nz);
                    // the slot read back belongs to the cell across the link:
                    // it must be solid (a fluid cell across a thin part writes
                    // that slot in the same step; a face cell's is rewritten)
                    if (h_flags[((uint64_t)tz * ny + ty) * nx + tx] != 1u) hw = false;
                }
            }
            if (hw) {
                h_flags[c] = FLAG_HALFWAY | ((w.mask & 0x7FFFEu) << 1);
                ++nhalfway;
            } else {
                keep.push_back((uint32_t)r);
            }
        }
    }
    lat->nhalfway = nhalfway;
    // a spinning part keeps one record per cell of its swept box (x fastest;
    // lattice.rs:465-478), re-voxelized every step
    const bool spin = geom_kind == SIMLAB_GEOM_SPIN;
    const uint32_t nrec = spin ? (uint32_t)((uint64_t)g->box_n[0] * g->box_n[1] * g->box_n[2]) : (uint32_t)keep.size();
    lat->nrec = nrec;
    lat->moving = spin;
    const size_t rb = real_bytes(su->storage);
    std::vector<uint32_t> h_mask(nrec), h_cell(nrec), h_q((size_t)9 * nrec);
    std::vector<unsigned char> h_fpost((size_t)18 * nrec * rb), h_mac((size_t)4 * nrec * rb), h_ft((size_t)6 * nrec * rb, 0);
    {
        const double *ui = su->u_init;
        const double usq = 1.5 * (ui[0] * ui[0] + ui[1] * ui[1] + ui[2] * ui[2]);
        double fe[19];   // feq_shifted (lattice.rs:191-200); f32 words except in FP64 storage
        for (int i = 0; i < 19; ++i) {
            const double cu = CV[i][0] * ui[0] + CV[i][1] * ui[1] + CV[i][2] * ui[2];
            fe[i] = weight(i) * (3.0 * cu + 4.5 * cu * cu - usq);
        }
        auto put = [&](std::vector<unsigned char> &v, size_t idx, double x) {
            if (rb == 8) {
                std::memcpy(&v[idx * 8], &x, 8);
            } else {
                const float f = (float)x;
                std::memcpy(&v[idx * 4], &f, 4);
            }
        };
        for (uint32_t r = 0; r < nrec; ++r) {
            uint32_t cell;
            if (spin) {
                const uint32_t bx = r % g->box_n[0], by = (r / g->box_n[0]) % g->box_n[1], bz = r / (g->box_n[0] * g->box_n[1]);
                cell = (uint32_t)(((uint64_t)(bz + g->box_lo[2]) * ny + (by + g->box_lo[1])) * nx + bx + g->box_lo[0]);
                h_mask[r] = 0;
                for (int k = 0; k < 9; ++k) h_q[(size_t)k * nrec + r] = 0x3C003C00u;   // q = 1
            } else {
                const simlab_wall_cell &w = g->walls[keep[r]];
                cell = w.cell;
                h_mask[r] = w.mask;
                for (int k = 0; k < 9; ++k)
                    h_q[(size_t)k * nrec + r] = (uint32_t)f16_bits_host(w.q[2 * k + 1]) | ((uint32_t)f16_bits_host(w.q[2 * k + 2]) << 16);
            }
            h_cell[r] = cell;
            for (int i = 1; i < 19; ++i) put(h_fpost, (size_t)(i - 1) * nrec + r, fe[i]);
            put(h_mac, r, 1.0);
            for (int k = 0; k < 3; ++k) put(h_mac, (size_t)(k + 1) * nrec + r, rb == 8 ? ui[k] : (double)(float)ui[k]);
            if ((h_flags[cell] & 3u) == 0) h_flags[cell] = (r + 1) << 2;   // record index + 1
        }
    }

    // probes
    std::vector<float4> h_verts;
    std::vector<int4> h_slices;
    if (pr) {
        for (uint64_t v = 0; v < pr->nvert; ++v) {
            const double *d = &pr->verts[6 * v];
            h_verts.push_back(make_float4((float)d[0], (float)d[1], (float)d[2], 0.0f));
            h_verts.push_back(make_float4((float)d[3], (float)d[4], (float)d[5], 0.0f));
        }
        h_slices.push_back(make_int4((int)pr->nslice, 0, 0, 0));
        for (uint64_t s = 0; s < pr->nslice; ++s) {
            const simlab_slice_def &d = pr->slices[s];
            for (int k = 0; k < 3; ++k) {
                const int64_t lo = (int64_t)d.origin[k], hi = lo + (int64_t)(d.nu ? d.nu - 1 : 0) * d.u[k] + (int64_t)(d.nv ? d.nv - 1 : 0) * d.v[k];
                if (std::min(lo, hi) < 0 || std::max(lo, hi) >= (int64_t)su->dims[k]) return fail(SIMLAB_ERR_INVALID, "slice leaves the grid");
            }
            h_slices.push_back(make_int4(d.origin[0], d.origin[1], d.origin[2], (int)d.nu));
            h_slices.push_back(make_int4(d.u[0], d.u[1], d.u[2], (int)d.nv));
            h_slices.push_back(make_int4(d.v[0], d.v[1], d.v[2], 0));
            lat->npix += d.nu * d.nv;
        }
        lat->nvert = (uint32_t)pr->nvert;
    }

    // device buffers
    size_t mem = 0;
    const size_t eb = elem_bytes(su->storage);
    try {
        lat->pops.alloc(pop_elems(n) * eb, mem);
        lat->flags.alloc(n * 4, mem);
        lat->rmask.alloc((size_t)nrec * 4, mem);
        lat->rcell.alloc((size_t)nrec * 4, mem);
        lat->rq.alloc((size_t)9 * nrec * 4, mem);
        lat->rfpost.alloc((size_t)18 * nrec * rb, mem);
        lat->rmac.alloc((size_t)4 * nrec * rb, mem);
        lat->rft.alloc((size_t)6 * nrec * rb, mem);
        lat->faces.alloc((size_t)lat->nface * 4, mem);
        lat->hist.alloc((size_t)KMAX * HIST_STRIDE * 8, mem);
        lat->counter.alloc(256, mem);
        lat->dump.alloc(su->debug_dump ? n * 16 : 16, mem);
        lat->mass_part.alloc(1024 * 8, mem);
        lat->mass_out.alloc(8, mem);
        lat->verts.alloc(h_verts.size() * sizeof(float4), mem);
        lat->vacc.alloc((size_t)lat->nvert * 8, mem);
        lat->slicedef.alloc(std::max<size_t>(1, h_slices.size()) * sizeof(int4), mem);
        lat->sacc.alloc((size_t)lat->npix * 20, mem);
    } catch (const HipError &e) {
        if (e.code == SIMLAB_ERR_OOM) {
            char b[160];
            std::snprintf(b, sizeof b, "not enough GPU memory for this grid (%.1f GB): %s", (double)(pop_elems(n) * eb + n * 4) / 1e9, e.what());
            return fail(SIMLAB_ERR_OOM, b);
        }
        throw;
    }
    hip_check(hipStreamCreateWithFlags(&lat->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&lat->stream2, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&lat->fork, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&lat->join, hipEventDisableTiming), "hipEventCreate");
    if (geom_kind == SIMLAB_GEOM_SPIN) {
        hip_check(hipStreamCreateWithFlags(&lat->stream3, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&lat->ev_vox, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&lat->ev_scat, hipEventDisableTiming), "hipEventCreate");
        if (const char *e = std::getenv("SIMLAB_CFD_SPIN_OVERLAP")) lat->spin_overlap = e[0] != '0';
    }
    for (uint64_t c = 0; c < n; ++c) lat->nbulk += (h_flags[c] == 0 || (h_flags[c] & FLAG_HALFWAY)) ? 1 : 0;
    hip_check(hipHostMalloc((void **)&lat->hist_host, (size_t)2 * KMAX * HIST_STRIDE * 8, hipHostMallocDefault), "hipHostMalloc");
    hipStream_t st = lat->stream;
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
        if (bytes) hip_check(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st), "upload");
    };
    up(lat->flags, h_flags.data(), n * 4);
    {
        // groups of POP_GROUP cells that are all plain fluid (flags 0) and
        // stay so: the swept box of a spinning part is re-voxelized every
        // step, so no group meeting it counts
        const uint64_t ng = (n + POP_GROUP - 1) / POP_GROUP;
        std::vector<uint32_t> pl((ng + 31) / 32, 0u);
        for (uint64_t g = 0; g < ng; ++g) {
            bool ok = true;
            const uint64_t c1 = std::min<uint64_t>(n, (g + 1) * POP_GROUP);
            for (uint64_t c = g * POP_GROUP; c < c1 && ok; ++c) ok = h_flags[c] == 0u;
            if (ok) pl[g >> 5] |= 1u << (g & 31);
        }
        if (geom_kind == SIMLAB_GEOM_SPIN) {
            const uint64_t nx = su->dims[0], ny = su->dims[1];
            for (uint64_t z = g->box_lo[2]; z < (uint64_t)g->box_lo[2] + g->box_n[2]; ++z)
                for (uint64_t y = g->box_lo[1]; y < (uint64_t)g->box_lo[1] + g->box_n[1]; ++y) {
                    const uint64_t r0 = (z * ny + y) * nx + g->box_lo[0];
                    for (uint64_t gg = r0 / POP_GROUP; gg <= (r0 + g->box_n[0] - 1) / POP_GROUP; ++gg)
                        pl[gg >> 5] &= ~(1u << (gg & 31));
                }
        }
        lat->plain.alloc(pl.size() * 4, mem);
        up(lat->plain, pl.data(), pl.size() * 4);
        const char *e = std::getenv("SIMLAB_CFD_PLAIN");
        lat->L.plain = (e && e[0] == '0') ? nullptr : lat->plain.as<uint32_t>();
        const char *sw = std::getenv("SIMLAB_CFD_SWIZZLE");
        lat->L.xcd_swizzle = (sw && sw[0] == '0') ? 0u : 1u;
        const char *ra = std::getenv("SIMLAB_CFD_ROWS");
        // 16- and 32-bit storage; every buffer offset of that addressing lies
        // within one z-plane of populations
        const uint64_t es = lat->storage == SIMLAB_STORAGE_FP32 ? 4 : lat->storage == SIMLAB_STORAGE_FP64 ? 8 : 2;
        const uint64_t plane = (uint64_t)lat->setup.dims[0] * lat->setup.dims[1] * 19 * es;
        lat->L.rows_aligned = (!(ra && ra[0] == '0') && es <= 4 && lat->setup.dims[0] % POP_GROUP == 0 &&
                               plane < (1ull << 30)) ? 1u : 0u;
    }
    up(lat->rmask, h_mask.data(), (size_t)nrec * 4);
    up(lat->rcell, h_cell.data(), (size_t)nrec * 4);
    up(lat->rq, h_q.data(), h_q.size() * 4);
    up(lat->rfpost, h_fpost.data(), h_fpost.size());
    up(lat->rmac, h_mac.data(), h_mac.size());
    up(lat->rft, h_ft.data(), h_ft.size());
    up(lat->faces, faces.data(), faces.size() * 4);
    up(lat->verts, h_verts.data(), h_verts.size() * sizeof(float4));
    up(lat->slicedef, h_slices.data(), h_slices.size() * sizeof(int4));
    hip_check(hipMemsetAsync(lat->counter.p, 0, 256, st), "memset");
    hip_check(hipMemsetAsync(lat->hist.p, 0, lat->hist.bytes, st), "memset");
    hip_check(hipMemsetAsync(lat->vacc.p, 0, lat->vacc.bytes, st), "memset");
    hip_check(hipMemsetAsync(lat->sacc.p, 0, lat->sacc.bytes, st), "memset");
    lat->memory_bytes = mem;

    DevLattice &L = lat->L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.ncell = (uint32_t)n;
    for (int a = 0; a < 3; ++a) L.per[a] = su->faces[2 * a] == SIMLAB_FACE_PERIODIC ? 1u : 0u;
    for (int f = 0; f < 6; ++f) L.vel_face[f] = su->faces[f] == SIMLAB_FACE_VELOCITY ? 1u : 0u;
    L.storage = su->storage;
    L.pops = lat->pops.p;
    L.flags = lat->flags.as<uint32_t>();
    L.nrec = nrec;
    L.rmask = lat->rmask.as<uint32_t>();
    L.rcell = lat->rcell.as<uint32_t>();
    L.rq = lat->rq.as<uint32_t>();
    L.rfpost = lat->rfpost.p;
    L.rmac = lat->rmac.p;
    L.rft = lat->rft.p;
    L.nface = lat->nface;
    L.faces = lat->faces.as<uint32_t>();
    L.hist = lat->hist.as<double>();
    L.tau0 = 0.5 + 3.0 * su->nu;   // lattice.rs:328
    L.smag2 = su->smagorinsky * su->smagorinsky;
    L.sponge_tau = su->sponge_tau;
    L.sigma = su->sponge_sigma;
    for (int k = 0; k < 3; ++k) {
        L.u_inf[k] = su->u_inf[k];
        L.u_init[k] = su->u_init[k];
        L.grav[k] = su->body_force[k];
        L.center[k] = su->center[k];
        L.sp_lo[k] = (double)su->sponge[2 * k];
        L.sp_hi[k] = (double)su->sponge[2 * k + 1];
        L.sp_ilo[k] = L.sp_lo[k] > 0.0 ? 1.0 / L.sp_lo[k] : 0.0;
        L.sp_ihi[k] = L.sp_hi[k] > 0.0 ? 1.0 / L.sp_hi[k] : 0.0;
    }
    L.abs_u = su->absorb_velocity != 0;
    L.total_p = su->total_pressure_inlet != 0;
    L.wall_model = su->wall_model != 0;
    L.moving = lat->moving ? 1 : 0;
    if (lat->moving) {
        // body-frame occupancy, triangles (v0, e1, e2) and the link scratch
        uint64_t m2 = 0;
        lat->occ.alloc(g->nocc_words * 4, m2);
        lat->tris.alloc(std::max<uint64_t>(1, g->ntri) * 3 * sizeof(float4), m2);
        lat->qscr.alloc((size_t)lat->nrec * 2 * 18 * 4, m2);
        mem += m2;
        hip_check(hipMemcpyAsync(lat->occ.p, g->occ_bits, g->nocc_words * 4, hipMemcpyHostToDevice, st), "upload");
        if (g->ntri)
            hip_check(hipMemcpyAsync(lat->tris.p, g->tris, g->ntri * 3 * sizeof(float4), hipMemcpyHostToDevice, st), "upload");
        for (int k = 0; k < 3; ++k) {
            L.box_lo[k] = g->box_lo[k];
            L.box_n[k] = g->box_n[k];
            L.occ_lo[k] = (float)g->occ_lo[k];
            L.occ_n[k] = (int32_t)g->occ_dims[k];
        }
        L.occ_invh = (float)(1.0 / g->occ_h);
        L.occ = lat->occ.as<uint32_t>();
        L.ntri = (uint32_t)g->ntri;
        L.tris = lat->tris.as<float4>();
        L.qscr = lat->qscr.as<uint32_t>();
        int32_t nn[3];
        std::vector<uint32_t> near = near_grid(*g, L.near_lo, nn);
        for (int k = 0; k < 3; ++k) L.near_n[k] = nn[k];
        L.near_words = (uint32_t)(near.size() / 2);
        lat->near.alloc(near.size() * 4, mem);
        hip_check(hipMemcpyAsync(lat->near.p, near.data(), near.size() * 4, hipMemcpyHostToDevice, st), "upload");
        hip_check(hipStreamSynchronize(st), "GPU wait failed");
        L.near = lat->near.as<uint32_t>();
    }
    L.forcing = (su->body_force[0] != 0.0 || su->body_force[1] != 0.0 || su->body_force[2] != 0.0) ? 1 : 0;
    L.sponge = 0;
    for (int k = 0; k < 6; ++k)
        if (su->sponge[k]) L.sponge = 1;
    L.nvert = lat->nvert;
    L.verts = lat->verts.as<float4>();
    L.vertex_offset = pr ? (float)pr->vertex_offset : 0.0f;
    L.vacc = lat->vacc.as<float>();
    L.npix = lat->npix;
    L.slicedef = lat->slicedef.as<int4>();
    L.sacc = lat->sacc.as<float>();
    if (const char *e = std::getenv("SIMLAB_CFD_LAUNCH")) lat->fused_launch = std::string(e) != "split";
    else lat->fused_launch = n <= FUSED_MAX_CELLS;
    choose_shape(L, default_block_threads(), 0, lat->fused_launch);
    {
        std::vector<uint32_t> hw;
        for (uint64_t c = 0; c < n; ++c)
            if (h_flags[c] & FLAG_HALFWAY) hw.push_back((uint32_t)c);
        L.nhw = (uint32_t)hw.size();
        L.hw_blocks = (L.nhw + SPECIAL_BLOCK - 1) / SPECIAL_BLOCK;
        lat->hw_cells.alloc(hw.size() * 4, mem);
        if (!hw.empty()) hip_check(hipMemcpyAsync(lat->hw_cells.p, hw.data(), hw.size() * 4, hipMemcpyHostToDevice, st), "upload");
        L.hw_cells = lat->hw_cells.as<uint32_t>();
    }
    // a spinning part's special kernel covers its whole swept box (2.1 M
    // cells for the 384^3 prop): beside the bulk kernel it ran until that one
    // ended or later, alone before it 85-100 us; the two-kernel step then
    // puts both on the main stream, 1 % faster in FP16 and FP32
    lat->concurrent = !lat->moving;
    if (const char *e = std::getenv("SIMLAB_CFD_SERIAL")) lat->concurrent = !(e[0] == '1');
    if (const char *e = std::getenv("SIMLAB_CFD_SIDE_SPECIAL")) lat->side_special = e[0] == '1' ? 1 : 0;
    lat->partials.alloc((size_t)KMAX * (L.wall_blocks + L.hw_blocks) * 8 * 8, mem);
    L.partials = lat->partials.as<double>();
    // x-face cells written by k_xfaces after each step: 16-bit storage and
    // lattices of up to FUSED_MAX_CELLS (the one-kernel step 9 % faster, the
    // whole step 4 %); in FP32, and on larger lattices, whose face rows have
    // left the caches by then (384^3: 10 % slower), it measured slower
    L.xdefer = (!L.per[0] && L.nx >= 4 && n <= FUSED_MAX_CELLS &&
                (su->storage == SIMLAB_STORAGE_FP16 || su->storage == SIMLAB_STORAGE_BF16)) ? 1u : 0u;
    if (const char *e = std::getenv("SIMLAB_CFD_XDEFER")) L.xdefer = L.xdefer && e[0] != '0';
    if (L.xdefer) {
        lat->xsrc.alloc((size_t)2 * L.nz * L.ny * 12 * sizeof(float), mem);
        hip_check(hipMemsetAsync(lat->xsrc.p, 0, lat->xsrc.bytes, st), "memset");
        L.xsrc = lat->xsrc.as<float>();
    } else {
        L.xsrc = nullptr;
    }
    lat->memory_bytes = mem;

    hip_check(launch_init(L, st), "init kernel");
    if (lat->moving) hip_check(launch_qclear(L, st), "link scratch clear");
    hip_check(hipStreamSynchronize(st), "setup");
    *out = lat.release();
    return SIMLAB_OK;
}

// Where a run keeps the per-block force partials and the per-chunk force
// history: the lattice's own buffers (kmax = KMAX steps per chunk), or an
// array the caller provides with room for kmax steps of both.
struct ForceScratch {
    double *partials = nullptr;   // [kmax][blocks][8]
    double *hist = nullptr;       // [kmax][HIST_STRIDE]
    size_t kmax = KMAX;
};

int run_steps(simlab_lattice *lat, const simlab_step_in *steps, size_t nsteps, double *ft, bool still,
              const ForceScratch *fs = nullptr) {
    if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
    if (!still && nsteps && !steps) return fail(SIMLAB_ERR_INVALID, "null steps");
    hip_check(hipSetDevice(lat->device), "hipSetDevice");
    // the kernels read the partials and history pointers from the lattice's
    // device view: point it at the given scratch for this call only
    struct ViewRestore {
        DevLattice &L;
        double *partials, *hist;
        ~ViewRestore() {
            L.partials = partials;
            L.hist = hist;
        }
    } restore{lat->L, lat->L.partials, lat->L.hist};
    size_t kmax = KMAX;
    if (fs) {
        lat->L.partials = fs->partials;
        lat->L.hist = fs->hist;
        kmax = fs->kmax;
    }
    simlab_step_in idn{};
    for (int k = 0; k < 3; ++k) {
        idn.rot[k][k] = 1.0;
        idn.rot_prev[k][k] = 1.0;
    }
    hipStream_t st = lat->stream;
    if (lat->scat_ready) {
        // hits of a step that a failed call never took: drop them
        hip_check(hipStreamSynchronize(lat->stream3), "GPU wait failed");
        hip_check(launch_qclear(lat->L, st), "hit scratch clear");
        lat->scat_ready = false;
    }
    if (lat->timing && lat->ev.size() < 2 * KMAX + 2) {
        while (lat->ev.size() < 2 * KMAX + 2) {
            hipEvent_t e;
            hip_check(hipEventCreate(&e), "hipEventCreate");
            lat->ev.push_back(e);
        }
    }
    lat->last_step_ms = 0.0;
    lat->last_main_ms = 0.0;
    lat->last_main_launches = 0;
    size_t done = 0;
    while (done < nsteps) {
        const size_t chunk = std::min<size_t>(kmax, nsteps - done);
        if (lat->timing) hip_check(hipEventRecord(lat->ev[2 * KMAX], st), "event");
        for (size_t k = 0; k < chunk; ++k) {
            const simlab_step_in &s = still ? idn : steps[done + k];
            const uint64_t step = lat->steps_done + k;
            const int par = (int)(step % 2);
            StepArgs a;
            pack_step(s, step, (uint32_t)k, a);
            const uint32_t tt = lat->L.bx * lat->L.by;
            const bool one_kernel = lat->fused_launch && tt <= SPECIAL_BLOCK && SPECIAL_BLOCK % tt == 0;
            // with timing on, the stream-collide kernel's dispatch records its
            // own start and end (no event packets between the kernels)
            hipEvent_t e0 = lat->timing ? lat->ev[2 * k] : nullptr, e1 = lat->timing ? lat->ev[2 * k + 1] : nullptr;
            // a spinning part is re-voxelized for this step's rotation first.
            // Its triangle hits come from stream3 when the previous step
            // launched them there; after the voxelization (which also clears
            // the previous step's hits), stream3 finds the next step's in the
            // other half of the hit scratch while this step's stream-collide
            // kernels run (none of them reads the hit scratch)
            if (lat->moving) {
                if (lat->scat_ready) {
                    hip_check(hipStreamWaitEvent(st, lat->ev_scat, 0), "event wait");
                    lat->scat_ready = false;
                } else {
                    hip_check(launch_scatter(lat->L, a, st), "re-voxelization kernels");
                }
                // the rotation of the previous voxelization lets it leave the
                // records of cells that stay far from the part as they are
                if (lat->vox_prev_valid) {
                    std::memcpy(a.vq, lat->vox_prev, sizeof a.vq);
                    a.vq_valid = 1;
                }
                // the hits of the step before this one in the call, in the
                // other half of the scratch, are cleared by this voxelization
                // (the scratch is clear between calls)
                a.hclear = done + k > 0 ? 1u : 0u;
                hip_check(launch_voxelize(lat->L, par, a, st), "re-voxelization kernels");
                std::memcpy(lat->vox_prev, a.r, sizeof lat->vox_prev);
                lat->vox_prev_valid = true;
                if (lat->spin_overlap && done + k + 1 < nsteps) {
                    StepArgs an;
                    pack_step(still ? idn : steps[done + k + 1], step + 1, (uint32_t)(k + 1), an);
                    hip_check(hipEventRecord(lat->ev_vox, st), "event");
                    hip_check(hipStreamWaitEvent(lat->stream3, lat->ev_vox, 0), "event wait");
                    hip_check(launch_scatter(lat->L, an, lat->stream3), "re-voxelization kernels");
                    hip_check(hipEventRecord(lat->ev_scat, lat->stream3), "event");
                    lat->scat_ready = true;
                } else if (done + k + 1 == nsteps) {
                    hip_check(launch_qclear_near(lat->L, a, st), "re-voxelization kernels");
                }
            }
            if (done + k == 0 && lat->fused_faces) {
                // first step of this call: rebuild this step's face inputs from
                // the stored state, then the next step's (constant where the
                // source cell is solid; overwritten by k_special otherwise). In
                // this order: with parity 0 the next step's face slots are the
                // face cells' own slots, which hold the source cells' parity-1
                // outputs toward them that the first gather reads.
                hip_check(launch_faces(lat->L, par, a.seed, st), "face kernel");
                hip_check(launch_faces(lat->L, 1 - par, (uint32_t)(step + 1) * 2654435761u, st), "face kernel");
            } else if (!lat->fused_faces) {
                hip_check(launch_faces(lat->L, par, a.seed, st), "face kernel");
            }
            if (one_kernel) {
                // special and bulk blocks in one launch: no cross-stream wait
                hip_check(launch_step_fused(lat->L, par, a, st, e0, e1), "stream-collide kernel");
            } else if (lat->concurrent) {
                // a special kernel of few blocks: k_bulk on the main stream,
                // right behind the previous step's work there, and k_special
                // on the side stream after the fork, its blocks taking slots
                // as bulk blocks end; the main stream then waits for the
                // special kernel, which has ended long before the bulk kernel,
                // so no cross-stream wait stands between one step's bulk kernel
                // and the next one's (12 instead of 23 us between them on the
                // 512x256x256 drone arm, +1.3 % MLUPS). A larger special
                // kernel goes on the main stream, dispatched first, so that its
                // blocks hold their slots before k_bulk fills the GPU and it
                // ends within the bulk kernel instead of after it
                const bool side = lat->side_special < 0 ? lat->L.wall_blocks + lat->L.hw_blocks <= SIDE_SPECIAL_BLOCKS
                                                        : lat->side_special == 1;
                hipStream_t s_sp = side ? lat->stream2 : st, s_bulk = side ? st : lat->stream2;
                hip_check(hipEventRecord(lat->fork, st), "event");
                hip_check(hipStreamWaitEvent(lat->stream2, lat->fork, 0), "event wait");
                hip_check(launch_stream_collide(lat->L, par, a, s_sp, s_bulk, e0, e1), "stream-collide kernels");
                hip_check(hipEventRecord(lat->join, lat->stream2), "event");
                hip_check(hipStreamWaitEvent(st, lat->join, 0), "event wait");
            } else {
                hip_check(launch_stream_collide(lat->L, par, a, st, st, e0, e1), "stream-collide kernels");
            }
            if (lat->L.xdefer) hip_check(launch_xfaces(lat->L, par, a, st), "x-face kernel");
            if (s.sample && (lat->nvert || lat->npix)) hip_check(launch_samplers(lat->L, par, a, st), "sampler kernels");
        }
        if (lat->timing) hip_check(hipEventRecord(lat->ev[2 * KMAX + 1], st), "event");
        hip_check(launch_reduce_hist(lat->L, (uint32_t)chunk, st), "force reduction kernel");
        hip_check(hipMemcpyAsync(lat->hist_host, lat->L.hist, chunk * HIST_STRIDE * 8, hipMemcpyDeviceToHost, st), "force copy");
        hip_check(hipStreamSynchronize(st), "GPU wait failed");
        if (lat->timing) {
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, lat->ev[2 * KMAX], lat->ev[2 * KMAX + 1]), "event time");
            lat->last_step_ms += ms;
            for (size_t k = 0; k < chunk; ++k) {
                hip_check(hipEventElapsedTime(&ms, lat->ev[2 * k], lat->ev[2 * k + 1]), "event time");
                lat->last_main_ms += ms;
            }
            lat->last_main_launches += chunk;
        }
        if (ft)
            for (size_t k = 0; k < chunk; ++k)
                for (int j = 0; j < 6; ++j) ft[(done + k) * 6 + j] = lat->hist_host[k * HIST_STRIDE + j];
        lat->steps_done += chunk;
        done += chunk;
    }
    return SIMLAB_OK;
}

}  // namespace

extern "C" {

const char *simlab_cfd_last_error(void) { return simlab::g_err.c_str(); }

int simlab_cfd_gpu_info(char *buf, size_t len) {
    return guarded([&]() -> int {
        if (!buf || len == 0) return fail(SIMLAB_ERR_INVALID, "null buffer");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            (void)hipGetLastError();
            return fail(SIMLAB_ERR_GPU, "no usable GPU: no HIP device found");
        }
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t p;
        hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
        // gpu.rs:56-63: "name · backend · GB max buffer". The device name can
        // be blank in the properties; then the architecture names the device
        std::string name(p.name, strnlen(p.name, sizeof p.name));
        while (!name.empty() && name.back() == ' ') name.pop_back();
        if (name.empty()) {
            std::string arch(p.gcnArchName, strnlen(p.gcnArchName, sizeof p.gcnArchName));
            arch = arch.substr(0, arch.find(':'));
            name = "AMD " + (arch.empty() ? std::string("GPU") : arch);
        }
        // the largest buffer one allocation can have now: the free device
        // memory (a HIP allocation is not capped below it)
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        std::snprintf(buf, len, "%s \xc2\xb7 HIP \xc2\xb7 %.1f GB max buffer", name.c_str(), (double)free_b / 1e9);
        return SIMLAB_OK;
    });
}

uint64_t simlab_cfd_max_cells(int32_t storage) {
    // layout limit: 32-bit cell indices with signed neighbour offsets
    (void)storage;
    return (1ull << 31) - 1;
}

int simlab_cfd_lattice_new(const simlab_lattice_setup *setup, const simlab_geometry *geom, const simlab_probes *probes,
                           simlab_lattice **out) {
    return guarded([&]() -> int {
        try {
            return new_lattice(setup, geom, probes, out);
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

void simlab_cfd_lattice_free(simlab_lattice *lat) {
    if (!lat) return;
    (void)hipSetDevice(lat->device);
    delete lat;
}

int simlab_cfd_lattice_info(const simlab_lattice *lat, simlab_lattice_info *out) {
    if (!lat || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof *out);
    out->steps_done = lat->steps_done;
    out->cells = lat->ncell;
    out->memory_bytes = lat->memory_bytes;
    for (int k = 0; k < 3; ++k) out->dims[k] = lat->setup.dims[k];
    out->nrec = lat->nrec;
    out->nface = lat->nface;
    out->npix = lat->npix;
    out->nvert = lat->nvert;
    out->storage = lat->storage;
    out->moving = lat->moving ? 1 : 0;
    out->tau0 = lat->L.tau0;
    out->bulk_cells = lat->nbulk;
    out->halfway_cells = lat->nhalfway;
    return SIMLAB_OK;
}

int simlab_cfd_lattice_run(simlab_lattice *lat, const simlab_step_in *steps, size_t nsteps, double *ft_out) {
    return guarded([&]() -> int {
        try {
            return run_steps(lat, steps, nsteps, ft_out, false);
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_run_still(simlab_lattice *lat, size_t nsteps, double *ft_out) {
    return guarded([&]() -> int {
        try {
            return run_steps(lat, nullptr, nsteps, ft_out, true);
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free run
namespace {
// The force partials and history of simlab_cfd_lattice_run_noalloc: a
// file-scope device array (each GPU the library runs on holds its own copy
// until the process ends). Host code never uses its name as an address: the
// address on each device comes from hipGetSymbolAddress, taken once per
// device on the entry point's first call there.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// the device's address of `scratch`; the device must be current
double *scratch_on(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) throw HipError(SIMLAB_ERR_GPU, "run_noalloc: device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int simlab_cfd_lattice_run_noalloc(simlab_lattice *lat, const simlab_step_in *steps, size_t nsteps, double *ft_out) {
    return guarded([&]() -> int {
        if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            // per step: one 8-double partial per wall block, one history row
            const size_t blocks = (size_t)lat->L.wall_blocks + lat->L.hw_blocks;
            const size_t per_step = (blocks * 8 + HIST_STRIDE) * sizeof(double);
            const size_t kmax = std::min<size_t>(KMAX, NOALLOC_SCRATCH_BYTES / per_step);
            if (kmax == 0)
                return fail(SIMLAB_ERR_INVALID, "run_noalloc: the force partials of one step (" +
                                                    std::to_string(per_step) + " bytes for " + std::to_string(blocks) +
                                                    " wall blocks) exceed the " +
                                                    std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB device scratch");
            double *base = scratch_on(lat->device);
            ForceScratch fs;
            fs.partials = base;
            fs.hist = base + kmax * blocks * 8;
            fs.kmax = kmax;
            // timing would create events on first use: off for this call
            struct TimingOff {
                simlab_lattice *l;
                bool was;
                ~TimingOff() { l->timing = was; }
            } off{lat, lat->timing};
            lat->timing = false;
            return run_steps(lat, steps, nsteps, ft_out, false, &fs);
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_macros(simlab_lattice *lat, float *out) {
    return guarded([&]() -> int {
        if (!lat || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
        if (!lat->setup.debug_dump) return fail(SIMLAB_ERR_STATE, "macros() needs debug_dump");   // lattice.rs:820-822
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            const int par = (int)(lat->steps_done % 2);   // lattice.rs:824
            hip_check(launch_macros(lat->L, par, lat->dump.as<float4>(), lat->stream), "dump kernel");
            hip_check(hipMemcpyAsync(out, lat->dump.p, lat->ncell * 16, hipMemcpyDeviceToHost, lat->stream), "GPU read failed");
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_records(simlab_lattice *lat, uint32_t *out) {
    return guarded([&]() -> int {
        if (!lat || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            const size_t nr = lat->nrec, rb = real_bytes(lat->storage);
            std::vector<uint32_t> mask(nr), cell(nr), q(9 * nr);
            std::vector<unsigned char> fpost(18 * nr * rb), mac(4 * nr * rb), ftv(6 * nr * rb);
            hipStream_t st = lat->stream;
            auto down = [&](void *dst, const DevBuf &b, size_t bytes) {
                if (bytes) hip_check(hipMemcpyAsync(dst, b.p, bytes, hipMemcpyDeviceToHost, st), "GPU read failed");
            };
            down(mask.data(), lat->rmask, nr * 4);
            down(cell.data(), lat->rcell, nr * 4);
            down(q.data(), lat->rq, 9 * nr * 4);
            down(fpost.data(), lat->rfpost, fpost.size());
            down(mac.data(), lat->rmac, mac.size());
            down(ftv.data(), lat->rft, ftv.size());
            hip_check(hipStreamSynchronize(st), "GPU wait failed");
            auto getf = [&](const std::vector<unsigned char> &v, size_t idx) -> uint32_t {
                float f;
                if (rb == 8) {
                    double d;
                    std::memcpy(&d, &v[idx * 8], 8);
                    f = (float)d;
                } else {
                    std::memcpy(&f, &v[idx * 4], 4);
                }
                uint32_t u;
                std::memcpy(&u, &f, 4);
                return u;
            };
            const size_t nout = std::max<size_t>(nr, 1);
            std::memset(out, 0, nout * 40 * 4);
            for (size_t r = 0; r < nr; ++r) {   // shader.rs:62-69 layout
                uint32_t *w = out + r * 40;
                w[0] = mask[r];
                w[1] = cell[r];
                for (int k = 0; k < 9; ++k) w[2 + k] = q[(size_t)k * nr + r];
                for (int i = 0; i < 18; ++i) w[11 + i] = getf(fpost, (size_t)i * nr + r);
                for (int k = 0; k < 4; ++k) w[29 + k] = getf(mac, (size_t)k * nr + r);
                for (int k = 0; k < 6; ++k) w[33 + k] = getf(ftv, (size_t)k * nr + r);
            }
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_reset_samples(simlab_lattice *lat) {
    return guarded([&]() -> int {
        if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            hip_check(hipMemsetAsync(lat->vacc.p, 0, lat->vacc.bytes, lat->stream), "memset");
            hip_check(hipMemsetAsync(lat->sacc.p, 0, lat->sacc.bytes, lat->stream), "memset");
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_vertex_means(simlab_lattice *lat, float *out) {
    return guarded([&]() -> int {
        if (!lat || (!out && lat->nvert)) return fail(SIMLAB_ERR_INVALID, "null argument");
        if (!lat->nvert) return SIMLAB_OK;
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            std::vector<float> v((size_t)lat->nvert * 2);
            hip_check(hipMemcpyAsync(v.data(), lat->vacc.p, v.size() * 4, hipMemcpyDeviceToHost, lat->stream), "GPU read failed");
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            for (uint32_t i = 0; i < lat->nvert; ++i) out[i] = v[2 * i + 1] > 0.0f ? v[2 * i] / v[2 * i + 1] : std::nanf("");
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_slice_means(simlab_lattice *lat, float *out) {
    return guarded([&]() -> int {
        if (!lat || (!out && lat->npix)) return fail(SIMLAB_ERR_INVALID, "null argument");
        if (!lat->npix) return SIMLAB_OK;
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            std::vector<float> v((size_t)lat->npix * 5);
            hip_check(hipMemcpyAsync(v.data(), lat->sacc.p, v.size() * 4, hipMemcpyDeviceToHost, lat->stream), "GPU read failed");
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            for (uint32_t i = 0; i < lat->npix; ++i) {
                const float c = v[5 * i + 4];
                for (int k = 0; k < 4; ++k) out[4 * i + k] = c > 0.0f ? v[5 * i + k] / c : std::nanf("");
            }
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_mass(simlab_lattice *lat, double *mass) {
    return guarded([&]() -> int {
        if (!lat || !mass) return fail(SIMLAB_ERR_INVALID, "null argument");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            const int par = (int)(lat->steps_done % 2);
            hip_check(launch_mass(lat->L, par, lat->mass_part.as<double>(), 1024, lat->mass_out.as<double>(), lat->stream), "mass kernel");
            hip_check(hipMemcpyAsync(mass, lat->mass_out.p, 8, hipMemcpyDeviceToHost, lat->stream), "GPU read failed");
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_get_pops(simlab_lattice *lat, void *out, size_t bytes) {
    return guarded([&]() -> int {
        if (!lat || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
        const size_t eb = elem_bytes(lat->storage);
        const size_t want = (size_t)19 * lat->ncell * eb;
        if (bytes != want) return fail(SIMLAB_ERR_INVALID, "population buffer has the wrong size");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            std::vector<char> dev(pop_elems(lat->ncell) * eb);
            hip_check(hipMemcpyAsync(dev.data(), lat->pops.p, dev.size(), hipMemcpyDeviceToHost, lat->stream), "GPU read failed");
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            reorder_pops(lat->ncell, eb, dev.data(), (char *)out, true);
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_set_pops(simlab_lattice *lat, const void *in, size_t bytes) {
    return guarded([&]() -> int {
        if (!lat || !in) return fail(SIMLAB_ERR_INVALID, "null argument");
        const size_t eb = elem_bytes(lat->storage);
        const size_t want = (size_t)19 * lat->ncell * eb;
        if (bytes != want) return fail(SIMLAB_ERR_INVALID, "population buffer has the wrong size");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            std::vector<char> dev(pop_elems(lat->ncell) * eb, 0);
            reorder_pops(lat->ncell, eb, dev.data(), (char *)const_cast<void *>(in), false);
            hip_check(hipMemcpyAsync(lat->pops.p, dev.data(), dev.size(), hipMemcpyHostToDevice, lat->stream), "GPU write failed");
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_set_timing(simlab_lattice *lat, int enable) {
    if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
    lat->timing = enable != 0;
    return SIMLAB_OK;
}

int simlab_cfd_lattice_timing(const simlab_lattice *lat, double *step_ms, double *main_ms, uint64_t *main_launches) {
    if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
    if (step_ms) *step_ms = lat->last_step_ms;
    if (main_ms) *main_ms = lat->last_main_ms;
    if (main_launches) *main_launches = lat->last_main_launches;
    return SIMLAB_OK;
}

// ---------------------------------------------------------------- autotuner
}  // extern "C"
namespace {

// a tuned choice holds for lattices of the same size and storage that run
// the same kernel variant (walls, moving part, body force, sponge)
struct TuneKey {
    uint32_t nx, ny, nz;
    int32_t storage;
    bool walls, moving, forcing, sponge;
    bool operator<(const TuneKey &o) const {
        return std::tie(nx, ny, nz, storage, walls, moving, forcing, sponge) <
               std::tie(o.nx, o.ny, o.nz, o.storage, o.walls, o.moving, o.forcing, o.sponge);
    }
};
std::mutex g_tune_mu;
struct Tuned {
    uint32_t bx, by;
    bool fused;
};
std::map<TuneKey, Tuned> g_tune;   // best tile and launch per lattice size

// tiles to try: the lattice's current tile first, then x extents of a whole
// row rounded to a wave, its whole-wave divisors (384: 192, 128, 64) and 64,
// 128, 256, with rows per tile so a tile has 64 .. 1024 threads
std::vector<std::pair<uint32_t, uint32_t>> tile_candidates(const DevLattice &L) {
    std::vector<std::pair<uint32_t, uint32_t>> out{{L.bx, L.by}};
    const uint32_t row = std::min<uint32_t>(1024, ((L.nx + 63) / 64) * 64);
    std::vector<uint32_t> widths;
    for (uint32_t k = 1; k <= row / 64; ++k)
        if (row % k == 0 && (row / k) % 64 == 0) widths.push_back(row / k);
    for (uint32_t bx : {64u, 128u, 256u})
        if (bx < row && std::find(widths.begin(), widths.end(), bx) == widths.end()) widths.push_back(bx);
    for (uint32_t bx : widths) {
        for (uint32_t t = 64; t <= 1024; t *= 2) {
            if (t < bx) continue;
            const uint32_t by = std::min<uint32_t>(t / bx, L.ny);
            if (std::find(out.begin(), out.end(), std::make_pair(bx, by)) == out.end()) out.push_back({bx, by});
        }
    }
    return out;
}

// Keeps the evolving state of a lattice (every buffer a checkpoint holds,
// plus a spinning part's hit scratch) in device copies while the autotuner
// steps it, and restores it, steps_done, timing, tile and launch mode when it
// goes out of scope; keep_tile() makes a new tile and launch mode stay.
struct TuneGuard {
    simlab_lattice *lat;
    std::vector<DevBuf *> bufs;
    std::vector<std::unique_ptr<DevBuf>> saved;
    uint64_t steps0;
    bool timing0, fused0;
    uint32_t bx0, by0;
    bool vox_valid0;
    float vox0[3][4];
    explicit TuneGuard(simlab_lattice *l)
        : lat(l), steps0(l->steps_done), timing0(l->timing), fused0(l->fused_launch), bx0(l->L.bx), by0(l->L.by),
          vox_valid0(l->vox_prev_valid) {
        std::memcpy(vox0, l->vox_prev, sizeof vox0);
        bufs = {&l->pops, &l->flags, &l->rmask, &l->rq, &l->rfpost, &l->rmac, &l->rft, &l->vacc, &l->sacc, &l->qscr};
        uint64_t mem = 0;
        for (DevBuf *b : bufs) {
            saved.emplace_back(new DevBuf);
            if (b->p) {
                saved.back()->alloc(b->bytes, mem);
                hip_check(hipMemcpyAsync(saved.back()->p, b->p, b->bytes, hipMemcpyDeviceToDevice, l->stream), "copy");
            }
        }
    }
    void put_back_state() {
        if (lat->scat_ready) {
            // hits of a next step the trial launched on the third stream
            hip_check(hipStreamSynchronize(lat->stream3), "GPU wait failed");
            lat->scat_ready = false;
        }
        for (size_t k = 0; k < bufs.size(); ++k)
            if (bufs[k]->p)
                hip_check(hipMemcpyAsync(bufs[k]->p, saved[k]->p, bufs[k]->bytes, hipMemcpyDeviceToDevice, lat->stream),
                          "copy");
        lat->steps_done = steps0;
        lat->vox_prev_valid = vox_valid0;
        std::memcpy(lat->vox_prev, vox0, sizeof vox0);
    }
    void keep_tile(uint32_t bx, uint32_t by, bool fused) {
        bx0 = bx;
        by0 = by;
        fused0 = fused;
    }
    ~TuneGuard() {
        try {
            put_back_state();
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
        } catch (...) {
        }
        lat->timing = timing0;
        set_shape(lat->L, bx0, by0);
        lat->fused_launch = fused0;
    }
};

}  // namespace
extern "C" {

int simlab_cfd_lattice_autotune(simlab_lattice *lat, uint32_t steps, int use_cache, uint32_t *tile_out, double *step_us_out) {
    return guarded([&]() -> int {
        if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
        if (steps == 0) steps = 20;
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            DevLattice &L = lat->L;
            const TuneKey key{L.nx, L.ny, L.nz, L.storage, lat->nrec != 0 || lat->nhalfway != 0, lat->moving,
                              L.forcing != 0, L.sponge != 0};
            if (use_cache) {
                std::lock_guard<std::mutex> g(g_tune_mu);
                auto it = g_tune.find(key);
                if (it != g_tune.end()) {
                    set_shape(L, it->second.bx, it->second.by);
                    lat->fused_launch = it->second.fused;
                    if (tile_out) { tile_out[0] = L.bx; tile_out[1] = L.by; }
                    if (step_us_out) *step_us_out = 0.0;
                    return SIMLAB_OK;
                }
            }
            // the trial steps advance the state: keep every buffer a
            // checkpoint holds and put it back after each trial; on any exit
            // (a failed trial throws) the guard restores the state, steps_done,
            // timing, tile and launch mode
            hipStream_t st = lat->stream;
            TuneGuard guard(lat);
            lat->timing = true;
            double best = 1e30;
            Tuned best_tile{L.bx, L.by, lat->fused_launch};
            for (const auto &c : tile_candidates(L)) {
                for (int fz = 0; fz < 2; ++fz) {
                    const uint32_t tt = c.first * c.second;
                    if (fz && (tt > SPECIAL_BLOCK || SPECIAL_BLOCK % tt != 0)) continue;
                    set_shape(L, c.first, c.second);
                    lat->fused_launch = fz != 0;
                    const int r1 = run_steps(lat, nullptr, 2, nullptr, true);   // warm-up
                    const int r2 = r1 == SIMLAB_OK ? run_steps(lat, nullptr, steps, nullptr, true) : r1;
                    if (r2 == SIMLAB_OK) {
                        const double us = lat->last_step_ms * 1e3 / steps;
                        if (us < best) {
                            best = us;
                            best_tile = Tuned{c.first, c.second, fz != 0};
                        }
                    }
                    guard.put_back_state();
                }
            }
            hip_check(hipStreamSynchronize(st), "GPU wait failed");
            guard.keep_tile(best_tile.bx, best_tile.by, best_tile.fused);
            {
                std::lock_guard<std::mutex> g(g_tune_mu);
                g_tune[key] = best_tile;
            }
            if (tile_out) { tile_out[0] = best_tile.bx; tile_out[1] = best_tile.by; }
            if (step_us_out) *step_us_out = best;
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

}  // extern "C"
namespace {
// the buffers a checkpoint holds, in order
std::vector<DevBuf *> ckpt_bufs(simlab_lattice *lat) {
    return {&lat->pops, &lat->flags, &lat->rmask, &lat->rq, &lat->rfpost, &lat->rmac, &lat->rft, &lat->vacc, &lat->sacc};
}
struct CkptHead {
    char magic[8];
    uint32_t dims[3];
    int32_t storage;
    uint32_t nrec, nbuf;
    uint64_t steps_done;
    uint64_t sizes[9];
};
}  // namespace
extern "C" {

int simlab_cfd_lattice_checkpoint_size(const simlab_lattice *lat, size_t *bytes) {
    if (!lat || !bytes) return fail(SIMLAB_ERR_INVALID, "null argument");
    size_t b = sizeof(CkptHead);
    for (DevBuf *d : ckpt_bufs(const_cast<simlab_lattice *>(lat))) b += d->bytes;
    *bytes = b;
    return SIMLAB_OK;
}

int simlab_cfd_lattice_checkpoint(simlab_lattice *lat, void *out, size_t bytes) {
    return guarded([&]() -> int {
        if (!lat || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
        size_t want = 0;
        simlab_cfd_lattice_checkpoint_size(lat, &want);
        if (bytes != want) return fail(SIMLAB_ERR_INVALID, "checkpoint buffer has the wrong size");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            CkptHead h{};
            std::memcpy(h.magic, "SLCKPT1", 8);
            for (int k = 0; k < 3; ++k) h.dims[k] = lat->setup.dims[k];
            h.storage = lat->storage;
            h.nrec = lat->nrec;
            const auto bufs = ckpt_bufs(lat);
            h.nbuf = (uint32_t)bufs.size();
            h.steps_done = lat->steps_done;
            char *o = (char *)out + sizeof(CkptHead);
            for (size_t j = 0; j < bufs.size(); ++j) {
                h.sizes[j] = bufs[j]->bytes;
                if (bufs[j]->bytes)
                    hip_check(hipMemcpyAsync(o, bufs[j]->p, bufs[j]->bytes, hipMemcpyDeviceToHost, lat->stream), "GPU read failed");
                o += bufs[j]->bytes;
            }
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            std::memcpy(out, &h, sizeof h);
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_restore(simlab_lattice *lat, const void *in, size_t bytes) {
    return guarded([&]() -> int {
        if (!lat || !in) return fail(SIMLAB_ERR_INVALID, "null argument");
        size_t want = 0;
        simlab_cfd_lattice_checkpoint_size(lat, &want);
        CkptHead h;
        if (bytes != want || bytes < sizeof h) return fail(SIMLAB_ERR_INVALID, "checkpoint does not fit this lattice");
        std::memcpy(&h, in, sizeof h);
        const auto bufs = ckpt_bufs(lat);
        bool ok = std::memcmp(h.magic, "SLCKPT1", 8) == 0 && h.storage == lat->storage && h.nrec == lat->nrec &&
                  h.nbuf == bufs.size();
        for (int k = 0; k < 3; ++k) ok = ok && h.dims[k] == lat->setup.dims[k];
        for (size_t j = 0; ok && j < bufs.size(); ++j) ok = h.sizes[j] == bufs[j]->bytes;
        if (!ok) return fail(SIMLAB_ERR_INVALID, "checkpoint does not fit this lattice");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            const char *o = (const char *)in + sizeof(CkptHead);
            for (DevBuf *d : bufs) {
                if (d->bytes) hip_check(hipMemcpyAsync(d->p, o, d->bytes, hipMemcpyHostToDevice, lat->stream), "GPU write failed");
                o += d->bytes;
            }
            hip_check(hipStreamSynchronize(lat->stream), "GPU wait failed");
            lat->steps_done = h.steps_done;
            lat->adj_ready = false;
            lat->vox_prev_valid = false;   // the next voxelization writes every box cell
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_lattice_get_tile(const simlab_lattice *lat, uint32_t *tile_out) {
    if (!lat || !tile_out) return fail(SIMLAB_ERR_INVALID, "null argument");
    tile_out[0] = lat->L.bx;
    tile_out[1] = lat->L.by;
    return SIMLAB_OK;
}

int simlab_cfd_lattice_set_tile(simlab_lattice *lat, uint32_t bx, uint32_t by) {
    if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
    if (bx == 0 || bx % 64 || bx > 1024 || by == 0 || bx * by > 1024)
        return fail(SIMLAB_ERR_INVALID, "tile must be a multiple of 64 cells wide and at most 1024 threads");
    set_shape(lat->L, bx, std::min<uint32_t>(by, lat->L.ny));
    return SIMLAB_OK;
}

// ---------------------------------------------------------------- batches
// lattices of more cells than this in a batch step with their own launches
constexpr uint64_t BATCH_SOLO_CELLS = 4u << 20;

struct simlab_batch {
    std::vector<simlab_lattice *> lats;
    DevBuf dlats, boffs;
    hipStream_t stream = nullptr;    // faces, special kernel, copies
    hipStream_t stream2 = nullptr;   // bulk kernel, concurrent with the special kernel
    hipEvent_t fork = nullptr, join = nullptr;
    int device = 0;
    ~simlab_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream2) (void)hipStreamSynchronize(stream2);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

int simlab_cfd_batch_new(simlab_lattice *const *lats, size_t n, simlab_batch **out) {
    return guarded([&]() -> int {
        if (!lats || !out || n == 0) return fail(SIMLAB_ERR_INVALID, "empty batch");
        *out = nullptr;
        for (size_t i = 0; i < n; ++i) {
            if (!lats[i]) return fail(SIMLAB_ERR_INVALID, "null lattice in batch");
            const simlab_lattice &a = *lats[i], &b = *lats[0];
            if (a.storage != b.storage) return fail(SIMLAB_ERR_INVALID, "a batch needs one population storage");
            if (a.L.forcing != b.L.forcing || a.L.sponge != b.L.sponge)
                return fail(SIMLAB_ERR_INVALID, "a batch needs the same body-force and sponge settings");
            if (a.device != b.device) return fail(SIMLAB_ERR_INVALID, "a batch lives on one GPU");
            if (a.moving) return fail(SIMLAB_ERR_INVALID, "a batch takes static geometry only");
            for (size_t j = 0; j < i; ++j)
                if (lats[j] == lats[i]) return fail(SIMLAB_ERR_INVALID, "a lattice appears twice in the batch");
        }
        try {
            auto b = std::make_unique<simlab_batch>();
            b->lats.assign(lats, lats + n);
            b->device = lats[0]->device;
            hip_check(hipSetDevice(b->device), "hipSetDevice");
            uint64_t mem = 0;
            b->dlats.alloc(n * sizeof(DevLattice), mem);
            b->boffs.alloc(2 * (n + 1) * 4, mem);
            hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&b->join, hipEventDisableTiming), "hipEventCreate");
            *out = b.release();
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

void simlab_cfd_batch_free(simlab_batch *b) { delete b; }

int simlab_cfd_batch_run_still(simlab_batch *b, size_t nsteps, double *ft_out) {
    return guarded([&]() -> int {
        if (!b) return fail(SIMLAB_ERR_INVALID, "null batch");
        const size_t n = b->lats.size();
        for (size_t i = 1; i < n; ++i)
            if (b->lats[i]->steps_done != b->lats[0]->steps_done)
                return fail(SIMLAB_ERR_INVALID, "the lattices of a batch must have done the same number of steps");
        try {
            hip_check(hipSetDevice(b->device), "hipSetDevice");
            hipStream_t st = b->stream;
            // One kernel per step for every member (the default), members
            // ordered largest first so the small ones' tiles fill the large
            // ones' tail; or (SIMLAB_CFD_BATCH=split) a special and a bulk
            // kernel on two streams, members above BATCH_SOLO_CELLS stepping
            // with their own launches.
            const char *mode = std::getenv("SIMLAB_CFD_BATCH");
            const bool fused = !(mode && std::string(mode) == "split");
            std::vector<size_t> mi;
            for (size_t i = 0; i < n; ++i) {
                simlab_lattice *l = b->lats[i];
                if (fused || l->ncell <= BATCH_SOLO_CELLS) {
                    mi.push_back(i);
                    continue;
                }
                const int rc = run_steps(l, nullptr, nsteps, ft_out ? ft_out + i * nsteps * 6 : nullptr, true);
                if (rc != SIMLAB_OK) return rc;
            }
            if (mi.empty()) return SIMLAB_OK;
            if (fused)
                std::stable_sort(mi.begin(), mi.end(),
                                 [&](size_t x, size_t y) { return b->lats[x]->ncell > b->lats[y]->ncell; });
            std::vector<simlab_lattice *> ml;
            for (size_t i : mi) ml.push_back(b->lats[i]);
            const size_t n = ml.size();   // the batched lattices from here on
            // device views and block offsets as the lattices stand now
            std::vector<DevLattice> hl(n);
            std::vector<uint32_t> boff(2 * (n + 1), 0);
            const uint32_t threads = SPECIAL_BLOCK;
            // in the one-kernel batch a member that steps with two kernels on
            // its own (by default those of more than FUSED_MAX_CELLS cells)
            // keeps its bulk kernel,
            // at its own tile, on the second stream beside the batched kernel
            // (which holds its special blocks); the others keep their x-face
            // writes deferred to the small kernel after the step
            std::vector<char> own_bulk(n, 0);
            for (size_t i = 0; i < n; ++i) {
                hip_check(hipStreamSynchronize(ml[i]->stream), "GPU wait failed");
                hl[i] = ml[i]->L;
                {
                    const uint32_t t0 = hl[i].bx * hl[i].by;
                    own_bulk[i] = fused && !(ml[i]->fused_launch && t0 <= threads && threads % t0 == 0);
                }
                // the two-kernel batch writes every face cell in the step
                if (!fused) hl[i].xdefer = 0;
                if (own_bulk[i]) {
                    boff[i + 1] = boff[i];
                    boff[n + 1 + i + 1] = boff[n + 1 + i] + hl[i].wall_blocks + hl[i].hw_blocks;
                    continue;
                }
                // the batch's blocks have SPECIAL_BLOCK threads: a tile that
                // does not divide that is replaced (in the batch's copy):
                // 128 x 1 for 16-bit storage, else whole rows of up to 256
                // cells (or 128-cell pieces) stacked to 256 threads
                uint32_t tt = hl[i].bx * hl[i].by;
                if (tt > threads || threads % tt != 0) {
                    const bool half = hl[i].storage == SIMLAB_STORAGE_FP16 || hl[i].storage == SIMLAB_STORAGE_BF16;
                    uint32_t bx = std::min<uint32_t>(((hl[i].nx + 63) / 64) * 64, threads);
                    if (threads % bx != 0 || (half && bx > 128)) bx = 128;
                    uint32_t by = 1;
                    while (!half && by * 2 * bx <= threads && by * 2 <= hl[i].ny) by *= 2;
                    set_shape(hl[i], bx, by);
                    tt = bx * by;
                }
                const uint64_t tiles = (uint64_t)hl[i].xblocks * hl[i].yblocks * hl[i].nz;
                boff[i + 1] = boff[i] + (uint32_t)((tiles + threads / tt - 1) / (threads / tt));
                boff[n + 1 + i + 1] = boff[n + 1 + i] + hl[i].wall_blocks + hl[i].hw_blocks;
            }
            if (fused) {
                // the one-kernel batch reads [special offsets][bulk offsets]
                std::vector<uint32_t> f(boff.size());
                std::copy(boff.begin() + (n + 1), boff.end(), f.begin());
                std::copy(boff.begin(), boff.begin() + (n + 1), f.begin() + (n + 1));
                boff.swap(f);
            }
            hip_check(hipMemcpyAsync(b->dlats.p, hl.data(), n * sizeof(DevLattice), hipMemcpyHostToDevice, st), "upload");
            hip_check(hipMemcpyAsync(b->boffs.p, boff.data(), boff.size() * 4, hipMemcpyHostToDevice, st), "upload");
            BatchLaunch bl;
            bl.lats = b->dlats.as<DevLattice>();
            bl.fused = fused;
            bl.rows = true;
            for (size_t i = 0; i < n; ++i) bl.rows = bl.rows && hl[i].rows_aligned != 0;
            if (fused) {
                bl.boff_sp = b->boffs.as<uint32_t>();
                bl.boff_bulk = b->boffs.as<uint32_t>() + (n + 1);
                bl.nsp = boff[n];
                bl.nbulk = boff[2 * n + 1];
            } else {
                bl.boff_bulk = b->boffs.as<uint32_t>();
                bl.boff_sp = b->boffs.as<uint32_t>() + (n + 1);
                bl.nbulk = boff[n];
                bl.nsp = boff[n + 1 + n];
            }
            bl.nl = (uint32_t)n;
            bl.threads = threads;
            bl.storage = hl[0].storage;
            bl.force = hl[0].forcing != 0;
            bl.sponge = hl[0].sponge != 0;
            simlab_step_in idn{};
            for (int k = 0; k < 3; ++k) idn.rot[k][k] = idn.rot_prev[k][k] = 1.0;
            const uint64_t step0 = ml[0]->steps_done;
            size_t done = 0;
            while (done < nsteps) {
                const size_t chunk = std::min<size_t>(KMAX, nsteps - done);
                for (size_t k = 0; k < chunk; ++k) {
                    const uint64_t step = step0 + done + k;
                    const int par = (int)(step % 2);
                    StepArgs a;
                    pack_step(idn, step, (uint32_t)k, a);
                    if (done + k == 0)
                        for (size_t i = 0; i < n; ++i) {
                            // as at the start of a lattice's own run: this step's
                            // face inputs, then the next step's
                            hip_check(launch_faces(hl[i], par, a.seed, st), "face kernel");
                            hip_check(launch_faces(hl[i], 1 - par, (uint32_t)(step + 1) * 2654435761u, st), "face kernel");
                        }
                    if (fused) {
                        bool forked = false;
                        for (size_t i = 0; i < n; ++i) {
                            if (!own_bulk[i]) continue;
                            if (!forked) {
                                hip_check(hipEventRecord(b->fork, st), "event");
                                hip_check(hipStreamWaitEvent(b->stream2, b->fork, 0), "event wait");
                                forked = true;
                            }
                            hip_check(launch_bulk_planes(hl[i], par, a, b->stream2, 0, hl[i].nz, hl[i].nz, 0),
                                      "stream-collide kernels");
                        }
                        hip_check(launch_batch_step(bl, par, a, st, st), "batched stream-collide kernel");
                        if (forked) {
                            hip_check(hipEventRecord(b->join, b->stream2), "event");
                            hip_check(hipStreamWaitEvent(st, b->join, 0), "event wait");
                        }
                        for (size_t i = 0; i < n; ++i)
                            if (hl[i].xdefer)
                                hip_check(launch_xfaces(hl[i], par, a, st), "x-face kernel");
                    } else {
                        // special kernel first on the main stream, bulk kernel on
                        // the side stream beside it (as a lattice's own split step)
                        hip_check(hipEventRecord(b->fork, st), "event");
                        hip_check(hipStreamWaitEvent(b->stream2, b->fork, 0), "event wait");
                        hip_check(launch_batch_step(bl, par, a, st, b->stream2), "batched stream-collide kernels");
                        hip_check(hipEventRecord(b->join, b->stream2), "event");
                        hip_check(hipStreamWaitEvent(st, b->join, 0), "event wait");
                    }
                }
                for (size_t i = 0; i < n; ++i) {
                    simlab_lattice *l = ml[i];
                    hip_check(launch_reduce_hist(hl[i], (uint32_t)chunk, st), "force reduction kernel");
                    hip_check(hipMemcpyAsync(l->hist_host, l->hist.p, chunk * HIST_STRIDE * 8, hipMemcpyDeviceToHost, st),
                              "force copy");
                }
                hip_check(hipStreamSynchronize(st), "GPU wait failed");
                if (ft_out)
                    for (size_t i = 0; i < n; ++i)
                        for (size_t k = 0; k < chunk; ++k)
                            for (int j = 0; j < 6; ++j)
                                ft_out[(mi[i] * nsteps + done + k) * 6 + j] = ml[i]->hist_host[k * HIST_STRIDE + j];
                done += chunk;
            }
            for (size_t i = 0; i < n; ++i) ml[i]->steps_done += nsteps;
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

// ---------------------------------------------------------------- partial bounce-back + adjoint
namespace {

const char *pbb_unsupported(const simlab_lattice *lat) {
    const simlab_lattice_setup &s = lat->setup;
    if (lat->storage != SIMLAB_STORAGE_FP32 && lat->storage != SIMLAB_STORAGE_FP64)
        return "partial bounce-back needs FP32 or FP64 storage";
    if (lat->nrec != 0 || lat->nhalfway != 0 || lat->has_solid)
        return "partial bounce-back needs a lattice without wall geometry";
    if (s.body_force[0] != 0.0 || s.body_force[1] != 0.0 || s.body_force[2] != 0.0)
        return "partial bounce-back does not take a body force";
    for (int k = 0; k < 6; ++k)
        if (s.sponge[k]) return "partial bounce-back does not take a sponge";
    if (s.wall_model || s.total_pressure_inlet) return "partial bounce-back needs no wall model and no total-pressure inlet";
    return nullptr;
}

}  // namespace

int simlab_cfd_lattice_set_solid_fraction(simlab_lattice *lat, const float *phi) {
    return guarded([&]() -> int {
        if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
        if (const char *m = pbb_unsupported(lat)) return fail(SIMLAB_ERR_INVALID, m);
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            uint64_t mem = 0;
            if (!lat->phi.p) lat->phi.alloc(lat->ncell * 4, mem);
            lat->memory_bytes += mem;
            if (phi) {
                for (uint64_t c = 0; c < lat->ncell; ++c)
                    if (!(phi[c] >= 0.0f && phi[c] <= 1.0f)) return fail(SIMLAB_ERR_INVALID, "solid fraction outside [0, 1]");
                hip_check(hipMemcpy(lat->phi.p, phi, lat->ncell * 4, hipMemcpyHostToDevice), "upload");
            } else {
                hip_check(hipMemset(lat->phi.p, 0, lat->ncell * 4), "memset");
            }
            lat->pbb = true;
            lat->adj_ready = false;
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_adjoint_forward(simlab_lattice *lat, size_t nsteps, size_t ckpt_every, size_t avg_from, int axis,
                               double *J, double *ft_out) {
    return guarded([&]() -> int {
        if (!lat) return fail(SIMLAB_ERR_INVALID, "null lattice");
        if (!lat->pbb) return fail(SIMLAB_ERR_STATE, "set a solid fraction first");
        if (nsteps == 0 || ckpt_every == 0 || avg_from >= nsteps || axis < 0 || axis > 2)
            return fail(SIMLAB_ERR_INVALID, "need nsteps > avg_from >= 0, ckpt_every > 0 and axis 0, 1 or 2");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            DevLattice &L = lat->L;
            hipStream_t st = lat->stream;
            const size_t state = pop_elems(lat->ncell) * elem_bytes(lat->storage);
            const size_t nck = (nsteps + ckpt_every - 1) / ckpt_every;
            const uint32_t nblk = L.xblocks * L.yblocks * L.nz;
            uint64_t mem = 0;
            try {
                if (lat->ckpt.bytes < nck * state) { lat->ckpt.reset(); lat->ckpt.alloc(nck * state, mem); }
                if (lat->pbb_part.bytes < (size_t)KMAX * nblk * 64) { lat->pbb_part.reset(); lat->pbb_part.alloc((size_t)KMAX * nblk * 64, mem); }
            } catch (const HipError &e) {
                if (e.code == SIMLAB_ERR_OOM) return fail(SIMLAB_ERR_OOM, "not enough GPU memory for the adjoint checkpoints");
                throw;
            }
            lat->memory_bytes += mem;
            // from the initial state, as a fresh lattice
            hip_check(launch_init(L, st), "init kernel");
            lat->steps_done = 0;
            StepArgs a0;
            simlab_step_in idn{};
            for (int k = 0; k < 3; ++k) idn.rot[k][k] = idn.rot_prev[k][k] = 1.0;
            pack_step(idn, 0, 0, a0);
            hip_check(launch_faces(L, 0, a0.seed, st), "face kernel");
            hip_check(launch_faces(L, 1, 2654435761u, st), "face kernel");
            double sum = 0.0;
            size_t done = 0;
            while (done < nsteps) {
                const size_t chunk = std::min<size_t>(KMAX, nsteps - done);
                for (size_t k = 0; k < chunk; ++k) {
                    const uint64_t t = done + k;
                    if (t % ckpt_every == 0)
                        hip_check(hipMemcpyAsync((char *)lat->ckpt.p + (t / ckpt_every) * state, L.pops, state,
                                                 hipMemcpyDeviceToDevice, st), "checkpoint copy");
                    StepArgs a;
                    pack_step(idn, t, (uint32_t)k, a);
                    DevLattice Lp = L;
                    Lp.partials = lat->pbb_part.as<double>();
                    hip_check(launch_pbb_step(Lp, (int)(t % 2), a, lat->phi.as<float>(), st), "partial bounce-back step");
                }
                hip_check(launch_reduce_partials(lat->pbb_part.as<double>(), nblk, (uint32_t)chunk, L.hist, st), "force reduction kernel");
                hip_check(hipMemcpyAsync(lat->hist_host, L.hist, chunk * HIST_STRIDE * 8, hipMemcpyDeviceToHost, st), "force copy");
                hip_check(hipStreamSynchronize(st), "GPU wait failed");
                for (size_t k = 0; k < chunk; ++k) {
                    const uint64_t t = done + k;
                    if (t >= avg_from) sum += lat->hist_host[k * HIST_STRIDE + axis];
                    if (ft_out)
                        for (int j = 0; j < 6; ++j) ft_out[t * 6 + j] = lat->hist_host[k * HIST_STRIDE + j];
                }
                done += chunk;
            }
            lat->steps_done = nsteps;
            lat->adj_steps = nsteps;
            lat->adj_every = ckpt_every;
            lat->adj_from = avg_from;
            lat->adj_axis = axis;
            lat->adj_ready = true;
            if (J) *J = sum / (double)(nsteps - avg_from);
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

int simlab_cfd_adjoint_backward(simlab_lattice *lat, double *dJ_dphi) {
    return guarded([&]() -> int {
        if (!lat || !dJ_dphi) return fail(SIMLAB_ERR_INVALID, "null argument");
        if (!lat->adj_ready) return fail(SIMLAB_ERR_STATE, "run adjoint_forward first");
        if (lat->setup.smagorinsky != 0.0) return fail(SIMLAB_ERR_INVALID, "the adjoint needs smagorinsky = 0");
        try {
            hip_check(hipSetDevice(lat->device), "hipSetDevice");
            DevLattice &L = lat->L;
            hipStream_t st = lat->stream;
            const size_t state = pop_elems(lat->ncell) * elem_bytes(lat->storage);
            const uint64_t T = lat->adj_steps, K = lat->adj_every;
            uint64_t mem = 0;
            try {
                if (lat->seg.bytes < (size_t)K * state) { lat->seg.reset(); lat->seg.alloc((size_t)K * state, mem); }
                if (!lat->adj.p) lat->adj.alloc(state, mem);
                if (!lat->grad.p) lat->grad.alloc(lat->ncell * 8, mem);
            } catch (const HipError &e) {
                if (e.code == SIMLAB_ERR_OOM) return fail(SIMLAB_ERR_OOM, "not enough GPU memory for the adjoint segment");
                throw;
            }
            lat->memory_bytes += mem;
            hip_check(hipMemsetAsync(lat->adj.p, 0, state, st), "memset");
            hip_check(hipMemsetAsync(lat->grad.p, 0, lat->ncell * 8, st), "memset");
            DevLattice A = L;
            A.pops = lat->adj.p;
            const double wj = 1.0 / (double)(T - lat->adj_from);
            simlab_step_in idn{};
            for (int k = 0; k < 3; ++k) idn.rot[k][k] = idn.rot_prev[k][k] = 1.0;
            const uint64_t nseg = (T + K - 1) / K;
            for (uint64_t sg = nseg; sg-- > 0;) {
                const uint64_t t0 = sg * K, t1 = std::min<uint64_t>(T, t0 + K);
                // recompute the segment: seg[k] = input state of step t0 + k
                char *S = (char *)lat->seg.p;
                hip_check(hipMemcpyAsync(S, (char *)lat->ckpt.p + sg * state, state, hipMemcpyDeviceToDevice, st), "copy");
                for (uint64_t t = t0; t + 1 < t1; ++t) {
                    char *next = S + (t + 1 - t0) * state;
                    hip_check(hipMemcpyAsync(next, S + (t - t0) * state, state, hipMemcpyDeviceToDevice, st), "copy");
                    DevLattice Ls = L;
                    Ls.pops = next;
                    Ls.partials = lat->pbb_part.as<double>();
                    StepArgs a;
                    pack_step(idn, t, 0, a);
                    hip_check(launch_pbb_step(Ls, (int)(t % 2), a, lat->phi.as<float>(), st), "partial bounce-back step");
                }
                for (uint64_t t = t1; t-- > t0;) {
                    DevLattice Ls = L;
                    Ls.pops = S + (t - t0) * state;
                    hip_check(launch_adjoint_step(Ls, A, (int)(t % 2), lat->phi.as<float>(), lat->grad.as<double>(),
                                                  t >= lat->adj_from ? wj : 0.0, lat->adj_axis, st), "adjoint step");
                }
            }
            hip_check(hipMemcpyAsync(dJ_dphi, lat->grad.p, lat->ncell * 8, hipMemcpyDeviceToHost, st), "GPU read failed");
            hip_check(hipStreamSynchronize(st), "GPU wait failed");
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

}  // extern "C"
