// Host side of the refined D3Q27 lattice: setup validation, per-level cell
// types and lists, wall records, slice pixels and moment buffers
// (crates/cfd/src/mlattice.rs:185-690), the recursive dispatch schedule of
// one coarse step (mlattice.rs:786-842), the run loop with its force
// read-back (mlattice.rs:713-784) and the read-backs (mlattice.rs:844-934),
// behind the C ABI of include/simlab_cfd.h (`simlab_cfd_multi_*`).

#include "hip_util.h"
#include "mkernels.h"
#include "simlab_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace simlab;
using namespace simlab::m27;

namespace {

size_t elem_bytes27(int storage) { return storage == SIMLAB_STORAGE_FP32 ? 4 : storage == SIMLAB_STORAGE_FP64 ? 8 : 2; }
size_t real_bytes27(int storage) { return storage == SIMLAB_STORAGE_FP64 ? 8 : 4; }

uint32_t level_dim(const simlab_multi_setup &s, int l, int k) {
    return l == 0 ? s.dims0[k] : 2 * (s.boxes[l - 1].n[k] - 1) + 1;
}

struct Level {
    DevLevel d{};
    uint32_t dims[3] = {0, 0, 0};
    uint64_t n = 0;
    double tau = 0.0;
    DevBuf pops, flags, lists, mom_child, mom_up, spix, dump;
    uint32_t src = 0;     // the set the next step reads
    uint32_t mslot = 1;   // the moments slot the latest step wrote
    uint64_t steps = 0;
};

// mlattice.rs:165-181: the 24-word record of a wall cell (F and T zero)
void rec_words(uint32_t cell, uint32_t mask, const float q[27], uint32_t *w) {
    for (uint32_t k = 0; k < RS; ++k) w[k] = 0;
    w[0] = mask;
    w[1] = cell;
    float qs[26];
    for (float &v : qs) v = 1.0f;
    for (int i = 0; i < 27; ++i)
        if (i != (int)REST) qs[i < (int)REST ? i : i - 1] = q[i];
    for (int k = 0; k < 13; ++k) w[R_Q + k] = (uint32_t)f16_bits_host(qs[2 * k]) | ((uint32_t)f16_bits_host(qs[2 * k + 1]) << 16);
}

}  // namespace

struct simlab_multi {
    int device = 0;
    simlab_multi_setup setup{};
    int storage = 0;
    std::vector<std::unique_ptr<Level>> lv;
    hipStream_t stream = nullptr;
    uint32_t nrec = 0, nvert = 0, npix = 0, nblk = 0;
    double vertex_offset = 0.0;
    uint64_t memory_bytes = 0;
    uint64_t fine_steps = 0;
    std::vector<uint32_t> rec_host;    // nrec x RS words (F, T filled at read-back)
    std::vector<float> probe_host;     // nvert x 8
    DevBuf rmask, rq, rft, probes, sacc, partials, hist;
    double *hist_host = nullptr;       // pinned, HIST_CAP x 8
    ~simlab_multi() {
        lv.clear();
        if (hist_host) (void)hipHostFree(hist_host);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int new_multi(const simlab_multi_setup *sp, const simlab_geometry *g, const simlab_probes *pr, simlab_multi **out) {
    if (!sp || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
    *out = nullptr;
    const simlab_multi_setup &s = *sp;
    if (s.nboxes > SIMLAB_MULTI_MAX_LEVELS - 1) return fail(SIMLAB_ERR_INVALID, "too many refinement levels");
    if (s.storage != SIMLAB_STORAGE_FP32 && s.storage != SIMLAB_STORAGE_FP16 && s.storage != SIMLAB_STORAGE_FP64)
        return fail(SIMLAB_ERR_INVALID, "the refined solver stores FP32, FP16 or FP64 populations");
    for (int k = 0; k < 3; ++k)
        if (s.dims0[k] < 3) return fail(SIMLAB_ERR_INVALID, "grid too small");
    const int nl = (int)s.nboxes + 1, finest = nl - 1;
    // mlattice.rs:192-209
    const bool periodic0[3] = {s.faces[0] == SIMLAB_FACE_PERIODIC, s.faces[2] == SIMLAB_FACE_PERIODIC,
                               s.faces[4] == SIMLAB_FACE_PERIODIC};
    if (nl > 1 && (periodic0[0] || periodic0[1] || periodic0[2]))
        return fail(SIMLAB_ERR_INVALID, "periodic faces only work without refinement");
    for (int l = 1; l < nl; ++l) {
        const simlab_level_box &b = s.boxes[l - 1];
        for (int k = 0; k < 3; ++k) {
            const int64_t hi = (int64_t)b.lo[k] + (int64_t)b.n[k] - 1;
            const int margin = l == 1 ? 2 : 3;
            if (b.lo[k] < margin || hi > (int64_t)level_dim(s, l - 1, k) - 1 - margin || b.n[k] < 8)
                return fail(SIMLAB_ERR_INVALID, "refinement box " + std::to_string(l) + " does not fit inside its parent");
        }
    }
    std::vector<uint64_t> ncell(nl);
    for (int l = 0; l < nl; ++l) {
        ncell[l] = (uint64_t)level_dim(s, l, 0) * level_dim(s, l, 1) * level_dim(s, l, 2);
        if (ncell[l] >= 0xffffffffull) return fail(SIMLAB_ERR_INVALID, "level " + std::to_string(l) + " has too many cells");
    }
    if (g && g->kind == SIMLAB_GEOM_SPIN)
        return fail(SIMLAB_ERR_INVALID, "the refined solver takes still parts only (spinning parts use the D3Q19 lattice)");
    const bool has_geom = g && g->kind == SIMLAB_GEOM_STATIC;
    if (has_geom && ((g->nwall && !g->walls) || (g->nsolid && !g->solid))) return fail(SIMLAB_ERR_INVALID, "null geometry array");
    if (pr && ((pr->nvert && !pr->verts) || (pr->nslice && !pr->slices))) return fail(SIMLAB_ERR_INVALID, "null probe array");

    auto M = std::make_unique<simlab_multi>();
    hip_check(hipGetDevice(&M->device), "hipGetDevice");
    M->setup = s;
    M->storage = s.storage;
    const uint64_t nfine = ncell[finest];
    // wall records (finest level)
    std::vector<std::pair<uint32_t, uint32_t>> rec_of;
    std::vector<uint32_t> solid;
    if (has_geom) {
        if (g->nwall >= (1ull << 28)) return fail(SIMLAB_ERR_INVALID, "too many wall cells");
        M->nrec = (uint32_t)g->nwall;
        M->rec_host.assign((size_t)M->nrec * RS, 0);
        for (uint64_t r = 0; r < g->nwall; ++r) {
            if (g->walls[r].cell >= nfine) return fail(SIMLAB_ERR_INVALID, "wall cell outside the finest level");
            rec_words(g->walls[r].cell, g->walls[r].mask, g->walls[r].q, &M->rec_host[r * RS]);
            rec_of.push_back({g->walls[r].cell, (uint32_t)r + 1});
        }
        solid.assign(g->solid, g->solid + g->nsolid);
    }
    std::sort(rec_of.begin(), rec_of.end());
    std::sort(solid.begin(), solid.end());
    // probes (mlattice.rs:237-240, 367-374)
    if (pr && pr->nvert) {
        M->nvert = (uint32_t)pr->nvert;
        M->probe_host.assign((size_t)M->nvert * 8, 0.0f);
        for (uint64_t v = 0; v < pr->nvert; ++v)
            for (int k = 0; k < 6; ++k) M->probe_host[8 * v + k] = (float)pr->verts[6 * v + k];
    }
    M->vertex_offset = pr ? pr->vertex_offset : 0.0;
    // each level's node 0 and cell size in finest-level nodes (mlattice.rs:246-264)
    std::vector<std::array<double, 3>> org(nl, {0.0, 0.0, 0.0});
    std::vector<double> scale(nl, 1.0);
    scale[0] = (double)((uint64_t)1 << finest);
    for (int l = 1; l < nl; ++l) {
        scale[l] = scale[l - 1] / 2.0;
        for (int k = 0; k < 3; ++k) org[l][k] = org[l - 1][k] + (double)s.boxes[l - 1].lo[k] * scale[l - 1];
    }
    // slice pixels: each from the finest level whose interior (without its
    // ghost layer) holds it (mlattice.rs:265-303)
    std::vector<std::vector<uint32_t>> pix(nl);
    uint32_t pbase = 0;
    for (uint64_t si = 0; pr && si < pr->nslice; ++si) {
        const simlab_slice_def &sd = pr->slices[si];
        for (uint32_t iv = 0; iv < sd.nv; ++iv)
            for (uint32_t iu = 0; iu < sd.nu; ++iu) {
                double p[3];
                for (int k = 0; k < 3; ++k) p[k] = (double)(sd.origin[k] + (int32_t)iu * sd.u[k] + (int32_t)iv * sd.v[k]);
                for (int l = nl - 1; l >= 0; --l) {
                    bool ok = true;
                    uint32_t cell[3] = {0, 0, 0};
                    for (int k = 0; k < 3; ++k) {
                        const double x = std::round((p[k] - org[l][k]) / scale[l]);
                        const double lo = l == 0 ? 1.0 : 2.0;
                        if (x < lo || x > (double)level_dim(s, l, k) - 1.0 - lo) {
                            ok = false;
                            break;
                        }
                        cell[k] = (uint32_t)x;
                    }
                    if (ok) {
                        pix[l].push_back(pbase + iv * sd.nu + iu);
                        pix[l].push_back((cell[2] * level_dim(s, l, 1) + cell[1]) * level_dim(s, l, 0) + cell[0]);
                        break;
                    }
                }
            }
        pbase += sd.nu * sd.nv;
    }
    M->npix = pbase;

    hip_check(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking), "hipStreamCreate");
    size_t total = 0;
    const size_t eb = elem_bytes27(s.storage), rb = real_bytes27(s.storage);
    try {
        M->rmask.alloc((size_t)M->nrec * 4, total);
        M->rq.alloc((size_t)M->nrec * 13 * 4, total);
        M->rft.alloc((size_t)M->nrec * 6 * rb, total);
        M->nblk = std::max<uint32_t>(1, (M->nrec + RED_CHUNK - 1) / RED_CHUNK);
        M->partials.alloc((size_t)HIST_CAP * M->nblk * 6 * 8, total);
        M->hist.alloc((size_t)HIST_CAP * 8 * 8, total);
        M->probes.alloc((size_t)M->nvert * 8 * 4, total);
        M->sacc.alloc((size_t)M->npix * 5 * 4, total);
        for (int l = 0; l < nl; ++l) {
            M->lv.push_back(std::make_unique<Level>());
            Level &L = *M->lv.back();
            for (int k = 0; k < 3; ++k) L.dims[k] = level_dim(s, l, k);
            L.n = ncell[l];
            L.pops.alloc(2 * set_elems(L.n) * eb, total);
            L.flags.alloc(L.n * 4, total);
        }
    } catch (const HipError &e) {
        char b[256];
        std::snprintf(b, sizeof b, "not enough GPU memory (%.1f GB): %s", (double)total / 1e9, e.what());
        return fail(SIMLAB_ERR_OOM, b);
    }
    hipStream_t st = M->stream;
    // records, SoA on the device
    if (M->nrec) {
        std::vector<uint32_t> mask(M->nrec), q((size_t)13 * M->nrec);
        for (uint32_t r = 0; r < M->nrec; ++r) {
            mask[r] = M->rec_host[(size_t)r * RS];
            for (int k = 0; k < 13; ++k) q[(size_t)k * M->nrec + r] = M->rec_host[(size_t)r * RS + R_Q + k];
        }
        hip_check(hipMemcpy(M->rmask.p, mask.data(), mask.size() * 4, hipMemcpyHostToDevice), "upload");
        hip_check(hipMemcpy(M->rq.p, q.data(), q.size() * 4, hipMemcpyHostToDevice), "upload");
        hip_check(hipMemset(M->rft.p, 0, M->rft.bytes), "clear");
    }
    if (M->nvert) hip_check(hipMemcpy(M->probes.p, M->probe_host.data(), M->probe_host.size() * 4, hipMemcpyHostToDevice), "upload");
    if (M->npix) hip_check(hipMemset(M->sacc.p, 0, M->sacc.bytes), "clear");

    // per level: cell types and lists (mlattice.rs:391-470), moment buffers
    // (mlattice.rs:475-500), constants (mlattice.rs:510-545)
    try {
        for (int l = 0; l < nl; ++l) {
            Level &L = *M->lv[l];
            const uint32_t nx = L.dims[0], ny = L.dims[1], nz = L.dims[2];
            const bool has_child = l + 1 < nl;
            int32_t ch_lo[3] = {0, 0, 0};
            uint32_t ch_n[3] = {1, 1, 1};
            if (has_child)
                for (int k = 0; k < 3; ++k) {
                    ch_lo[k] = s.boxes[l].lo[k];
                    ch_n[k] = s.boxes[l].n[k];
                }
            std::vector<uint32_t> flags(L.n, 0), faces, ghosts, restricted;
            for (uint64_t cc = 0; cc < L.n; ++cc) {
                const uint32_t x = (uint32_t)(cc % nx), y = (uint32_t)((cc / nx) % ny), z = (uint32_t)(cc / ((uint64_t)nx * ny));
                uint32_t ty = 0;
                if (l == 0) {
                    const bool on[6] = {x == 0, x == nx - 1, y == 0, y == ny - 1, z == 0, z == nz - 1};
                    for (int f = 0; f < 6; ++f)
                        if (on[f] && s.faces[f] != SIMLAB_FACE_PERIODIC) {
                            const uint32_t t = s.faces[f] == SIMLAB_FACE_VELOCITY ? 1u : 2u;
                            ty = ty == 1u ? 1u : t;   // a velocity face wins at edges
                        }
                }
                if (ty) {
                    flags[cc] = T_FACE | (ty << 3);
                    faces.push_back((uint32_t)cc);
                } else if (l > 0 && (x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1)) {
                    flags[cc] = T_GHOST;
                    ghosts.push_back((uint32_t)cc);
                } else if (has_child) {
                    const int32_t p[3] = {(int32_t)x, (int32_t)y, (int32_t)z};
                    bool in = true;
                    for (int k = 0; k < 3; ++k)
                        if (p[k] < ch_lo[k] + 2 || p[k] > ch_lo[k] + (int32_t)ch_n[k] - 3) in = false;
                    if (in) {
                        flags[cc] = T_RESTRICT;
                        restricted.push_back((uint32_t)cc);
                    }
                }
            }
            if (l == finest) {
                for (uint32_t c : solid)
                    if (c < L.n && flags[c] == 0) flags[c] = T_SOLID;
                for (const auto &rf : rec_of)
                    if ((flags[rf.first] & 7u) == T_FLUID || (flags[rf.first] & 7u) == T_SOLID)
                        flags[rf.first] = (flags[rf.first] & 7u) | (rf.second << 3);
            }
            std::vector<uint32_t> lists = faces;
            lists.insert(lists.end(), ghosts.begin(), ghosts.end());
            lists.insert(lists.end(), restricted.begin(), restricted.end());
            L.lists.alloc(lists.size() * 4, total);
            L.spix.alloc(pix[l].size() * 4, total);
            hip_check(hipMemcpy(L.flags.p, flags.data(), flags.size() * 4, hipMemcpyHostToDevice), "upload");
            if (!lists.empty()) hip_check(hipMemcpy(L.lists.p, lists.data(), lists.size() * 4, hipMemcpyHostToDevice), "upload");
            if (!pix[l].empty()) hip_check(hipMemcpy(L.spix.p, pix[l].data(), pix[l].size() * 4, hipMemcpyHostToDevice), "upload");
            // moments: rest at the initial velocity, this level's tau
            L.tau = 0.5 + 3.0 * s.nu0 * (double)((uint64_t)1 << l);
            const uint64_t ch_nodes = has_child ? (uint64_t)ch_n[0] * ch_n[1] * ch_n[2] : 0;
            const uint32_t pn[3] = {(nx - 1) / 2 + 1, (ny - 1) / 2 + 1, (nz - 1) / 2 + 1};
            const uint64_t pnn = l > 0 ? (uint64_t)pn[0] * pn[1] * pn[2] : 0;
            L.mom_child.alloc((size_t)2 * ch_nodes * MW * rb, total);
            L.mom_up.alloc((size_t)pnn * MW * rb, total);
            auto init_m = [&](uint64_t nodes, DevBuf &buf) {
                if (!nodes) return;
                if (rb == 8) {
                    std::vector<double> v(nodes * MW, 0.0);
                    for (uint64_t k = 0; k < nodes; ++k) {
                        v[k * MW] = 1.0;
                        for (int a = 0; a < 3; ++a) v[k * MW + 1 + a] = s.u_init[a];
                        v[k * MW + 10] = L.tau;
                    }
                    hip_check(hipMemcpy(buf.p, v.data(), v.size() * 8, hipMemcpyHostToDevice), "upload");
                } else {
                    std::vector<float> v(nodes * MW, 0.0f);
                    for (uint64_t k = 0; k < nodes; ++k) {
                        v[k * MW] = 1.0f;
                        for (int a = 0; a < 3; ++a) v[k * MW + 1 + a] = (float)s.u_init[a];
                        v[k * MW + 10] = (float)L.tau;
                    }
                    hip_check(hipMemcpy(buf.p, v.data(), v.size() * 4, hipMemcpyHostToDevice), "upload");
                }
            };
            init_m(2 * ch_nodes, L.mom_child);
            init_m(pnn, L.mom_up);
            if (s.debug_dump) L.dump.alloc(L.n * 16, total);
            DevLevel &d = L.d;
            d.nx = nx;
            d.ny = ny;
            d.nz = nz;
            d.n = (uint32_t)L.n;
            d.storage = s.storage;
            for (int k = 0; k < 3; ++k) d.per[k] = l == 0 && periodic0[k] ? 1u : 0u;
            d.pops = L.pops.p;
            d.flags = L.flags.as<uint32_t>();
            d.lists = L.lists.as<uint32_t>();
            d.nface = (uint32_t)faces.size();
            d.nghost = (uint32_t)ghosts.size();
            d.nrestrict = (uint32_t)restricted.size();
            d.tau0 = L.tau;
            d.smag2 = s.smagorinsky * s.smagorinsky;
            d.sponge_tau = l == 0 ? s.sponge_tau : 0.0;
            d.sigma = l == 0 ? s.sponge_sigma : 0.0;
            const double scale_l = (double)((uint64_t)1 << (finest - l));
            d.n_sponge = 0;
            d.forcing = 0;
            for (int k = 0; k < 3; ++k) {
                d.sp_lo[k] = l == 0 ? (double)s.sponge[2 * k] : 0.0;
                d.sp_hi[k] = l == 0 ? (double)s.sponge[2 * k + 1] : 0.0;
                d.u_inf[k] = s.u_inf[k];
                d.u_init[k] = s.u_init[k];
                d.grav[k] = s.body_force0[k] / (double)((uint64_t)1 << l);
                d.center[k] = l == finest ? s.center[k] : (s.center[k] - org[l][k]) / scale_l;
                if (d.sp_lo[k] > 0.0 || d.sp_hi[k] > 0.0) d.n_sponge = 1;
                if (d.grav[k] != 0.0) d.forcing = 1;
            }
            d.absorb = s.absorb_velocity ? 1 : 0;
            d.total_p = s.total_pressure_inlet ? 1 : 0;
            d.has_child = has_child ? 1 : 0;
            d.has_parent = l > 0 ? 1 : 0;
            for (int k = 0; k < 3; ++k) {
                d.ch_lo[k] = ch_lo[k];
                d.ch_n[k] = ch_n[k];
                d.pn[k] = pn[k];
            }
            d.ch_nodes = (uint32_t)ch_nodes;
            d.pnn = (uint32_t)pnn;
            d.mom_child = L.mom_child.p;
            d.mom_up = L.mom_up.p;
            d.nrec = l == finest ? M->nrec : 0;
            d.rmask = l == finest ? M->rmask.as<uint32_t>() : nullptr;
            d.rq = l == finest ? M->rq.as<uint32_t>() : nullptr;
            d.rft = l == finest ? M->rft.p : nullptr;
            d.npix = (uint32_t)(pix[l].size() / 2);
            d.spix = L.spix.as<uint32_t>();
            d.sacc = M->sacc.as<float>();
            d.dump = s.debug_dump ? L.dump.as<float4>() : nullptr;
        }
    } catch (const HipError &e) {
        if (e.code != SIMLAB_ERR_OOM) throw;
        char b[256];
        std::snprintf(b, sizeof b, "not enough GPU memory (%.1f GB): %s", (double)total / 1e9, e.what());
        return fail(SIMLAB_ERR_OOM, b);
    }
    for (int l = 0; l < nl; ++l) {
        DevLevel &d = M->lv[l]->d;
        d.mom_parent = l > 0 ? M->lv[l - 1]->mom_child.p : nullptr;
        d.mom_child_up = l + 1 < nl ? M->lv[l + 1]->mom_up.p : nullptr;
    }
    hip_check(hipHostMalloc((void **)&M->hist_host, (size_t)HIST_CAP * 8 * 8, hipHostMallocDefault), "hipHostMalloc");
    M->memory_bytes = total;
    // initial state (mlattice.rs:675-688): feq(1, u_init) in both sets, seed 0
    for (auto &L : M->lv) hip_check(m_init(L->d, st), "init kernel");
    hip_check(hipStreamSynchronize(st), "init kernel");
    *out = M.release();
    return SIMLAB_OK;
}

// mlattice.rs:690-705: the seed of one dispatch
uint32_t seed_of(const simlab_multi *M, uint32_t hslot, uint32_t src, uint32_t mw) {
    return ((uint32_t)M->fine_steps + hslot) * 2654435761u + (src * 7919u + mw * 104729u);
}

StepArgs args_of(const simlab_step_in &sin) {
    StepArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.omega[k] = sin.omega[k];
        for (int j = 0; j < 3; ++j) a.rot[k][j] = sin.rot[k][j];
    }
    return a;
}

// mlattice.rs:786-842: one step of level l and, recursively, the two steps
// of each finer level inside it, with the ghost fills before and the
// restriction after them
void schedule(simlab_multi *M, int l, const simlab_step_in *fine, uint64_t &fine_i) {
    const int nl = (int)M->lv.size(), finest = nl - 1;
    Level &L = *M->lv[l];
    hipStream_t st = M->stream;
    simlab_step_in still{};
    for (int k = 0; k < 3; ++k) still.rot[k][k] = still.rot_prev[k][k] = 1.0;
    const simlab_step_in &sin = l == finest ? fine[fine_i] : still;
    const uint32_t lsrc = L.src, lmw = 1 - L.mslot;
    const uint32_t hslot = (uint32_t)(fine_i % HIST_CAP);
    StepArgs base = args_of(sin);
    base.src = lsrc;
    base.mw = lmw;
    base.seed = seed_of(M, hslot, lsrc, lmw);
    if (l == 0 && L.d.nface > 0) hip_check(m_face(L.d, base, st), "face kernel");
    hip_check(m_step(L.d, base, st), "step kernel");
    L.src = 1 - lsrc;
    L.mslot = lmw;
    L.steps += 1;
    if (l == finest) {
        if (M->nrec > 0)
            hip_check(m_reduce1(L.d, M->partials.as<double>() + (size_t)fine_i * M->nblk * 6, st), "reduce kernel");
        if (sin.sample) {
            if (M->nvert > 0) {
                StepArgs after = base;
                after.src = 1 - lsrc;
                hip_check(m_sample_verts(L.d, after, M->probes.as<float>(), M->nvert, M->vertex_offset, st), "sampler");
            }
            for (auto &lk : M->lv)
                if (lk->d.npix > 0) hip_check(m_sample_slices(lk->d, lk->src, st), "sampler");
        }
        fine_i += 1;
    }
    if (l + 1 < nl) {
        const uint32_t mold = 1 - lmw;
        Level &C = *M->lv[l + 1];
        for (int sub = 0; sub < 2; ++sub) {
            StepArgs e = args_of(still);
            e.src = C.src;
            e.mold = mold;
            e.mnew = lmw;
            e.blend = 0.5f * (float)sub;
            e.seed = seed_of(M, 0, C.src, 0);
            if (C.d.nghost > 0) hip_check(m_ghost(C.d, e, st), "ghost kernel");
            schedule(M, l + 1, fine, fine_i);
        }
        StepArgs e = args_of(still);
        e.src = L.src;
        e.seed = seed_of(M, 0, L.src, 0);
        if (L.d.nrestrict > 0) hip_check(m_restrict(L.d, e, st), "restrict kernel");
    }
}

int run_multi(simlab_multi *M, const simlab_step_in *steps, size_t nsteps, double *ft) {
    if (!M || (nsteps && !steps)) return fail(SIMLAB_ERR_INVALID, "null argument");
    const uint64_t ss = (uint64_t)1 << (M->lv.size() - 1);
    if (nsteps % ss != 0) return fail(SIMLAB_ERR_INVALID, "step count must be a multiple of " + std::to_string(ss));
    hip_check(hipSetDevice(M->device), "hipSetDevice");
    const uint64_t max_coarse = std::max<uint64_t>(HIST_CAP / ss, 1);
    uint64_t done = 0;
    while (done < nsteps) {
        const uint64_t nc = std::min<uint64_t>((nsteps - done) / ss, max_coarse);
        uint64_t fine_i = 0;
        for (uint64_t k = 0; k < nc; ++k) schedule(M, 0, steps + done, fine_i);
        const uint32_t nf = (uint32_t)fine_i;
        if (M->nrec > 0) {
            hip_check(m_reduce2(M->partials.as<double>(), M->nblk, nf, M->hist.as<double>(), M->stream), "reduce kernel");
            hip_check(hipMemcpyAsync(M->hist_host, M->hist.p, (size_t)nf * 64, hipMemcpyDeviceToHost, M->stream), "read-back");
        }
        hip_check(hipStreamSynchronize(M->stream), "run");
        if (ft)
            for (uint32_t k = 0; k < nf; ++k)
                for (int j = 0; j < 6; ++j) ft[(done + k) * 6 + j] = M->nrec > 0 ? M->hist_host[(size_t)k * 8 + j] : 0.0;
        M->fine_steps += nf;
        done += nf;
    }
    return SIMLAB_OK;
}

template <class F> int guarded_hip(F &&f) {
    return guarded([&]() -> int {
        try {
            return f();
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}

}  // namespace

extern "C" {

int simlab_cfd_multi_new(const simlab_multi_setup *setup, const simlab_geometry *geom, const simlab_probes *probes,
                         simlab_multi **out) {
    return guarded_hip([&] { return new_multi(setup, geom, probes, out); });
}

void simlab_cfd_multi_free(simlab_multi *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    delete m;
}

int simlab_cfd_multi_info(const simlab_multi *m, simlab_multi_info *out) {
    if (!m || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof *out);
    out->levels = (uint32_t)m->lv.size();
    out->sub_steps = 1u << (m->lv.size() - 1);
    out->fine_steps = m->fine_steps;
    out->memory_bytes = m->memory_bytes;
    out->nrec = m->nrec;
    out->nvert = m->nvert;
    out->npix = m->npix;
    out->storage = m->storage;
    for (size_t l = 0; l < m->lv.size(); ++l) {
        for (int k = 0; k < 3; ++k) out->dims[l][k] = m->lv[l]->dims[k];
        out->cells[l] = m->lv[l]->n;
        out->tau[l] = m->lv[l]->tau;
    }
    return SIMLAB_OK;
}

int simlab_cfd_multi_run(simlab_multi *m, const simlab_step_in *steps, size_t nsteps, double *ft_out) {
    return guarded_hip([&] { return run_multi(m, steps, nsteps, ft_out); });
}

int simlab_cfd_multi_run_still(simlab_multi *m, size_t nsteps, double *ft_out) {
    return guarded_hip([&] {
        simlab_step_in s{};
        for (int k = 0; k < 3; ++k) s.rot[k][k] = s.rot_prev[k][k] = 1.0;
        std::vector<simlab_step_in> v(nsteps, s);
        return run_multi(m, v.data(), nsteps, ft_out);
    });
}

int simlab_cfd_multi_macros(simlab_multi *m, uint32_t level, float *out) {
    return guarded_hip([&]() -> int {
        if (!m || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
        if (level >= m->lv.size()) return fail(SIMLAB_ERR_INVALID, "no such level");
        if (!m->setup.debug_dump) return fail(SIMLAB_ERR_STATE, "macros() needs debug_dump");
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        Level &L = *m->lv[level];
        hip_check(m_dump(L.d, L.src, m->stream), "dump kernel");
        hip_check(hipMemcpyAsync(out, L.dump.p, L.n * 16, hipMemcpyDeviceToHost, m->stream), "read-back");
        hip_check(hipStreamSynchronize(m->stream), "read-back");
        return SIMLAB_OK;
    });
}

int simlab_cfd_multi_records(simlab_multi *m, uint32_t *out) {
    return guarded_hip([&]() -> int {
        if (!m || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        std::memset(out, 0, (size_t)std::max<uint32_t>(m->nrec, 1) * RS * 4);
        if (!m->nrec) return SIMLAB_OK;
        const size_t rb = real_bytes27(m->storage);
        std::vector<uint8_t> ft((size_t)m->nrec * 6 * rb);
        hip_check(hipStreamSynchronize(m->stream), "read-back");
        hip_check(hipMemcpy(ft.data(), m->rft.p, ft.size(), hipMemcpyDeviceToHost), "read-back");
        for (uint32_t r = 0; r < m->nrec; ++r) {
            uint32_t *w = out + (size_t)r * RS;
            std::memcpy(w, &m->rec_host[(size_t)r * RS], RS * 4);
            for (int k = 0; k < 6; ++k) {
                const size_t i = (size_t)k * m->nrec + r;
                const float v = rb == 8 ? (float)((const double *)ft.data())[i] : ((const float *)ft.data())[i];
                std::memcpy(&w[R_FT + k], &v, 4);
            }
        }
        return SIMLAB_OK;
    });
}

int simlab_cfd_multi_reset_samples(simlab_multi *m) {
    return guarded_hip([&]() -> int {
        if (!m) return fail(SIMLAB_ERR_INVALID, "null argument");
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(m->stream), "reset");
        if (m->nvert) hip_check(hipMemcpy(m->probes.p, m->probe_host.data(), m->probe_host.size() * 4, hipMemcpyHostToDevice), "reset");
        if (m->npix) hip_check(hipMemset(m->sacc.p, 0, m->sacc.bytes), "reset");
        return SIMLAB_OK;
    });
}

int simlab_cfd_multi_vertex_means(simlab_multi *m, float *out) {
    return guarded_hip([&]() -> int {
        if (!m || (m->nvert && !out)) return fail(SIMLAB_ERR_INVALID, "null argument");
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(m->stream), "read-back");
        std::vector<float> v((size_t)m->nvert * 8);
        if (m->nvert) hip_check(hipMemcpy(v.data(), m->probes.p, v.size() * 4, hipMemcpyDeviceToHost), "read-back");
        for (uint32_t i = 0; i < m->nvert; ++i) out[i] = v[8 * i + 7] > 0.0f ? v[8 * i + 6] / v[8 * i + 7] : NAN;
        return SIMLAB_OK;
    });
}

int simlab_cfd_multi_slice_means(simlab_multi *m, float *out) {
    return guarded_hip([&]() -> int {
        if (!m || (m->npix && !out)) return fail(SIMLAB_ERR_INVALID, "null argument");
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(m->stream), "read-back");
        std::vector<float> v((size_t)m->npix * 5);
        if (m->npix) hip_check(hipMemcpy(v.data(), m->sacc.p, v.size() * 4, hipMemcpyDeviceToHost), "read-back");
        for (uint32_t p = 0; p < m->npix; ++p) {
            const float c = v[5 * p + 4];
            for (int k = 0; k < 4; ++k) out[4 * p + k] = c > 0.0f ? v[5 * p + k] / c : NAN;
        }
        return SIMLAB_OK;
    });
}

int simlab_cfd_multi_get_pops(simlab_multi *m, uint32_t level, double *out) {
    return guarded_hip([&]() -> int {
        if (!m || !out) return fail(SIMLAB_ERR_INVALID, "null argument");
        if (level >= m->lv.size()) return fail(SIMLAB_ERR_INVALID, "no such level");
        hip_check(hipSetDevice(m->device), "hipSetDevice");
        Level &L = *m->lv[level];
        DevBuf tmp;
        size_t t = 0;
        tmp.alloc((size_t)Q * L.n * 8, t);
        hip_check(m_get_pops(L.d, L.src, tmp.as<double>(), m->stream), "read kernel");
        hip_check(hipMemcpyAsync(out, tmp.p, (size_t)Q * L.n * 8, hipMemcpyDeviceToHost, m->stream), "read-back");
        hip_check(hipStreamSynchronize(m->stream), "read-back");
        return SIMLAB_OK;
    });
}

}  // extern "C"
