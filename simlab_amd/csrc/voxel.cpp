// Host-side geometry preparation: triangle surface -> lattice solid cells and
// Bouzidi wall links. Restates crates/cfd/src/voxel.rs (the reference runs it
// on the CPU with rayon; here OpenMP). The results are independent of the
// thread count: votes are OR-ed atomically, bins and walls are sorted.
//
//  * inside test (voxel.rs:73-216): ray parity along lines parallel to each
//    axis through every grid point, majority vote over the three axes;
//    consistently oriented lines use the non-zero winding rule, odd lines
//    (holes) do not vote; canonical edge order so a line through a shared
//    edge counts one crossing.
//  * wall links (voxel.rs:363-490): every link from a fluid cell into a
//    solid one; q from the nearest Moller-Trumbore hit (0.5 if none); links
//    crossing a sheet thinner than a cell are walls from both sides.

#include "simlab_internal.h"
#include "voxel_gpu.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace simlab {

namespace {

using P3 = std::array<double, 3>;

// voxel.rs:68-71 scan-line offsets (grid units)
constexpr double EPS_B = 1.234'567'1e-5;
constexpr double EPS_C = 2.718'281'8e-5;

struct Grid {
    P3 origin;
    double h;
    size_t dims[3];
    size_t len() const { return dims[0] * dims[1] * dims[2]; }
    size_t idx(size_t i, size_t j, size_t k) const { return (k * dims[1] + j) * dims[0] + i; }
};

struct Hit {
    uint32_t j;
    double ta;
    int sign;
};

// voxel.rs:186-216
void mark_line(const Grid &g, int a, int b, int c, size_t j, size_t k, const Hit *line, size_t nline, int mode,
               std::atomic<uint32_t> *votes) {
    size_t na = g.dims[a];
    uint32_t vbit = 1u << (3 + a), ibit = 1u << a;
    size_t h = 0;
    int wind = 0;
    for (size_t i = 0; i < na; ++i) {
        double ta = (double)i;
        while (h < nline && line[h].ta < ta) {
            wind += line[h].sign;
            ++h;
        }
        size_t pos[3];
        pos[a] = i;
        pos[b] = j;
        pos[c] = k;
        size_t id = g.idx(pos[0], pos[1], pos[2]);
        uint32_t bits = 0;
        if (mode != 0) {
            bits |= vbit;
            bool ins = mode == 1 ? wind != 0 : (h % 2 == 1);
            if (ins) bits |= ibit;
        }
        if (bits) votes[id].fetch_or(bits, std::memory_order_relaxed);
    }
}

// vertices in grid units (voxel.rs:103-106)
std::vector<P3> grid_verts(const Surface &s, const Grid &g) {
    std::vector<P3> gv(s.verts.size());
    for (size_t v = 0; v < s.verts.size(); ++v)
        for (int k = 0; k < 3; ++k) gv[v][k] = (s.verts[v][k] - g.origin[k]) / g.h;
    return gv;
}

// the triangles whose c-range meets row k of the scan lines along axis a
// (voxel.rs:107-120), in triangle order
std::vector<std::vector<uint32_t>> axis_rows(const Surface &s, const Grid &g, const std::vector<P3> &gv, int a) {
    int c = (a + 2) % 3;
    size_t nc = g.dims[c];
    std::vector<std::vector<uint32_t>> rows(nc);
    for (size_t t = 0; t < s.tris.size(); ++t) {
        double lo = INFINITY, hi = -INFINITY;
        for (int q = 0; q < 3; ++q) {
            lo = std::min(lo, gv[s.tris[t][q]][c]);
            hi = std::max(hi, gv[s.tris[t][q]][c]);
        }
        double k0f = std::max(std::ceil(lo - EPS_C), 0.0);
        double k1f = std::floor(hi - EPS_C);
        if (k1f < 0.0) continue;
        size_t k0 = k0f >= 1.8e19 ? SIZE_MAX : (size_t)k0f;
        size_t k1 = std::min((size_t)k1f, nc ? nc - 1 : 0);
        for (size_t k = k0; k <= k1 && k < nc; ++k) rows[k].push_back((uint32_t)t);
    }
    return rows;
}

// voxel.rs:97-184
void parity_axis(const Surface &s, const Grid &g, int a, std::atomic<uint32_t> *votes) {
    int b = (a + 1) % 3, c = (a + 2) % 3;
    size_t nb = g.dims[b], nc = g.dims[c];
    std::vector<P3> gv = grid_verts(s, g);
    std::vector<std::vector<uint32_t>> rows = axis_rows(s, g, gv, a);
#pragma omp parallel
    {
        std::vector<Hit> hits;
        std::vector<char> have;
#pragma omp for schedule(dynamic, 1)
        for (int64_t kk = 0; kk < (int64_t)nc; ++kk) {
            size_t k = (size_t)kk;
            double pc = (double)k + EPS_C;
            hits.clear();
            for (uint32_t t : rows[k]) {
                const auto &tri = s.tris[t];
                uint32_t ids[3] = {tri[0], tri[1], tri[2]};
                const P3 *p[3] = {&gv[ids[0]], &gv[ids[1]], &gv[ids[2]]};
                double lo = INFINITY, hi = -INFINITY;
                for (int q = 0; q < 3; ++q) {
                    lo = std::min(lo, (*p[q])[b]);
                    hi = std::max(hi, (*p[q])[b]);
                }
                double j0f = std::max(std::ceil(lo - EPS_B), 0.0);
                double j1f = std::floor(hi - EPS_B);
                if (j1f < 0.0) continue;
                size_t j0 = (size_t)j0f;
                size_t j1 = std::min((size_t)j1f, nb ? nb - 1 : 0);
                for (size_t j = j0; j <= j1 && j < nb; ++j) {
                    double pb = (double)j + EPS_B;
                    auto e = [&](int u, int v) -> double {
                        uint32_t iu = ids[u], iv = ids[v];
                        const P3 &pu = iu < iv ? *p[u] : *p[v];
                        const P3 &pv = iu < iv ? *p[v] : *p[u];
                        double sg = iu < iv ? 1.0 : -1.0;
                        return sg * ((pv[b] - pu[b]) * (pc - pu[c]) - (pv[c] - pu[c]) * (pb - pu[b]));
                    };
                    double e0 = e(1, 2), e1 = e(2, 0), e2 = e(0, 1);
                    bool inside = (e0 > 0.0 && e1 > 0.0 && e2 > 0.0) || (e0 < 0.0 && e1 < 0.0 && e2 < 0.0);
                    if (!inside) continue;
                    double sum = e0 + e1 + e2;
                    double ta = (e0 * (*p[0])[a] + e1 * (*p[1])[a] + e2 * (*p[2])[a]) / sum;
                    hits.push_back({(uint32_t)j, ta, sum > 0.0 ? 1 : -1});
                }
            }
            // sort by (line, position); total order on the position
            std::sort(hits.begin(), hits.end(), [](const Hit &x, const Hit &y) {
                if (x.j != y.j) return x.j < y.j;
                return total_cmp_less(x.ta, y.ta);
            });
            size_t i0 = 0;
            while (i0 < hits.size()) {
                uint32_t j = hits[i0].j;
                size_t i1 = i0;
                while (i1 < hits.size() && hits[i1].j == j) ++i1;
                int wsum = 0;
                for (size_t q = i0; q < i1; ++q) wsum += hits[q].sign;
                int mode = wsum == 0 ? 1 : ((i1 - i0) % 2 == 0 ? 2 : 0);
                mark_line(g, a, b, c, j, k, hits.data() + i0, i1 - i0, mode, votes);
                i0 = i1;
            }
            have.assign(nb, 0);
            for (const Hit &h : hits) have[h.j] = 1;
            for (size_t j = 0; j < nb; ++j)
                if (!have[j]) mark_line(g, a, b, c, j, k, nullptr, 0, 1, votes);
        }
    }
}

// voxel.rs:74-95
std::vector<uint8_t> inside_grid(const Surface &s, const Grid &g) {
    size_t n = g.len();
    std::vector<uint8_t> out(n, 0);
    if (n == 0 || s.tris.empty()) return out;
    std::vector<std::atomic<uint32_t>> votes(n);
    for (auto &v : votes) v.store(0, std::memory_order_relaxed);
    for (int a = 0; a < 3; ++a) parity_axis(s, g, a, votes.data());
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < (int64_t)n; ++l) {
        uint32_t v = votes[l].load(std::memory_order_relaxed);
        uint32_t valid = (v >> 3) & 7;
        uint32_t ins = v & 7 & valid;
        int nv = __builtin_popcount(valid), ni = __builtin_popcount(ins);
        out[l] = (nv > 0 && 2 * ni > nv) ? 1 : 0;
    }
    return out;
}

// voxel.rs:218-296 Bins: triangles near each bin of a unit grid
struct Bins {
    P3 lo;
    size_t dims[3];
    std::vector<uint32_t> start, list;

    void build(const Surface &s, P3 lo_, const size_t d[3], double dilate) {
        lo = lo_;
        for (int k = 0; k < 3; ++k) dims[k] = d[k];
        size_t nb = dims[0] * dims[1] * dims[2];
        auto range = [&](size_t t, size_t r0[3], size_t r1[3]) -> bool {
            for (int k = 0; k < 3; ++k) {
                double mn = INFINITY, mx = -INFINITY;
                for (int q = 0; q < 3; ++q) {
                    mn = std::min(mn, s.verts[s.tris[t][q]][k]);
                    mx = std::max(mx, s.verts[s.tris[t][q]][k]);
                }
                mn = mn - dilate - lo[k];
                mx = mx + dilate - lo[k];
                if (mx < 0.0 || mn >= (double)dims[k]) return false;
                r0[k] = (size_t)std::max(std::floor(mn), 0.0);
                r1[k] = std::min((size_t)std::floor(mx), dims[k] - 1);
            }
            return true;
        };
        auto idx = [&](size_t i, size_t j, size_t k) { return (k * dims[1] + j) * dims[0] + i; };
        std::vector<uint32_t> counts(nb, 0);
        for (size_t t = 0; t < s.tris.size(); ++t) {
            size_t r0[3], r1[3];
            if (!range(t, r0, r1)) continue;
            for (size_t k = r0[2]; k <= r1[2]; ++k)
                for (size_t j = r0[1]; j <= r1[1]; ++j)
                    for (size_t i = r0[0]; i <= r1[0]; ++i) counts[idx(i, j, k)] += 1;
        }
        start.assign(nb + 1, 0);
        for (size_t i = 0; i < nb; ++i) start[i + 1] = start[i] + counts[i];
        list.assign(start[nb], 0);
        std::vector<uint32_t> cursor(start.begin(), start.begin() + nb);
        for (size_t t = 0; t < s.tris.size(); ++t) {
            size_t r0[3], r1[3];
            if (!range(t, r0, r1)) continue;
            for (size_t k = r0[2]; k <= r1[2]; ++k)
                for (size_t j = r0[1]; j <= r1[1]; ++j)
                    for (size_t i = r0[0]; i <= r1[0]; ++i) list[cursor[idx(i, j, k)]++] = (uint32_t)t;
        }
        for (size_t i = 0; i < nb; ++i) std::sort(list.begin() + start[i], list.begin() + start[i + 1]);
    }

    std::pair<const uint32_t *, size_t> get(const P3 &p) const {
        size_t ix[3];
        for (int k = 0; k < 3; ++k) {
            double f = std::floor(p[k] - lo[k]);
            if (f < 0.0 || f >= (double)dims[k]) return {nullptr, 0};
            ix[k] = (size_t)f;
        }
        size_t b = (ix[2] * dims[1] + ix[1]) * dims[0] + ix[0];
        return {list.data() + start[b], (size_t)(start[b + 1] - start[b])};
    }
};

}  // namespace

bool total_cmp_less(double x, double y) {
    int64_t a, b;
    std::memcpy(&a, &x, 8);
    std::memcpy(&b, &y, 8);
    a ^= (int64_t)((uint64_t)(a >> 63) >> 1);
    b ^= (int64_t)((uint64_t)(b >> 63) >> 1);
    return a < b;
}

// voxel.rs:298-326 Moller-Trumbore segment test, t in [0, 1] along o + t d
bool seg_tri(const double o[3], const double d[3], const double *t0, const double *t1, const double *t2, double &tt) {
    auto sub = [](const double *a, const double *b, double *r) { for (int k = 0; k < 3; ++k) r[k] = a[k] - b[k]; };
    auto cross = [](const double *a, const double *b, double *r) {
        r[0] = a[1] * b[2] - a[2] * b[1];
        r[1] = a[2] * b[0] - a[0] * b[2];
        r[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    double e1[3], e2[3], pv[3], tv[3], qv[3];
    sub(t1, t0, e1);
    sub(t2, t0, e2);
    cross(d, e2, pv);
    double det = dot(e1, pv);
    if (std::fabs(det) < 1e-14) return false;
    double inv = 1.0 / det;
    sub(o, t0, tv);
    double u = dot(tv, pv) * inv;
    if (!(u >= 0.0 && u <= 1.0)) return false;
    cross(tv, e1, qv);
    double v = dot(d, qv) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    double t = dot(e2, qv) * inv;
    if (!(t >= -1e-9 && t <= 1.0)) return false;
    tt = std::max(t, 0.0);
    return true;
}

std::vector<uint8_t> inside_grid_public(const Surface &s, const double origin[3], double h, const size_t dims[3]) {
    Grid g{{origin[0], origin[1], origin[2]}, h, {dims[0], dims[1], dims[2]}};
    return inside_grid(s, g);
}

namespace {
struct Thin {
    size_t nl;
    int d;
    float q;
};
struct Cell {
    size_t l;
    uint32_t mask;
    float q[27];
    std::vector<Thin> thin;
};

void hcheck(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("GPU voxelizer: ") + what + ": " + hipGetErrorString(e));
}
// device memory owned by the GPU voxelizer for one call
struct DMem {
    void *p = nullptr;
    DMem() = default;
    DMem(const DMem &) = delete;
    DMem &operator=(const DMem &) = delete;
    ~DMem() {
        if (p) (void)hipFree(p);
    }
    template <class T> T *alloc(size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        hcheck(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 16)), "hipMalloc");
        return (T *)p;
    }
    template <class T> T *as() const { return (T *)p; }
};
template <class T> void upload(DMem &d, const std::vector<T> &v, hipStream_t st) {
    d.alloc<T>(v.size());
    if (!v.empty()) hcheck(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st), "upload");
}

// the inside test on the GPU: the same votes as inside_grid. False if a
// scan line crossed more surfaces than a thread keeps (the caller then
// takes the CPU test); d_inside keeps the result on the device.
bool inside_grid_gpu(const Surface &s, const Grid &g, const DMem &d_tris, DMem &d_inside, std::vector<uint8_t> &inside,
                     hipStream_t st) {
    const size_t n = g.len();
    DMem d_votes, d_over, d_gv, d_rs, d_rl;
    uint32_t *votes = d_votes.alloc<uint32_t>(n);
    uint32_t *over = d_over.alloc<uint32_t>(1);
    hcheck(hipMemsetAsync(votes, 0, n * 4, st), "clear");
    hcheck(hipMemsetAsync(over, 0, 4, st), "clear");
    std::vector<P3> gv = grid_verts(s, g);
    std::vector<double> gvf(gv.size() * 3);
    for (size_t v = 0; v < gv.size(); ++v)
        for (int k = 0; k < 3; ++k) gvf[3 * v + k] = gv[v][k];
    upload(d_gv, gvf, st);
    const uint64_t gdims[3] = {g.dims[0], g.dims[1], g.dims[2]};
    for (int a = 0; a < 3; ++a) {
        std::vector<std::vector<uint32_t>> rows = axis_rows(s, g, gv, a);
        std::vector<uint32_t> start(rows.size() + 1, 0), list;
        for (size_t k = 0; k < rows.size(); ++k) {
            start[k + 1] = start[k] + (uint32_t)rows[k].size();
            list.insert(list.end(), rows[k].begin(), rows[k].end());
        }
        upload(d_rs, start, st);
        upload(d_rl, list, st);
        hcheck(gvox_parity(d_gv.as<double>(), d_tris.as<uint32_t>(), d_rs.as<uint32_t>(), d_rl.as<uint32_t>(), a, gdims,
                           votes, over, st),
               "parity kernel");
        hcheck(hipStreamSynchronize(st), "parity kernel");   // the row lists are freed with the next upload
    }
    uint32_t ov = 0;
    hcheck(hipMemcpy(&ov, over, 4, hipMemcpyDeviceToHost), "read");
    if (ov) return false;
    uint8_t *ins = d_inside.alloc<uint8_t>(n);
    hcheck(gvox_majority(votes, n, ins, st), "majority kernel");
    inside.resize(n);
    hcheck(hipMemcpyAsync(inside.data(), ins, n, hipMemcpyDeviceToHost, st), "read");
    hcheck(hipStreamSynchronize(st), "majority kernel");
    return true;
}
}  // namespace

// A lattice velocity set for the link search: nd directions, their
// velocities and opposites, and the rest direction (skipped).
struct LinkSet {
    int nd;
    int rest;
    int c[27][3];
    int opp[27];
};
LinkSet d3q19_links() {
    LinkSet v{19, 0, {}, {}};
    for (int d = 0; d < 19; ++d) {
        for (int k = 0; k < 3; ++k) v.c[d][k] = CV[d][k];
        v.opp[d] = opp(d);
    }
    return v;
}
// d3q27.rs:5-15: direction i = (i % 3 - 1, (i / 3) % 3 - 1, i / 9 - 1), rest 13, opposite 26 - i
LinkSet d3q27_links() {
    LinkSet v{27, 13, {}, {}};
    for (int d = 0; d < 27; ++d) {
        v.c[d][0] = d % 3 - 1;
        v.c[d][1] = (d / 3) % 3 - 1;
        v.c[d][2] = d / 9 - 1;
        v.opp[d] = 26 - d;
    }
    return v;
}

// voxel.rs:348-359 + 363-490 (voxelize_links for the D3Q19 or D3Q27 set).
// gpu (D3Q19 only): the inside test and the link search on the GPU (the
// same decisions; voxel_gpu.hip)
StaticVoxels voxelize_static_impl(const Surface &s, const size_t dims[3], bool gpu, const LinkSet &ls) {
    StaticVoxels out;
    const int nd = ls.nd, rest = ls.rest;
    double maxlen = 0.0;
    for (int d = 0; d < nd; ++d)
        maxlen = std::max(maxlen, std::sqrt((double)(ls.c[d][0] * ls.c[d][0] + ls.c[d][1] * ls.c[d][1] + ls.c[d][2] * ls.c[d][2])));
    if (nd != 19) gpu = false;
    P3 blo = {INFINITY, INFINITY, INFINITY}, bhi = {-INFINITY, -INFINITY, -INFINITY};
    for (const auto &p : s.verts)
        for (int k = 0; k < 3; ++k) {
            blo[k] = std::min(blo[k], p[k]);
            bhi[k] = std::max(bhi[k], p[k]);
        }
    int64_t slo[3], shi[3];
    for (int k = 0; k < 3; ++k) {
        slo[k] = std::max((int64_t)std::floor(blo[k] - 2.0), (int64_t)-1);
        shi[k] = std::min((int64_t)std::ceil(bhi[k] + 2.0), (int64_t)dims[k]);
    }
    for (int k = 0; k < 3; ++k) {
        out.lo[k] = (size_t)std::max(slo[k], (int64_t)0);
        out.hi[k] = (size_t)std::max(std::min(shi[k], (int64_t)dims[k] - 1), (int64_t)0);
    }
    if (s.tris.empty()) return out;
    for (int k = 0; k < 3; ++k)
        if (slo[k] > shi[k]) return out;
    size_t gd[3];
    for (int k = 0; k < 3; ++k) gd[k] = (size_t)(shi[k] - slo[k] + 1);
    Grid g{{(double)slo[0], (double)slo[1], (double)slo[2]}, 1.0, {gd[0], gd[1], gd[2]}};
    std::vector<uint8_t> inside;
    hipStream_t st = nullptr;
    DMem d_tris, d_inside, d_verts;
    bool on_gpu = false;
    if (gpu) {
        std::vector<uint32_t> tf(s.tris.size() * 3);
        std::vector<double> vf(s.verts.size() * 3);
        for (size_t t = 0; t < s.tris.size(); ++t)
            for (int k = 0; k < 3; ++k) tf[3 * t + k] = s.tris[t][k];
        for (size_t v = 0; v < s.verts.size(); ++v)
            for (int k = 0; k < 3; ++k) vf[3 * v + k] = s.verts[v][k];
        upload(d_tris, tf, st);
        upload(d_verts, vf, st);
        on_gpu = inside_grid_gpu(s, g, d_tris, d_inside, inside, st);
    }
    if (!on_gpu) inside = inside_grid(s, g);
    Bins bins;
    bins.build(s, {(double)slo[0] - 0.5, (double)slo[1] - 0.5, (double)slo[2] - 0.5}, gd, maxlen + 0.01);
    auto in_domain = [&](size_t i, size_t j, size_t k) {
        int64_t p[3] = {(int64_t)i + slo[0], (int64_t)j + slo[1], (int64_t)k + slo[2]};
        for (int a = 0; a < 3; ++a)
            if (p[a] < 0 || p[a] >= (int64_t)dims[a]) return false;
        return true;
    };
    auto gidx = [&](size_t i, size_t j, size_t k) -> size_t {
        return ((size_t)((int64_t)k + slo[2]) * dims[1] + (size_t)((int64_t)j + slo[1])) * dims[0] +
               (size_t)((int64_t)i + slo[0]);
    };
    auto rel = [&](size_t l, size_t &i, size_t &j, size_t &k) {
        i = l % gd[0];
        j = (l / gd[0]) % gd[1];
        k = l / (gd[0] * gd[1]);
    };
    size_t glen = g.len();
    for (size_t l = 0; l < glen; ++l) {
        if (!inside[l]) continue;
        size_t i, j, k;
        rel(l, i, j, k);
        if (in_domain(i, j, k)) out.solid.push_back((uint32_t)gidx(i, j, k));
    }
    auto is_solid = [&](int64_t i, int64_t j, int64_t k) -> bool {
        if (i < 0 || j < 0 || k < 0 || i >= (int64_t)gd[0] || j >= (int64_t)gd[1] || k >= (int64_t)gd[2]) return false;
        return inside[g.idx((size_t)i, (size_t)j, (size_t)k)] != 0;
    };
    std::vector<Cell> cells;
    if (on_gpu) {
        // every fluid point with a wall link, in any order; sorted by point below
        DMem d_bs, d_bl, d_out, d_cnt;
        upload(d_bs, bins.start, st);
        upload(d_bl, bins.list, st);
        GvoxLinkArgs a{};
        a.verts = d_verts.as<double>();
        a.tris = d_tris.as<uint32_t>();
        a.inside = d_inside.as<uint8_t>();
        a.bin_start = d_bs.as<uint32_t>();
        a.bin_list = d_bl.as<uint32_t>();
        for (int k = 0; k < 3; ++k) {
            a.bin_lo[k] = bins.lo[k];
            a.bin_dims[k] = bins.dims[k];
            a.gd[k] = gd[k];
            a.slo[k] = slo[k];
            a.dims[k] = dims[k];
        }
        a.count = d_cnt.alloc<unsigned long long>(1);
        a.capacity = std::min<uint64_t>(glen, (uint64_t)1 << 20);
        unsigned long long found = 0;
        for (int pass = 0; pass < 2; ++pass) {
            a.out = d_out.alloc<GvoxCell>(a.capacity);
            hcheck(hipMemsetAsync(a.count, 0, 8, st), "clear");
            hcheck(gvox_links(a, st), "link kernel");
            hcheck(hipMemcpyAsync(&found, a.count, 8, hipMemcpyDeviceToHost, st), "read");
            hcheck(hipStreamSynchronize(st), "link kernel");
            if (found <= a.capacity) break;
            a.capacity = found;   // room for all of them: once more
        }
        std::vector<GvoxCell> gc(found);
        if (found) hcheck(hipMemcpy(gc.data(), a.out, found * sizeof(GvoxCell), hipMemcpyDeviceToHost), "read");
        std::sort(gc.begin(), gc.end(), [](const GvoxCell &x, const GvoxCell &y) { return x.l < y.l; });
        cells.resize(found);
        for (size_t r = 0; r < found; ++r) {
            Cell &c = cells[r];
            c.l = gc[r].l;
            c.mask = gc[r].mask;
            for (int d = 0; d < 27; ++d) c.q[d] = d < 19 ? gc[r].q[d] : 1.0f;
            size_t i, j, k;
            rel(c.l, i, j, k);
            for (int d = 1; d < 19; ++d)
                if (gc[r].thin_q[d] >= 0.0f) {
                    const size_t ni = (size_t)((int64_t)i + CV[d][0]), nj = (size_t)((int64_t)j + CV[d][1]),
                                 nk = (size_t)((int64_t)k + CV[d][2]);
                    c.thin.push_back({g.idx(ni, nj, nk), opp(d), gc[r].thin_q[d]});
                }
        }
    }
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    std::vector<std::vector<Cell>> parts(on_gpu ? 0 : nthreads);
    // contiguous static chunks per thread keep the concatenation in l order
#pragma omp parallel num_threads(nthreads) if (!on_gpu)
    if (!on_gpu) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        size_t chunk = (glen + nthreads - 1) / nthreads;
        size_t l0 = std::min(glen, (size_t)tid * chunk), l1 = std::min(glen, l0 + chunk);
        for (size_t l = l0; l < l1; ++l) {
            if (inside[l]) continue;
            size_t i, j, k;
            rel(l, i, j, k);
            if (!in_domain(i, j, k)) continue;
            P3 p = {(double)((int64_t)i + slo[0]), (double)((int64_t)j + slo[1]), (double)((int64_t)k + slo[2])};
            auto cand = bins.get(p);
            Cell cell;
            cell.l = l;
            cell.mask = 0;
            for (float &q : cell.q) q = 1.0f;
            for (int d = 0; d < nd; ++d) {
                if (d == rest) continue;
                const int *c = ls.c[d];
                int64_t ni = (int64_t)i + c[0], nj = (int64_t)j + c[1], nk = (int64_t)k + c[2];
                double dv[3] = {(double)c[0], (double)c[1], (double)c[2]};
                double tmin = INFINITY, tmax = -INFINITY;
                int hits = 0;
                for (size_t q = 0; q < cand.second; ++q) {
                    uint32_t t = cand.first[q];
                    double tt;
                    if (seg_tri(p.data(), dv, s.verts[s.tris[t][0]].data(), s.verts[s.tris[t][1]].data(),
                                s.verts[s.tris[t][2]].data(), tt)) {
                        tmin = std::min(tmin, tt);
                        tmax = std::max(tmax, tt);
                        hits += 1;
                    }
                }
                bool in_region = ni >= 0 && nj >= 0 && nk >= 0 && ni < (int64_t)gd[0] && nj < (int64_t)gd[1] &&
                                 nk < (int64_t)gd[2];
                bool nsolid = in_region ? is_solid(ni, nj, nk) : (hits % 2 == 1);
                if (nsolid) {
                    cell.q[d] = hits > 0 ? (float)std::min(std::max(tmin, 1e-3), 1.0) : 0.5f;
                    cell.mask |= 1u << d;
                } else if (hits >= 2 && tmax - tmin > 1e-6) {
                    cell.q[d] = (float)std::min(std::max(tmin, 1e-3), 1.0);
                    cell.mask |= 1u << d;
                    if (in_region)
                        cell.thin.push_back({g.idx((size_t)ni, (size_t)nj, (size_t)nk), ls.opp[d],
                                             (float)std::min(std::max(1.0 - tmax, 1e-3), 1.0)});
                }
            }
            if (cell.mask) parts[tid].push_back(std::move(cell));
        }
    }
    for (auto &pt : parts)
        for (auto &c : pt) cells.push_back(std::move(c));
    // make thin-sheet links reciprocal (voxel.rs:462-480)
    std::unordered_map<size_t, size_t> at;
    for (size_t r = 0; r < cells.size(); ++r) at[cells[r].l] = r;
    size_t nper = cells.size();
    for (size_t r = 0; r < nper; ++r) {
        for (size_t t = 0; t < cells[r].thin.size(); ++t) {
            Thin th = cells[r].thin[t];
            size_t i, j, k;
            rel(th.nl, i, j, k);
            if (inside[th.nl] || !in_domain(i, j, k)) continue;
            auto it = at.find(th.nl);
            size_t rr;
            if (it == at.end()) {
                Cell nc;
                nc.l = th.nl;
                nc.mask = 0;
                for (float &q : nc.q) q = 1.0f;
                cells.push_back(std::move(nc));
                rr = cells.size() - 1;
                at[th.nl] = rr;
            } else {
                rr = it->second;
            }
            if (!(cells[rr].mask & (1u << th.d))) {
                cells[rr].mask |= 1u << th.d;
                cells[rr].q[th.d] = th.q;
            }
        }
    }
    std::sort(cells.begin(), cells.end(), [](const Cell &x, const Cell &y) { return x.l < y.l; });
    out.walls.reserve(cells.size());
    for (const Cell &c : cells) {
        size_t i, j, k;
        rel(c.l, i, j, k);
        simlab_wall_cell w;
        w.cell = (uint32_t)gidx(i, j, k);
        w.mask = c.mask;
        std::memcpy(w.q, c.q, sizeof w.q);
        out.walls.push_back(w);
    }
    return out;
}

StaticVoxels voxelize_static(const Surface &s, const size_t dims[3]) {
    return voxelize_static_impl(s, dims, false, d3q19_links());
}
StaticVoxels voxelize_static_gpu(const Surface &s, const size_t dims[3]) {
    return voxelize_static_impl(s, dims, true, d3q19_links());
}
StaticVoxels voxelize_static27(const Surface &s, const size_t dims[3]) {
    return voxelize_static_impl(s, dims, false, d3q27_links());
}

// voxel.rs:562-584: split triangles 1 -> 4 until no edge exceeds max_edge
Surface subdivide(const Surface &s, double max_edge) {
    Surface r;
    r.verts = s.verts;
    std::vector<std::array<uint32_t, 3>> stack = s.tris;
    auto len = [](const P3 &a, const P3 &b) {
        return std::sqrt(std::pow(a[0] - b[0], 2) + std::pow(a[1] - b[1], 2) + std::pow(a[2] - b[2], 2));
    };
    while (!stack.empty()) {
        auto t = stack.back();
        stack.pop_back();
        P3 a = r.verts[t[0]], b = r.verts[t[1]], c = r.verts[t[2]];
        if (std::max(std::max(len(a, b), len(b, c)), len(c, a)) <= max_edge || r.tris.size() > 50'000'000) {
            r.tris.push_back(t);
            continue;
        }
        auto mid = [&](const P3 &p, const P3 &q) {
            r.verts.push_back({(p[0] + q[0]) / 2.0, (p[1] + q[1]) / 2.0, (p[2] + q[2]) / 2.0});
            return (uint32_t)(r.verts.size() - 1);
        };
        uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
        stack.push_back({t[0], ab, ca});
        stack.push_back({ab, t[1], bc});
        stack.push_back({ca, bc, t[2]});
        stack.push_back({ab, bc, ca});
    }
    return r;
}

// voxel.rs:507-560: body-frame occupancy bits + subdivided triangles
SpinVoxels voxelize_spin(const Surface &s, const double axis[3], double h) {
    SpinVoxels v;
    P3 blo = {INFINITY, INFINITY, INFINITY}, bhi = {-INFINITY, -INFINITY, -INFINITY};
    for (const auto &p : s.verts)
        for (int k = 0; k < 3; ++k) {
            blo[k] = std::min(blo[k], p[k]);
            bhi[k] = std::max(bhi[k], p[k]);
        }
    double m = 2.0 * h;
    for (int k = 0; k < 3; ++k) {
        v.occ_lo[k] = blo[k] - m;
        v.occ_dims[k] = (size_t)std::ceil((bhi[k] - blo[k] + 2.0 * m) / h) + 1;
    }
    v.occ_h = h;
    Grid g{{v.occ_lo[0], v.occ_lo[1], v.occ_lo[2]}, h, {v.occ_dims[0], v.occ_dims[1], v.occ_dims[2]}};
    std::vector<uint8_t> inside = inside_grid(s, g);
    v.occ_bits.assign((g.len() + 31) / 32, 0);
    for (size_t l = 0; l < inside.size(); ++l)
        if (inside[l]) v.occ_bits[l >> 5] |= 1u << (l & 31);
    Surface fine = subdivide(s, 1.5);
    v.tris.reserve(fine.tris.size() * 12);
    for (const auto &t : fine.tris) {
        const P3 &a = fine.verts[t[0]], &b = fine.verts[t[1]], &c = fine.verts[t[2]];
        float row[12] = {(float)a[0], (float)a[1], (float)a[2], 0.0f,
                         (float)(b[0] - a[0]), (float)(b[1] - a[1]), (float)(b[2] - a[2]), 0.0f,
                         (float)(c[0] - a[0]), (float)(c[1] - a[1]), (float)(c[2] - a[2]), 0.0f};
        v.tris.insert(v.tris.end(), row, row + 12);
    }
    v.radius = 0.0;
    v.axial_min = INFINITY;
    v.axial_max = -INFINITY;
    for (const auto &p : s.verts) {
        double a = p[0] * axis[0] + p[1] * axis[1] + p[2] * axis[2];
        double r2 = std::max(p[0] * p[0] + p[1] * p[1] + p[2] * p[2] - a * a, 0.0);
        v.radius = std::max(v.radius, std::sqrt(r2));
        v.axial_min = std::min(v.axial_min, a);
        v.axial_max = std::max(v.axial_max, a);
    }
    return v;
}

}  // namespace simlab

// ---------------------------------------------------------------- C ABI
struct simlab_static_voxels {
    simlab::StaticVoxels v;
};
struct simlab_spin_voxels {
    simlab::SpinVoxels v;
};

namespace {
bool make_surface(const double *verts, uint64_t nverts, const uint32_t *tris, uint64_t ntris, simlab::Surface &s) {
    s.verts.resize(nverts);
    for (uint64_t i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) s.verts[i][k] = verts[3 * i + k];
    s.tris.resize(ntris);
    for (uint64_t t = 0; t < ntris; ++t)
        for (int k = 0; k < 3; ++k) {
            if (tris[3 * t + k] >= nverts) return false;
            s.tris[t][k] = tris[3 * t + k];
        }
    return true;
}
}  // namespace

extern "C" {

int simlab_cfd_voxelize_static(const double *verts, uint64_t nverts, const uint32_t *tris, uint64_t ntris,
                               const uint64_t dims[3], simlab_static_voxels **out) {
    return simlab::guarded([&]() -> int {
        if (!out || !dims || (nverts && !verts) || (ntris && !tris)) return simlab::fail(SIMLAB_ERR_INVALID, "null argument");
        simlab::Surface s;
        if (!make_surface(verts, nverts, tris, ntris, s)) return simlab::fail(SIMLAB_ERR_INVALID, "triangle index out of range");
        size_t d[3] = {(size_t)dims[0], (size_t)dims[1], (size_t)dims[2]};
        auto *v = new simlab_static_voxels;
        v->v = simlab::voxelize_static(s, d);
        *out = v;
        return SIMLAB_OK;
    });
}

int simlab_cfd_voxelize_static27(const double *verts, uint64_t nverts, const uint32_t *tris, uint64_t ntris,
                                 const uint64_t dims[3], simlab_static_voxels **out) {
    return simlab::guarded([&]() -> int {
        if (!out || !dims || (nverts && !verts) || (ntris && !tris)) return simlab::fail(SIMLAB_ERR_INVALID, "null argument");
        simlab::Surface s;
        if (!make_surface(verts, nverts, tris, ntris, s)) return simlab::fail(SIMLAB_ERR_INVALID, "triangle index out of range");
        size_t d[3] = {(size_t)dims[0], (size_t)dims[1], (size_t)dims[2]};
        auto *v = new simlab_static_voxels;
        v->v = simlab::voxelize_static27(s, d);
        *out = v;
        return SIMLAB_OK;
    });
}

int simlab_cfd_voxelize_static_gpu(const double *verts, uint64_t nverts, const uint32_t *tris, uint64_t ntris,
                                   const uint64_t dims[3], simlab_static_voxels **out) {
    return simlab::guarded([&]() -> int {
        if (!out || !dims || (nverts && !verts) || (ntris && !tris)) return simlab::fail(SIMLAB_ERR_INVALID, "null argument");
        simlab::Surface s;
        if (!make_surface(verts, nverts, tris, ntris, s)) return simlab::fail(SIMLAB_ERR_INVALID, "triangle index out of range");
        size_t d[3] = {(size_t)dims[0], (size_t)dims[1], (size_t)dims[2]};
        auto *v = new simlab_static_voxels;
        try {
            v->v = simlab::voxelize_static_gpu(s, d);
        } catch (...) {
            delete v;
            throw;
        }
        *out = v;
        return SIMLAB_OK;
    });
}

int simlab_cfd_voxels_view(const simlab_static_voxels *v, uint64_t *nsolid, const uint32_t **solid, uint64_t *nwall,
                           const simlab_wall_cell **walls, uint64_t lo[3], uint64_t hi[3]) {
    if (!v) return simlab::fail(SIMLAB_ERR_INVALID, "null voxels");
    if (nsolid) *nsolid = v->v.solid.size();
    if (solid) *solid = v->v.solid.data();
    if (nwall) *nwall = v->v.walls.size();
    if (walls) *walls = v->v.walls.data();
    for (int k = 0; k < 3; ++k) {
        if (lo) lo[k] = v->v.lo[k];
        if (hi) hi[k] = v->v.hi[k];
    }
    return SIMLAB_OK;
}

void simlab_cfd_voxels_free(simlab_static_voxels *v) { delete v; }

int simlab_cfd_inside_grid(const double *verts, uint64_t nverts, const uint32_t *tris, uint64_t ntris,
                           const double origin[3], double h, const uint64_t dims[3], uint8_t *out) {
    return simlab::guarded([&]() -> int {
        if (!out || !dims || !origin || (nverts && !verts) || (ntris && !tris)) return simlab::fail(SIMLAB_ERR_INVALID, "null argument");
        simlab::Surface s;
        if (!make_surface(verts, nverts, tris, ntris, s)) return simlab::fail(SIMLAB_ERR_INVALID, "triangle index out of range");
        size_t d[3] = {(size_t)dims[0], (size_t)dims[1], (size_t)dims[2]};
        auto r = simlab::inside_grid_public(s, origin, h, d);
        std::memcpy(out, r.data(), r.size());
        return SIMLAB_OK;
    });
}

int simlab_cfd_voxelize_spin(const double *verts, uint64_t nverts, const uint32_t *tris, uint64_t ntris,
                             const double axis[3], double h, simlab_spin_voxels **out) {
    return simlab::guarded([&]() -> int {
        if (!out || !axis || (nverts && !verts) || (ntris && !tris)) return simlab::fail(SIMLAB_ERR_INVALID, "null argument");
        if (!(h > 0.0)) return simlab::fail(SIMLAB_ERR_INVALID, "occupancy spacing must be positive");
        simlab::Surface s;
        if (!make_surface(verts, nverts, tris, ntris, s)) return simlab::fail(SIMLAB_ERR_INVALID, "triangle index out of range");
        auto *v = new simlab_spin_voxels;
        v->v = simlab::voxelize_spin(s, axis, h);
        *out = v;
        return SIMLAB_OK;
    });
}

int simlab_cfd_spin_voxels_view(const simlab_spin_voxels *v, double occ_lo[3], double *occ_h, uint64_t occ_dims[3],
                                uint64_t *nocc_words, const uint32_t **occ_bits, uint64_t *ntri, const float **tris,
                                double *radius, double axial[2]) {
    if (!v) return simlab::fail(SIMLAB_ERR_INVALID, "null voxels");
    for (int k = 0; k < 3; ++k) {
        if (occ_lo) occ_lo[k] = v->v.occ_lo[k];
        if (occ_dims) occ_dims[k] = v->v.occ_dims[k];
    }
    if (occ_h) *occ_h = v->v.occ_h;
    if (nocc_words) *nocc_words = v->v.occ_bits.size();
    if (occ_bits) *occ_bits = v->v.occ_bits.data();
    if (ntri) *ntri = v->v.tris.size() / 12;
    if (tris) *tris = v->v.tris.data();
    if (radius) *radius = v->v.radius;
    if (axial) {
        axial[0] = v->v.axial_min;
        axial[1] = v->v.axial_max;
    }
    return SIMLAB_OK;
}

void simlab_cfd_spin_voxels_free(simlab_spin_voxels *v) { delete v; }

}  // extern "C"
