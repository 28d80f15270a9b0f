// gfx950 kernels of the static voxelizer (crates/cfd/src/voxel.rs:74-216,
// 298-326, 363-460 run on the GPU). Every decision is the CPU voxelizer's
// (voxel.cpp): the same double-precision operations in the same order, no
// fused multiply-adds, so the inside votes, the wall links and the q values
// come out bit for bit the same.

#include "voxel_gpu.h"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace simlab {
namespace {

#pragma clang fp contract(off)

// voxel.rs:68-71 scan-line offsets (grid units)
constexpr double EPS_B = 1.234'567'1e-5;
constexpr double EPS_C = 2.718'281'8e-5;

// D3Q19 velocity set (shader.rs:22-42 order)
__constant__ int GCV[19][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                               {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1}, {0, -1, -1}, {1, -1, 0},
                               {-1, 1, 0}, {1, 0, -1},  {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};

// total order of doubles (f64::total_cmp): x < y
__device__ __forceinline__ bool total_less(double x, double y) {
    long long a = __double_as_longlong(x), b = __double_as_longlong(y);
    a ^= (long long)((unsigned long long)(a >> 63) >> 1);
    b ^= (long long)((unsigned long long)(b >> 63) >> 1);
    return a < b;
}

struct LHit {
    double ta;
    int sign;
};

// voxel.rs:97-216 for one scan line: the crossings of the line (j, k) along
// axis a with the triangles of row k, sorted along the line, then the vote
// of every point of the line (wind or parity, mode as voxel.rs:170-180)
__global__ void __launch_bounds__(64) k_parity(const double *gv, const uint3 *tris, const uint32_t *row_start,
                                               const uint32_t *row_list, int a, uint64_t d0, uint64_t d1, uint64_t d2,
                                               uint32_t *votes, uint32_t *overflow) {
    const uint64_t dims[3] = {d0, d1, d2};
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const uint64_t nb = dims[b], nc = dims[c], na = dims[a];
    const uint64_t line = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nb * nc) return;
    const uint64_t j = line % nb, k = line / nb;
    const double pc = (double)k + EPS_C;
    const double pb = (double)j + EPS_B;
    LHit hits[GVOX_MAX_LINE_HITS];
    int nh = 0;
    bool over = false;
    for (uint32_t r = row_start[k]; r < row_start[k + 1]; ++r) {
        const uint3 tv = tris[row_list[r]];
        const uint32_t ids[3] = {tv.x, tv.y, tv.z};
        const double *p[3] = {gv + 3 * (size_t)ids[0], gv + 3 * (size_t)ids[1], gv + 3 * (size_t)ids[2]};
        double lo = INFINITY, hi = -INFINITY;
        for (int q = 0; q < 3; ++q) {
            lo = p[q][b] < lo ? p[q][b] : lo;
            hi = hi < p[q][b] ? p[q][b] : hi;
        }
        double j0f = ceil(lo - EPS_B);
        j0f = j0f < 0.0 ? 0.0 : j0f;
        const double j1f = floor(hi - EPS_B);
        if (j1f < 0.0) continue;
        if ((double)j < j0f || (double)j > j1f) continue;
        auto e = [&](int u, int v) -> double {
            const uint32_t iu = ids[u], iv = ids[v];
            const double *pu = iu < iv ? p[u] : p[v];
            const double *pv = iu < iv ? p[v] : p[u];
            const double sg = iu < iv ? 1.0 : -1.0;
            return sg * ((pv[b] - pu[b]) * (pc - pu[c]) - (pv[c] - pu[c]) * (pb - pu[b]));
        };
        const double e0 = e(1, 2), e1 = e(2, 0), e2 = e(0, 1);
        const bool inside = (e0 > 0.0 && e1 > 0.0 && e2 > 0.0) || (e0 < 0.0 && e1 < 0.0 && e2 < 0.0);
        if (!inside) continue;
        const double sum = e0 + e1 + e2;
        const double ta = (e0 * p[0][a] + e1 * p[1][a] + e2 * p[2][a]) / sum;
        if (nh == GVOX_MAX_LINE_HITS) {
            over = true;
            continue;
        }
        // insertion in total order of the position along the line
        int q = nh++;
        while (q > 0 && total_less(ta, hits[q - 1].ta)) {
            hits[q] = hits[q - 1];
            --q;
        }
        hits[q].ta = ta;
        hits[q].sign = sum > 0.0 ? 1 : -1;
    }
    if (over) atomicOr(overflow, 1u);
    int wsum = 0;
    for (int q = 0; q < nh; ++q) wsum += hits[q].sign;
    const int mode = wsum == 0 ? 1 : (nh % 2 == 0 ? 2 : 0);
    const uint32_t vbit = 1u << (3 + a), ibit = 1u << a;
    int h = 0, wind = 0;
    for (uint64_t i = 0; i < na; ++i) {
        const double ta = (double)i;
        while (h < nh && hits[h].ta < ta) {
            wind += hits[h].sign;
            ++h;
        }
        uint64_t pos[3];
        pos[a] = i;
        pos[b] = j;
        pos[c] = k;
        const uint64_t id = (pos[2] * dims[1] + pos[1]) * dims[0] + pos[0];
        uint32_t bits = 0;
        if (mode != 0) {
            bits |= vbit;
            const bool ins = mode == 1 ? wind != 0 : (h % 2 == 1);
            if (ins) bits |= ibit;
        }
        if (bits) atomicOr(votes + id, bits);
    }
}

__global__ void __launch_bounds__(256) k_majority(const uint32_t *votes, uint64_t n, uint8_t *inside) {
    const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const uint32_t v = votes[l];
    const uint32_t valid = (v >> 3) & 7u;
    const uint32_t ins = v & 7u & valid;
    const int nv = __popc(valid), ni = __popc(ins);
    inside[l] = (nv > 0 && 2 * ni > nv) ? 1 : 0;
}

// voxel.rs:298-326 Moller-Trumbore segment test, t in [0, 1] along o + t d
__device__ bool seg_tri(const double o[3], const double d[3], const double *t0, const double *t1, const double *t2,
                        double &tt) {
    double e1[3], e2[3], pv[3], tv[3], qv[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = t1[k] - t0[k];
        e2[k] = t2[k] - t0[k];
    }
    pv[0] = d[1] * e2[2] - d[2] * e2[1];
    pv[1] = d[2] * e2[0] - d[0] * e2[2];
    pv[2] = d[0] * e2[1] - d[1] * e2[0];
    const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (fabs(det) < 1e-14) return false;
    const double inv = 1.0 / det;
    for (int k = 0; k < 3; ++k) tv[k] = o[k] - t0[k];
    const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (!(u >= 0.0 && u <= 1.0)) return false;
    qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
    qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
    qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
    const double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    const double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    if (!(t >= -1e-9 && t <= 1.0)) return false;
    tt = t < 0.0 ? 0.0 : t;   // std::max(t, 0.0)
    return true;
}

// min(max(x, 1e-3), 1) as std::min / std::max evaluate it
__device__ __forceinline__ double clamp_q(double x) {
    const double m = x < 1e-3 ? 1e-3 : x;
    return 1.0 < m ? 1.0 : m;
}

// voxel.rs:404-460: the links of one fluid grid point
__global__ void __launch_bounds__(128) k_links(GvoxLinkArgs A) {
    const uint64_t glen = A.gd[0] * A.gd[1] * A.gd[2];
    const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= glen || A.inside[l]) return;
    const int64_t i = (int64_t)(l % A.gd[0]), j = (int64_t)((l / A.gd[0]) % A.gd[1]), k = (int64_t)(l / (A.gd[0] * A.gd[1]));
    const int64_t pg[3] = {i + A.slo[0], j + A.slo[1], k + A.slo[2]};
    for (int a = 0; a < 3; ++a)
        if (pg[a] < 0 || pg[a] >= (int64_t)A.dims[a]) return;
    const double p[3] = {(double)pg[0], (double)pg[1], (double)pg[2]};
    // the triangles binned near p (Bins::get; none outside the bins): with
    // none, a link into a solid neighbour is still a wall (q = 1/2)
    uint32_t c0 = 0, c1 = 0;
    {
        uint64_t ix[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const double f = floor(p[a] - A.bin_lo[a]);
            if (f < 0.0 || f >= (double)A.bin_dims[a]) ok = false;
            else ix[a] = (uint64_t)f;
        }
        if (ok) {
            const uint64_t bn = (ix[2] * A.bin_dims[1] + ix[1]) * A.bin_dims[0] + ix[0];
            c0 = A.bin_start[bn];
            c1 = A.bin_start[bn + 1];
        }
    }
    GvoxCell cell;
    cell.l = l;
    cell.mask = 0;
    cell.pad = 0;
    for (int d = 0; d < 19; ++d) {
        cell.q[d] = 1.0f;
        cell.thin_q[d] = -1.0f;
    }
    for (int d = 1; d < 19; ++d) {
        const int64_t ni = i + GCV[d][0], nj = j + GCV[d][1], nk = k + GCV[d][2];
        const double dv[3] = {(double)GCV[d][0], (double)GCV[d][1], (double)GCV[d][2]};
        double tmin = INFINITY, tmax = -INFINITY;
        int hits = 0;
        for (uint32_t q = c0; q < c1; ++q) {
            const uint32_t t = A.bin_list[q];
            double tt;
            if (seg_tri(p, dv, A.verts + 3 * (size_t)A.tris[3 * (size_t)t], A.verts + 3 * (size_t)A.tris[3 * (size_t)t + 1],
                        A.verts + 3 * (size_t)A.tris[3 * (size_t)t + 2], tt)) {
                tmin = tt < tmin ? tt : tmin;
                tmax = tmax < tt ? tt : tmax;
                hits += 1;
            }
        }
        const bool in_region = ni >= 0 && nj >= 0 && nk >= 0 && ni < (int64_t)A.gd[0] && nj < (int64_t)A.gd[1] &&
                               nk < (int64_t)A.gd[2];
        const bool nsolid = in_region ? A.inside[((uint64_t)nk * A.gd[1] + (uint64_t)nj) * A.gd[0] + (uint64_t)ni] != 0
                                      : (hits % 2 == 1);
        if (nsolid) {
            cell.q[d] = hits > 0 ? (float)clamp_q(tmin) : 0.5f;
            cell.mask |= 1u << d;
        } else if (hits >= 2 && tmax - tmin > 1e-6) {
            cell.q[d] = (float)clamp_q(tmin);
            cell.mask |= 1u << d;
            if (in_region) cell.thin_q[d] = (float)clamp_q(1.0 - tmax);
        }
    }
    if (!cell.mask) return;
    const unsigned long long slot = atomicAdd(A.count, 1ull);
    if (slot < A.capacity) A.out[slot] = cell;
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

hipError_t gvox_parity(const double *gv, const uint32_t *tris, const uint32_t *row_start, const uint32_t *row_list,
                       int a, const uint64_t gdims[3], uint32_t *votes, uint32_t *overflow, hipStream_t s) {
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const uint64_t lines = gdims[b] * gdims[c];
    if (lines == 0) return hipSuccess;
    k_parity<<<cdiv(lines, 64), 64, 0, s>>>(gv, (const uint3 *)tris, row_start, row_list, a, gdims[0], gdims[1], gdims[2],
                                            votes, overflow);
    return hipGetLastError();
}

hipError_t gvox_majority(const uint32_t *votes, uint64_t n, uint8_t *inside, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_majority<<<cdiv(n, 256), 256, 0, s>>>(votes, n, inside);
    return hipGetLastError();
}

hipError_t gvox_links(const GvoxLinkArgs &a, hipStream_t s) {
    const uint64_t glen = a.gd[0] * a.gd[1] * a.gd[2];
    if (glen == 0) return hipSuccess;
    k_links<<<cdiv(glen, 128), 128, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace simlab
