// ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this file; only tests/, the smoke check and the CPU-baseline leg of
// bench.py do, and only as the checker.
//
// A plain CPU restatement of the reference's uniform D3Q19 lattice
// (offgridevices/simlab crates/cfd): the WGSL kernels generated by
// crates/cfd/src/shader.rs and the host bookkeeping of
// crates/cfd/src/lattice.rs, one loop iteration per GPU thread, in the same
// order of operations as the WGSL text. Templated on the arithmetic type:
// float for the FP32/FP16/BF16 storages (WGSL f32), double for the FP64
// oracle run. OpenMP over host cores; it is also the CPU baseline.
//
// Pinned by (tests/test_oracle_*.py): the reference's own tests restated
// (crates/cfd/tests/channel.rs, momentum.rs), closed forms (tau = 1 gives
// feq, halfway bounce-back, Poiseuille), invariants (mass, zero force on a
// uniform state), the committed golden vectors, and brute force on tiny
// lattices; the absorbing layer and the samplers by the closed forms of
// tests/test_oracle_layers_probes.py; the Spalding wall model by
// tests/test_oracle_wall_model.py; the refill's mean stress by
// tests/test_oracle_spin.py.

#include "../include/simlab_cfd.h"
#include "half.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using namespace oracle_half;

// shader.rs:22-42 velocity set, shader.rs:44-52 opp, shader.rs:54-60 weights
constexpr int CV[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                           {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                           {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                           {1, 0, -1}, {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};
inline int opp(int i) { return i == 0 ? 0 : (i % 2 == 1 ? i + 1 : i - 1); }
inline double weight_d(int i) { return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

// Oracle storages: the four public ones plus FP16 with the reference's exact
// per-store hash (shader.rs:510-517) for the known-answer checks.
enum Store { ST_F32 = 0, ST_F16 = 1, ST_BF16 = 2, ST_F64 = 3, ST_F16REF = 4 };

constexpr float FS = 32768.0f;          // shader.rs:73 FP16_SCALE
constexpr float INV_S = 1.0f / 32768.0f;
constexpr uint32_t RS = 40;             // shader.rs:63

// This library's dither stream for FP16/BF16 stores: one hash per cell (the
// reference's mixing function applied to the cell index), then xorshift32
// words; direction i takes 16 bits of word i/2. Face-cell (guarded) stores
// take the top 16 bits of the reference's per-store hash instead (key =
// cell * 19 + slot direction). Stochastic rounding itself is "add the dither
// below the kept mantissa bits, truncate".
uint32_t cell_hash(uint32_t n, uint32_t seed) {
    uint32_t h = (n * 0x9E3779B1u) ^ (seed * 0x85EBCA77u + 0x632BE5ABu);
    h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
    return h;
}
inline uint32_t xorshift32(uint32_t x) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; return x; }
void dither16(uint32_t n, uint32_t seed, uint16_t d[19]) {
    uint32_t w = cell_hash(n, seed);
    for (int i = 0; i < 19; i += 2) {
        d[i] = (uint16_t)(w & 0xffff);
        if (i + 1 < 19) d[i + 1] = (uint16_t)(w >> 16);
        w = xorshift32(w);
    }
}
// No clamp before the conversion: round toward zero saturates finite values
// beyond the half range at +-65504 (as v_cvt_pkrtz_f16_f32 does; the
// reference clamps at +-65000 instead, shader.rs:516 — a difference only for
// |f~| > 1.98, far outside a stable flow).
uint16_t enc_f16_sr(float x, uint16_t d) {
    return float_to_half_rtz(bitsf(fbits(x * FS) + (uint32_t)(d >> 3)));
}
uint16_t enc_bf16_sr(float x, uint16_t d) { return (uint16_t)((fbits(x) + (uint32_t)d) >> 16); }
float dec_bf16(uint16_t b) { return bitsf((uint32_t)b << 16); }

template <class T> struct V3 {
    T x, y, z;
    V3 operator+(const V3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(T s) const { return {x * s, y * s, z * s}; }
    T &operator[](int k) { return k == 0 ? x : (k == 1 ? y : z); }
    T operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
template <class T> inline T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(const V3<T> &a, const V3<T> &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// A wall record (shader.rs:62-69), kept unpacked so the FP64 oracle keeps
// FP64 payloads; `records()` packs the reference's 40 words.
template <class T> struct Rec {
    uint32_t mask = 0, cell = 0;
    uint32_t q[9] = {};
    T fpost[18] = {};   // directions 1..18, shifted by w_i
    T mac[4] = {};      // rho, ux, uy, uz
    T ft[6] = {};       // F, T
};

template <class T> struct Mom { T rho1; V3<T> j; T s[6]; };

struct Geo {
    int moving = 0;
    int box_lo[3] = {0, 0, 0};
    uint32_t box_n[3] = {1, 1, 1};
    float occ_lo[3] = {0, 0, 0};
    float occ_invh = 1.0f;
    int occ_n[3] = {1, 1, 1};
    std::vector<uint32_t> occ;
    std::vector<float> tris;   // ntri x 12
};

struct StepU {   // the per-step uniform (lattice.rs:123-132), as f32
    float r[3][3], q[3][3], omega[3];
    uint32_t seed;
};

template <class T> struct Lat {
    uint32_t NX, NY, NZ;
    uint64_t N;
    bool per[3];
    int store;
    T tau0, smag2, sponge_tau, sigma;
    bool abs_u, total_p, wall_model, forcing, sponge_on, debug_dump;
    T sp_lo[3], sp_hi[3];
    V3<T> u_inf, u_init, grav, center;
    T w[19];
    std::vector<float> pf;
    std::vector<double> pd;
    std::vector<uint16_t> ph;
    std::vector<uint32_t> flags, faces;
    std::vector<Rec<T>> recs;
    std::vector<T> facebuf;
    std::vector<uint32_t> qscr;   // moving: 9 links x (min, max, count) per box cell
    Geo geo;
    // probes
    std::vector<float> verts;     // nvert x 8 (pos.xyz,0, normal.xyz,0)
    std::vector<int32_t> slicedef;
    float voff = 0.0f;
    std::vector<float> vacc, sacc;
    uint32_t npix = 0;
    uint64_t steps_done = 0;
    int nthreads = 0;
    // test mode: the face update of step k+1 done right after step k (the
    // GPU library's order) instead of at the start of step k+1
    bool fused = false;
    bool first_of_run = true;

    // ---- slots (shader.rs:147-162) ----
    inline uint64_t slot(int i, uint64_t cell) const { return (uint64_t)i * N + cell; }
    inline T ld(int i, uint64_t cell) const {
        uint64_t s = slot(i, cell);
        switch (store) {
            case ST_F32: return (T)pf[s];
            case ST_F64: return (T)pd[s];
            case ST_BF16: return (T)dec_bf16(ph[s]);
            default: return (T)(half_to_float(ph[s]) * INV_S);
        }
    }
    // key: the storing thread's cell n (shader.rs:158 uses n * 19 + i)
    inline void st(int i, uint64_t cell, T v, uint32_t n, uint32_t seed, const uint16_t *dz) {
        uint64_t s = slot(i, cell);
        switch (store) {
            case ST_F32: pf[s] = (float)v; break;
            case ST_F64: pd[s] = (double)v; break;
            case ST_BF16: ph[s] = enc_bf16_sr((float)v, dz[i]); break;
            case ST_F16: ph[s] = enc_f16_sr((float)v, dz[i]); break;
            default: ph[s] = float_to_half_rne(sround_ref((float)v * FS, n * 19u + (uint32_t)i, seed)); break;
        }
    }

    // ---- neighbours, shader.rs:247-252 (always wrapping) ----
    inline void coords(uint64_t n, uint32_t &x, uint32_t &y, uint32_t &z) const {
        z = (uint32_t)(n / ((uint64_t)NX * NY));
        uint64_t r = n - (uint64_t)z * NX * NY;
        y = (uint32_t)(r / NX);
        x = (uint32_t)(r - (uint64_t)y * NX);
    }
    inline uint64_t cidx(uint32_t x, uint32_t y, uint32_t z) const { return ((uint64_t)z * NY + y) * NX + x; }
    inline uint64_t nb(uint32_t x, uint32_t y, uint32_t z, int dx, int dy, int dz) const {
        uint32_t xx = dx > 0 ? (x == NX - 1 ? 0 : x + 1) : dx < 0 ? (x == 0 ? NX - 1 : x - 1) : x;
        uint32_t yy = dy > 0 ? (y == NY - 1 ? 0 : y + 1) : dy < 0 ? (y == 0 ? NY - 1 : y - 1) : y;
        uint32_t zz = dz > 0 ? (z == NZ - 1 ? 0 : z + 1) : dz < 0 ? (z == 0 ? NZ - 1 : z - 1) : z;
        return cidx(xx, yy, zz);
    }

    // AA access (shader.rs:254-339), named by the parity q of the step whose
    // inputs (pre) or outputs (post) are meant:
    //   load_in(p) = gather_pre(p)        store_out(p) = scatter_post(p)
    //   load_prev_out in module p = gather_post(1-p)
    //   store_prev_out in module p = scatter_post(1-p, guarded)
    //   store_in(p) = scatter_pre(p)
    void gather_pre(uint64_t n, int q, T f[19]) const {
        uint32_t x, y, z; coords(n, x, y, z);
        for (int i = 0; i < 19; ++i)
            f[i] = (q == 0 || i == 0) ? ld(i, n) : ld(opp(i), nb(x, y, z, -CV[i][0], -CV[i][1], -CV[i][2]));
    }
    void gather_post(uint64_t n, int q, T f[19]) const {
        uint32_t x, y, z; coords(n, x, y, z);
        for (int i = 0; i < 19; ++i)
            f[i] = (q == 0 || i == 0) ? ld(opp(i), n) : ld(i, nb(x, y, z, CV[i][0], CV[i][1], CV[i][2]));
    }
    void scatter_post(uint64_t n, int q, const T f[19], bool guard, uint32_t seed) {
        uint32_t x, y, z; coords(n, x, y, z);
        uint16_t dz[19];
        if (store == ST_F16 || store == ST_BF16) {
            if (guard) {   // face cells / refills: one hash per store, key n*19 + slot direction
                for (int j = 0; j < 19; ++j) dz[j] = (uint16_t)(cell_hash((uint32_t)n * 19u + (uint32_t)j, seed) >> 16);
            } else {
                dither16((uint32_t)n, seed, dz);
            }
        }
        for (int i = 0; i < 19; ++i) {
            if (q == 0 || i == 0) { st(opp(i), n, f[i], (uint32_t)n, seed, dz); continue; }
            if (guard) {   // shader.rs:287-297
                bool ok = true;
                const uint32_t c[3] = {x, y, z}, lim[3] = {NX, NY, NZ};
                for (int k = 0; k < 3; ++k) {
                    if (CV[i][k] == 1 && !(per[k] || c[k] < lim[k] - 1)) ok = false;
                    if (CV[i][k] == -1 && !(per[k] || c[k] > 0)) ok = false;
                }
                if (!ok) continue;
            }
            st(i, nb(x, y, z, CV[i][0], CV[i][1], CV[i][2]), f[i], (uint32_t)n, seed, dz);
        }
    }
    void scatter_pre(uint64_t n, int q, const T f[19], uint32_t seed) {
        uint32_t x, y, z; coords(n, x, y, z);
        uint16_t dz[19]; if (store == ST_F16 || store == ST_BF16) dither16((uint32_t)n, seed, dz);
        for (int i = 0; i < 19; ++i) {
            if (q == 0 || i == 0) st(i, n, f[i], (uint32_t)n, seed, dz);
            else st(opp(i), nb(x, y, z, -CV[i][0], -CV[i][1], -CV[i][2]), f[i], (uint32_t)n, seed, dz);
        }
    }

    // ---- shader.rs:341-396 ----
    // the refill's share of one fluid neighbour (shader.rs:1062-1078): the
    // traceless non-equilibrium stress of its recorded post-collision state
    static void add_neighbour_stress(const Rec<T> &R2, T pis[6]) {
        T fp[19];
        fp[0] = 0;
        for (int j = 1; j < 19; ++j) fp[j] = R2.fpost[j - 1];
        Mom<T> mo2 = moments(fp);
        T r2 = R2.mac[0];
        V3<T> u2 = {R2.mac[1], R2.mac[2], R2.mac[3]};
        pis[3] = pis[3] + (mo2.s[3] - r2 * u2.x * u2.y);
        pis[4] = pis[4] + (mo2.s[4] - r2 * u2.x * u2.z);
        pis[5] = pis[5] + (mo2.s[5] - r2 * u2.y * u2.z);
        T tr = (mo2.s[0] + mo2.s[1] + mo2.s[2]) / (T)3;
        T uu = dot(u2, u2);
        pis[0] = pis[0] + (mo2.s[0] - tr - r2 * (u2.x * u2.x - uu / (T)3));
        pis[1] = pis[1] + (mo2.s[1] - tr - r2 * (u2.y * u2.y - uu / (T)3));
        pis[2] = pis[2] + (mo2.s[2] - tr - r2 * (u2.z * u2.z - uu / (T)3));
    }

    static Mom<T> moments(const T f[19]) {
        Mom<T> m;
        T r = 0, jx = 0, jy = 0, jz = 0, s[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 19; ++i) {   // same sum order as the generated text
            r = r + f[i];
        }
        for (int i = 0; i < 19; ++i) { if (CV[i][0] == 1) jx = jx + f[i]; else if (CV[i][0] == -1) jx = jx - f[i]; }
        for (int i = 0; i < 19; ++i) { if (CV[i][1] == 1) jy = jy + f[i]; else if (CV[i][1] == -1) jy = jy - f[i]; }
        for (int i = 0; i < 19; ++i) { if (CV[i][2] == 1) jz = jz + f[i]; else if (CV[i][2] == -1) jz = jz - f[i]; }
        const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
        for (int k = 0; k < 6; ++k)
            for (int i = 0; i < 19; ++i) {
                int c = CV[i][pa[k]] * CV[i][pb[k]];
                if (c == 1) s[k] = s[k] + f[i]; else if (c == -1) s[k] = s[k] - f[i];
            }
        m.rho1 = r; m.j = {jx, jy, jz};
        for (int k = 0; k < 6; ++k) m.s[k] = s[k];
        return m;
    }
    void feq(T rho1, V3<T> u, T f[19]) const {
        T rho = (T)1 + rho1;
        T usq = (T)1.5 * dot(u, u);
        for (int i = 0; i < 19; ++i) {
            T cu = (T)0;
            for (int k = 0; k < 3; ++k) { if (CV[i][k] == 1) cu = cu + u[k]; else if (CV[i][k] == -1) cu = cu - u[k]; }
            f[i] = w[i] * (rho1 + rho * ((T)3 * cu + (T)4.5 * cu * cu - usq));
        }
    }
    void fneq(const T p[6], T f[19]) const {
        T tr = (p[0] + p[1] + p[2]) * ((T)1 / (T)3);
        const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
        for (int i = 0; i < 19; ++i) {
            T t = -tr;
            for (int k = 0; k < 6; ++k) {
                int v = CV[i][pa[k]] * CV[i][pb[k]] * (pa[k] == pb[k] ? 1 : 2);
                if (v != 0) t = t + (T)v * p[k];
            }
            f[i] = (T)4.5 * w[i] * t;
        }
    }

    // ---- shader.rs:528-538 ---- layer weight; with the viscosity ramp and
    // the relaxation to feq(1, t) pinned by one- and two-step closed forms
    T sponge(uint32_t x, uint32_t y, uint32_t z) const {
        T p[3] = {(T)x, (T)y, (T)z};
        T hi[3] = {(T)(NX - 1) - p[0], (T)(NY - 1) - p[1], (T)(NZ - 1) - p[2]};
        T s = 0;
        for (int k = 0; k < 3; ++k) {
            if (sp_lo[k] > 0) s = std::max(s, (T)1 - p[k] / sp_lo[k]);
            if (sp_hi[k] > 0) s = std::max(s, (T)1 - hi[k] / sp_hi[k]);
        }
        return std::min(std::max(s, (T)0), (T)1);
    }

    static float qof(const Rec<T> &r, int i) {
        uint32_t w = r.q[(i - 1) / 2];
        return half_to_float((uint16_t)(((i - 1) % 2) ? (w >> 16) : (w & 0xffff)));
    }

    // ---- shader.rs:546-597 wall_tau (Spalding) ---- pinned by
    // tests/test_oracle_wall_model.py: the wall normal and distance from the
    // cut links, and the friction velocity against Spalding's law solved
    // independently
    T wall_tau(const Rec<T> &r, uint32_t mask, V3<T> pc, V3<T> u, T tau_les, const V3<T> &om) const {
        V3<T> nsum = {0, 0, 0};
        for (int i = 1; i < 19; ++i)
            if (mask & (1u << i)) {
                T q = (T)qof(r, i);
                V3<T> ci = {(T)CV[i][0], (T)CV[i][1], (T)CV[i][2]};
                nsum = nsum + ci * ((i < 7 ? (T)1 / (T)18 : (T)1 / (T)36) * ((T)1.5 - q));
            }
        T nl = std::sqrt(dot(nsum, nsum));
        if (nl < (T)1e-6) return tau_les;
        V3<T> nw = {nsum.x / nl, nsum.y / nl, nsum.z / nl};
        T y = 2;
        for (int i = 1; i < 19; ++i)
            if (mask & (1u << i)) {
                V3<T> ci = {(T)CV[i][0], (T)CV[i][1], (T)CV[i][2]};
                T cn = dot(ci, nw);
                if (cn > (T)0.3) {
                    T q = (T)qof(r, i);
                    y = std::min(y, std::max(q, (T)0.02) * cn);
                }
            }
        if (y > (T)1.9) return tau_les;
        V3<T> uwall = cross(om, pc + nw * y);
        V3<T> ur = u - uwall;
        V3<T> tvec = ur - nw * dot(ur, nw);
        T ut = std::sqrt(dot(tvec, tvec));
        T nu0 = (tau0 - (T)0.5) / (T)3;
        if (ut < (T)1e-7) return tau_les;
        const T kap = (T)0.41;
        const T eb = std::exp(-kap * (T)5.2);
        T ust = std::sqrt(nu0 * ut / y);
        for (int it = 0; it < 12; ++it) {
            T up = ut / ust;
            T yp = y * ust / nu0;
            T ku = std::min(kap * up, (T)30);
            T g = up + eb * (std::exp(ku) - (T)1 - ku - (T)0.5 * ku * ku - ku * ku * ku / (T)6) - yp;
            T dgdup = (T)1 + eb * kap * (std::exp(ku) - (T)1 - ku - (T)0.5 * ku * ku);
            T dg = dgdup * (-up / ust) - y / nu0;
            T step = g / dg;
            ust = std::min(std::max(ust - step, (T)0.1 * ust), (T)10 * ust);
            if (std::fabs(step) < (T)1e-5 * ust) break;
        }
        T nuw = ust * ust * y / ut;
        return std::max((T)0.5 + (T)3 * std::max(nuw, nu0), tau_les);
    }

    // ---- shader.rs:795-862 main_step (+ gen_main :399-458) for one cell ----
    void main_cell(uint64_t n, int p, const StepU &su, bool moving) {
        uint32_t fg = flags[n];
        if (fg & 3u) return;
        uint32_t x, y, z; coords(n, x, y, z);
        T f[19];
        gather_pre(n, p, f);
        uint32_t rec = fg >> 2;
        V3<T> F = {0, 0, 0}, Tq = {0, 0, 0};
        uint32_t mask = 0;
        V3<T> om = {(T)su.omega[0], (T)su.omega[1], (T)su.omega[2]};
        Rec<T> *R = nullptr;
        if (rec != 0) {
            R = &recs[rec - 1];
            mask = R->mask;
            T fin[19];
            for (int i = 0; i < 19; ++i) fin[i] = f[i];
            V3<T> pc = V3<T>{(T)x, (T)y, (T)z} - center;
            for (int i = 1; i < 19; ++i) {
                if (!(mask & (1u << i))) continue;
                int o = opp(i);
                T wi = w[i];
                T q = std::max((T)qof(*R, i), (T)0.001);
                T fpi = R->fpost[i - 1] + wi;
                T fpo = R->fpost[o - 1] + wi;
                V3<T> ci = {(T)CV[i][0], (T)CV[i][1], (T)CV[i][2]};
                V3<T> xw = pc + ci * q;
                V3<T> uw = cross(om, xw);
                T cu = dot(ci, uw);
                T fb;
                if (q >= (T)0.5) fb = (fpi + ((T)2 * q - (T)1) * fpo - (T)6 * wi * cu) / ((T)2 * q);
                else if (!(mask & (1u << o))) fb = (T)2 * q * fpi + ((T)1 - (T)2 * q) * (fin[i] + wi) - (T)6 * wi * cu;
                else fb = fpi - (T)6 * wi * cu;
                f[o] = fb - wi;
                V3<T> fl = (ci - uw) * fpi + (ci + uw) * fb;
                F = F + fl;
                Tq = Tq + cross(xw, fl);
            }
        }
        Mom<T> mo = moments(f);
        T rho = (T)1 + mo.rho1;
        V3<T> u = {mo.j.x / rho + (T)0.5 * grav.x, mo.j.y / rho + (T)0.5 * grav.y, mo.j.z / rho + (T)0.5 * grav.z};
        T pi[6];
        pi[0] = mo.s[0] - rho * u.x * u.x - mo.rho1 / (T)3;
        pi[1] = mo.s[1] - rho * u.y * u.y - mo.rho1 / (T)3;
        pi[2] = mo.s[2] - rho * u.z * u.z - mo.rho1 / (T)3;
        pi[3] = mo.s[3] - rho * u.x * u.y;
        pi[4] = mo.s[4] - rho * u.x * u.z;
        pi[5] = mo.s[5] - rho * u.y * u.z;
        T pp = pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] + (T)2 * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]);
        T qn = std::sqrt((T)2 * pp);
        T tau = (T)0.5 * (tau0 + std::sqrt(tau0 * tau0 + (T)18 * smag2 * qn / rho));
        if (wall_model && mask != 0)
            tau = wall_tau(*R, mask, V3<T>{(T)x, (T)y, (T)z} - center, u, tau, om);
        T sp = sponge_on ? sponge(x, y, z) : (T)0;
        if (sp > 0 && sponge_tau > tau) tau = tau + sp * sp * (sponge_tau - tau);
        T omg = (T)1 / tau;
        T fe[19], fq[19];
        feq(mo.rho1, u, fe);
        fneq(pi, fq);
        for (int i = 0; i < 19; ++i) f[i] = fe[i] + ((T)1 - omg) * fq[i];
        if (sp > 0 && sigma > 0) {
            V3<T> ut = abs_u ? u_inf : u;
            T ft[19];
            feq((T)0, ut, ft);
            T sg = sigma * sp * sp;
            for (int i = 0; i < 19; ++i) f[i] = f[i] + sg * (ft[i] - fe[i]);
        }
        V3<T> u_out = u;
        if (forcing) {   // shader.rs:442-455 Guo forcing
            V3<T> G = grav * rho;
            T uG = dot(u, G);
            T fo = (T)1 - (T)0.5 * omg;
            for (int i = 0; i < 19; ++i) {
                T cu = 0, cg = 0;
                for (int k = 0; k < 3; ++k) {
                    if (CV[i][k] == 1) { cu = cu + u[k]; cg = cg + G[k]; }
                    else if (CV[i][k] == -1) { cu = cu - u[k]; cg = cg - G[k]; }
                }
                f[i] = f[i] + w[i] * (fo * ((T)3 * (cg - uG) + (T)9 * cu * cg) - (T)1.5 * ((T)1 - omg) * cg);
            }
        }
        if (rec != 0) {
            for (int i = 1; i < 19; ++i) R->fpost[i - 1] = f[i];
            R->mac[0] = rho; R->mac[1] = u_out.x; R->mac[2] = u_out.y; R->mac[3] = u_out.z;
            if (moving) {
                F = F + V3<T>{R->ft[0], R->ft[1], R->ft[2]};
                Tq = Tq + V3<T>{R->ft[3], R->ft[4], R->ft[5]};
            }
            R->ft[0] = F.x; R->ft[1] = F.y; R->ft[2] = F.z;
            R->ft[3] = Tq.x; R->ft[4] = Tq.y; R->ft[5] = Tq.z;
        }
        scatter_post(n, p, f, false, su.seed);
    }

    // ---- shader.rs:630-687 face_gather / face_write ----
    void face_gather(uint32_t k, int p) {
        uint64_t b = faces[k];
        uint32_t x, y, z; coords(b, x, y, z);
        if (!per[0]) x = std::min(std::max(x, 1u), NX - 2);
        if (!per[1]) y = std::min(std::max(y, 1u), NY - 2);
        if (!per[2]) z = std::min(std::max(z, 1u), NZ - 2);
        uint64_t m = cidx(x, y, z);
        T rho = 1;
        V3<T> u = u_inf;
        T pi[6] = {0, 0, 0, 0, 0, 0};
        if ((flags[m] & 3u) == 0) {
            T f[19];
            gather_post(m, 1 - p, f);
            Mom<T> mo = moments(f);
            rho = (T)1 + mo.rho1;
            u = {mo.j.x / rho - (T)0.5 * grav.x, mo.j.y / rho - (T)0.5 * grav.y, mo.j.z / rho - (T)0.5 * grav.z};
            pi[0] = mo.s[0] - rho * u.x * u.x - mo.rho1 / (T)3;
            pi[1] = mo.s[1] - rho * u.y * u.y - mo.rho1 / (T)3;
            pi[2] = mo.s[2] - rho * u.z * u.z - mo.rho1 / (T)3;
            pi[3] = mo.s[3] - rho * u.x * u.y;
            pi[4] = mo.s[4] - rho * u.x * u.z;
            pi[5] = mo.s[5] - rho * u.y * u.z;
        }
        T *o = &facebuf[(size_t)k * 10];
        o[0] = rho; o[1] = u.x; o[2] = u.y; o[3] = u.z;
        for (int i = 0; i < 6; ++i) o[4 + i] = pi[i];
    }
    void face_write(uint32_t k, int p, uint32_t seed) {
        uint64_t b = faces[k];
        const T *o = &facebuf[(size_t)k * 10];
        T rho = o[0];
        V3<T> u = {o[1], o[2], o[3]};
        T pi[6];
        for (int i = 0; i < 6; ++i) pi[i] = o[4 + i];
        uint32_t ty = flags[b] & 3u;
        if (ty == 2) {
            if (total_p) rho = ((T)1 / (T)3 + (T)0.5 * dot(u_inf, u_inf)) / ((T)1 / (T)3 + (T)0.5 * dot(u, u));
            else u = u_inf;
        } else {
            rho = 1;
        }
        T fe[19], fq[19], f[19];
        feq(rho - (T)1, u, fe);
        fneq(pi, fq);
        for (int i = 0; i < 19; ++i) f[i] = fe[i] + fq[i];
        scatter_post(b, 1 - p, f, true, seed);
    }

    // ---- shader.rs:599-621 cell_macro (module parity p = the next step's) ----
    bool cell_macro(uint64_t m, int p, T &rho, V3<T> &u) const {
        uint32_t fg = flags[m];
        if (fg & 3u) return false;
        uint32_t rec = fg >> 2;
        if (rec != 0) {
            const Rec<T> &R = recs[rec - 1];
            rho = R.mac[0]; u = {R.mac[1], R.mac[2], R.mac[3]};
            return true;
        }
        T f[19];
        gather_post(m, 1 - p, f);
        Mom<T> mo = moments(f);
        rho = (T)1 + mo.rho1;
        u = {mo.j.x / rho - (T)0.5 * grav.x, mo.j.y / rho - (T)0.5 * grav.y, mo.j.z / rho - (T)0.5 * grav.z};
        return true;
    }

    // ---- moving geometry, shader.rs:865-1104 (f32, as the WGSL) ----
    static V3<float> to_body(const StepU &s, V3<float> v, bool prev) {
        const float (*m)[3] = prev ? s.q : s.r;   // m[k] = column k of the rotation
        return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
    }
    static V3<float> from_body(const StepU &s, V3<float> b) {
        return {s.r[0][0] * b.x + s.r[1][0] * b.y + s.r[2][0] * b.z, s.r[0][1] * b.x + s.r[1][1] * b.y + s.r[2][1] * b.z,
                s.r[0][2] * b.x + s.r[1][2] * b.y + s.r[2][2] * b.z};
    }
    bool occ_at(V3<float> b) const {
        int g[3];
        for (int k = 0; k < 3; ++k) g[k] = (int)std::nearbyint((b[k] - geo.occ_lo[k]) * geo.occ_invh);
        for (int k = 0; k < 3; ++k) if (g[k] < 0 || g[k] >= geo.occ_n[k]) return false;
        uint32_t l = (uint32_t)((g[2] * geo.occ_n[1] + g[1]) * geo.occ_n[0] + g[0]);
        return ((geo.occ[l >> 5] >> (l & 31)) & 1u) != 0;
    }
    V3<float> centerf() const { return {(float)center.x, (float)center.y, (float)center.z}; }
    bool solid_at(int px, int py, int pz, const StepU &s, bool before) const {
        V3<float> c = centerf();
        V3<float> pc = {(float)px - c.x, (float)py - c.y, (float)pz - c.z};
        return occ_at(to_body(s, pc, before));
    }
    static float ray_tri(V3<float> o, V3<float> d, V3<float> v0, V3<float> e1, V3<float> e2) {
        V3<float> pv = cross(d, e2);
        float det = dot(e1, pv);
        if (std::fabs(det) < 1e-12f) return 9.0f;
        float inv = 1.0f / det;
        V3<float> tv = o - v0;
        float a = dot(tv, pv) * inv;
        if (a < 0.0f || a > 1.0f) return 9.0f;
        V3<float> qv = cross(tv, e1);
        float b = dot(d, qv) * inv;
        if (b < 0.0f || a + b > 1.0f) return 9.0f;
        float t = dot(e2, qv) * inv;
        if (t < -1e-4f || t > 1.0f) return 9.0f;
        return t > 0.0f ? t : 0.0f;   // max(t, 0) as +0 for t = -0: hit bits must sort as the floats do
    }
    void tri_scatter(const StepU &s) {
        size_t ntri = geo.tris.size() / 12;
        V3<float> c = centerf();
        for (size_t t = 0; t < ntri; ++t) {
            const float *tr = &geo.tris[t * 12];
            V3<float> a = c + from_body(s, {tr[0], tr[1], tr[2]});
            V3<float> e1 = from_body(s, {tr[4], tr[5], tr[6]});
            V3<float> e2 = from_body(s, {tr[8], tr[9], tr[10]});
            V3<float> nrm = cross(e1, e2);
            float nl = std::sqrt(dot(nrm, nrm));
            if (nl < 1e-12f) continue;
            V3<float> nn = {nrm.x / nl, nrm.y / nl, nrm.z / nl};
            V3<float> b = a + e1, cc = a + e2;
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                float mn = std::min(std::min(a[k], b[k]), cc[k]), mx = std::max(std::max(a[k], b[k]), cc[k]);
                lo[k] = std::max((int)std::ceil(mn - 1.415f), geo.box_lo[k]);
                hi[k] = std::min((int)std::floor(mx + 1.415f), geo.box_lo[k] + (int)geo.box_n[k] - 1);
            }
            for (int z = lo[2]; z <= hi[2]; ++z)
                for (int y = lo[1]; y <= hi[1]; ++y)
                    for (int x = lo[0]; x <= hi[0]; ++x) {
                        V3<float> pp = {(float)x, (float)y, (float)z};
                        float d0 = dot(nn, pp - a);
                        if (std::fabs(d0) > 1.415f) continue;
                        uint32_t lx = x - geo.box_lo[0], ly = y - geo.box_lo[1], lz = z - geo.box_lo[2];
                        uint32_t k = (lz * geo.box_n[1] + ly) * geo.box_n[0] + lx;
                        for (int i = 1; i < 19; i += 2) {
                            V3<float> ci = {(float)CV[i][0], (float)CV[i][1], (float)CV[i][2]};
                            float d1 = d0 + dot(nn, ci);
                            if (d0 * d1 > 0.0f && std::fabs(d0) > 1e-6f) continue;
                            float tt = ray_tri(pp, ci, a, e1, e2);
                            if (tt <= 1.0f) {
                                size_t o = ((size_t)k * 9 + (i - 1) / 2) * 3;
                                uint32_t tb = fbits(tt);   // tt >= +0 (ray_tri)
                                qscr[o] = std::min(qscr[o], tb);
                                qscr[o + 1] = std::max(qscr[o + 1], tb);
                                qscr[o + 2] += 1;
                            }
                        }
                    }
        }
    }
    void qclear() {
        for (size_t k = 0; k < qscr.size() / 3; ++k) { qscr[3 * k] = 0x7f7fffffu; qscr[3 * k + 1] = 0; qscr[3 * k + 2] = 0; }
    }
    void voxelize_cell(uint32_t k, int p, const StepU &s) {
        const uint32_t *BN = geo.box_n;
        uint32_t lx = k % BN[0], ly = (k / BN[0]) % BN[1], lz = k / (BN[0] * BN[1]);
        int px = (int)lx + geo.box_lo[0], py = (int)ly + geo.box_lo[1], pz = (int)lz + geo.box_lo[2];
        uint64_t n = cidx(px, py, pz);
        V3<float> c = centerf();
        V3<float> pc = {(float)px - c.x, (float)py - c.y, (float)pz - c.z};
        bool solid_now = solid_at(px, py, pz, s, false);
        bool solid_before = solid_at(px, py, pz, s, true);
        Rec<T> &R = recs[k];
        if (solid_now) {
            flags[n] = 1u | ((k + 1u) << 2);
            R.mask = 0;
            V3<T> F = {0, 0, 0};
            if (!solid_before) { T rho = R.mac[0]; F = V3<T>{R.mac[1], R.mac[2], R.mac[3]} * rho; }
            V3<T> pcT = {(T)pc.x, (T)pc.y, (T)pc.z};
            V3<T> Tq = cross(pcT, F);
            R.ft[0] = F.x; R.ft[1] = F.y; R.ft[2] = F.z; R.ft[3] = Tq.x; R.ft[4] = Tq.y; R.ft[5] = Tq.z;
            return;
        }
        float qv[19], qx[19];
        uint32_t nh[19];
        for (int i = 1; i < 19; ++i) { qv[i] = 9.0f; qx[i] = -1.0f; nh[i] = 0; }
        for (int i = 1; i < 19; ++i) {
            if (i & 1) {
                size_t o = ((size_t)k * 9 + (i - 1) / 2) * 3;
                uint32_t n_ = qscr[o + 2];
                if (n_ > 0) { qv[i] = bitsf(qscr[o]); qx[i] = bitsf(qscr[o + 1]); nh[i] = n_; }
            } else {
                int q[3] = {(int)lx + CV[i][0], (int)ly + CV[i][1], (int)lz + CV[i][2]};
                if (q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < (int)BN[0] && q[1] < (int)BN[1] && q[2] < (int)BN[2]) {
                    uint32_t k2 = (uint32_t)((q[2] * (int)BN[1] + q[1]) * (int)BN[0] + q[0]);
                    size_t o = ((size_t)k2 * 9 + (i - 2) / 2) * 3;
                    uint32_t n_ = qscr[o + 2];
                    if (n_ > 0) { qv[i] = 1.0f - bitsf(qscr[o + 1]); qx[i] = 1.0f - bitsf(qscr[o]); nh[i] = n_; }
                }
            }
        }
        uint32_t mask = 0;
        for (int i = 1; i < 19; ++i) {
            if (solid_at(px + CV[i][0], py + CV[i][1], pz + CV[i][2], s, false)) {
                if (nh[i] == 0) qv[i] = 1.0f;
                mask |= 1u << i;
            } else if (nh[i] >= 2 && qx[i] - qv[i] > 1e-5f) {
                mask |= 1u << i;
            } else {
                qv[i] = 1.0f;
            }
        }
        for (int wv = 0; wv < 9; ++wv)   // pack2x16float: RNE (parity unpinned vs Metal)
            R.q[wv] = (uint32_t)float_to_half_rne(std::min(qv[2 * wv + 1], 1.0f)) |
                      ((uint32_t)float_to_half_rne(std::min(qv[2 * wv + 2], 1.0f)) << 16);
        R.mask = mask;
        V3<T> F = {0, 0, 0};
        if (solid_before) {
            T rs = 0, cnt = 0;
            T pis[6] = {0, 0, 0, 0, 0, 0};
            for (int i = 1; i < 19; ++i) {
                bool fluid_now = !solid_at(px + CV[i][0], py + CV[i][1], pz + CV[i][2], s, false);
                bool fluid_before = !solid_at(px + CV[i][0], py + CV[i][1], pz + CV[i][2], s, true);
                if (fluid_now && fluid_before) {
                    int q[3] = {(int)lx + CV[i][0], (int)ly + CV[i][1], (int)lz + CV[i][2]};
                    if (q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < (int)BN[0] && q[1] < (int)BN[1] && q[2] < (int)BN[2]) {
                        uint32_t k2 = (uint32_t)((q[2] * (int)BN[1] + q[1]) * (int)BN[0] + q[0]);
                        const Rec<T> &R2 = recs[k2];
                        rs = rs + R2.mac[0];
                        cnt = cnt + 1;
                        add_neighbour_stress(R2, pis);
                    }
                }
            }
            // (the mean non-equilibrium stress: add_neighbour_stress, pinned
            // by tests/test_oracle_spin.py)
            T rho = cnt > 0 ? rs / cnt : (T)1;
            V3<float> uwf = cross(V3<float>{s.omega[0], s.omega[1], s.omega[2]}, pc);
            V3<T> uw = {(T)uwf.x, (T)uwf.y, (T)uwf.z};
            T fe[19];
            feq(rho - (T)1, uw, fe);
            if (cnt > 0) {
                for (int j = 0; j < 6; ++j) pis[j] = pis[j] / cnt;
                T fq[19];
                fneq(pis, fq);
                for (int j = 0; j < 19; ++j) fe[j] = fe[j] + fq[j];
            }
            scatter_post(n, 1 - p, fe, true, s.seed);
            F = uw * (-rho);
            for (int i = 1; i < 19; ++i) R.fpost[i - 1] = fe[i];
            R.mac[0] = rho; R.mac[1] = uw.x; R.mac[2] = uw.y; R.mac[3] = uw.z;
        }
        V3<T> pcT = {(T)pc.x, (T)pc.y, (T)pc.z};
        V3<T> Tq = cross(pcT, F);
        R.ft[0] = F.x; R.ft[1] = F.y; R.ft[2] = F.z; R.ft[3] = Tq.x; R.ft[4] = Tq.y; R.ft[5] = Tq.z;
        flags[n] = (k + 1u) << 2;
    }

    // ---- shader.rs:736-784 samplers (module parity p) ----
    // (pinned: a linear density field at rest is read exactly, vertices at
    // centre + body position + offset * normal, slices cell by cell)
    void sample_verts(int p, const StepU &s) {
        size_t nv = verts.size() / 8;
        V3<float> c = centerf();
        for (size_t v = 0; v < nv; ++v) {
            V3<float> pb = {verts[8 * v], verts[8 * v + 1], verts[8 * v + 2]};
            V3<float> nbv = {verts[8 * v + 4], verts[8 * v + 5], verts[8 * v + 6]};
            V3<float> sp = c + from_body(s, pb + nbv * voff);
            V3<float> b0 = {std::floor(sp.x), std::floor(sp.y), std::floor(sp.z)};
            V3<float> fr = sp - b0;
            float acc = 0.0f, wsum = 0.0f;
            for (int k = 0; k < 8; ++k) {
                int o[3] = {k & 1, (k >> 1) & 1, (k >> 2) & 1};
                int cc[3] = {(int)b0.x + o[0], (int)b0.y + o[1], (int)b0.z + o[2]};
                if (cc[0] < 1 || cc[1] < 1 || cc[2] < 1 || cc[0] >= (int)NX - 1 || cc[1] >= (int)NY - 1 || cc[2] >= (int)NZ - 1) continue;
                float wv[3];
                for (int a = 0; a < 3; ++a) wv[a] = o[a] == 1 ? fr[a] : 1.0f - fr[a];
                float wgt = wv[0] * wv[1] * wv[2];
                T rho; V3<T> u;
                bool ok = cell_macro(cidx(cc[0], cc[1], cc[2]), p, rho, u);
                if (ok && wgt > 0.0f) { acc += wgt * (float)rho; wsum += wgt; }
            }
            if (wsum > 1e-6f) { vacc[2 * v] += acc / wsum; vacc[2 * v + 1] += 1.0f; }
        }
    }
    void sample_slices(int p) {
        int nsl = slicedef.empty() ? 0 : slicedef[0];
        for (uint32_t k = 0; k < npix; ++k) {
            uint32_t acc = 0;
            for (int s = 0; s < nsl; ++s) {
                const int32_t *o = &slicedef[4 * (1 + 3 * s)], *du = &slicedef[4 * (2 + 3 * s)], *dv = &slicedef[4 * (3 + 3 * s)];
                uint32_t nu = (uint32_t)o[3], nv = (uint32_t)du[3];
                if (k < acc + nu * nv) {
                    uint32_t l = k - acc, iv = l / nu, iu = l - iv * nu;
                    int cc[3];
                    for (int a = 0; a < 3; ++a) cc[a] = o[a] + (int)iu * du[a] + (int)iv * dv[a];
                    T rho; V3<T> u;
                    if (cell_macro(cidx(cc[0], cc[1], cc[2]), p, rho, u)) {
                        float *b = &sacc[5 * (size_t)k];
                        b[0] += (float)rho; b[1] += (float)u.x; b[2] += (float)u.y; b[3] += (float)u.z; b[4] += 1.0f;
                    }
                    break;
                }
                acc += nu * nv;
            }
        }
    }

    // lattice.rs:675-686 per-step seed
    uint32_t seed_of(uint64_t step) const { return (uint32_t)step * 2654435761u; }

    // lattice.rs:675-686: the step's uniform (rotation columns as f32, seed)
    StepU stepu(const simlab_step_in &si) const {
        StepU s;
        for (int k = 0; k < 3; ++k)
            for (int r = 0; r < 3; ++r) { s.r[k][r] = (float)si.rot[r][k]; s.q[k][r] = (float)si.rot_prev[r][k]; }
        for (int k = 0; k < 3; ++k) s.omega[k] = (float)si.omega[k];
        s.seed = seed_of(steps_done);
        return s;
    }
    // the re-voxelization stage of a step (lattice.rs:720-726)
    void spin_stage(const simlab_step_in &si) {
        int p = (int)(steps_done % 2);
        StepU s = stepu(si);
        tri_scatter(s);
        for (uint32_t k = 0; k < (uint32_t)recs.size(); ++k) voxelize_cell(k, p, s);
        qclear();
    }

    // lattice.rs:715-738: one step
    void step(const simlab_step_in &si, double ft[6]) {
        int p = (int)(steps_done % 2);
        StepU s = stepu(si);
        bool moving = geo.moving != 0;
        if (moving) spin_stage(si);
        int64_t nf = (int64_t)faces.size();
        if (!fused || first_of_run) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int64_t k = 0; k < nf; ++k) face_gather((uint32_t)k, p);
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int64_t k = 0; k < nf; ++k) face_write((uint32_t)k, p, s.seed);
            if (fused) {   // inputs of the step after this one too (after the gather above)
                for (int64_t k = 0; k < nf; ++k) face_gather((uint32_t)k, 1 - p);
                for (int64_t k = 0; k < nf; ++k) face_write((uint32_t)k, 1 - p, seed_of(steps_done + 1));
            }
        }
        first_of_run = false;
        int64_t nn = (int64_t)N;
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int64_t n = 0; n < nn; ++n) main_cell((uint64_t)n, p, s, moving);
        if (fused) {   // the next step's face update, right after this step
            for (int64_t k = 0; k < nf; ++k) face_gather((uint32_t)k, 1 - p);
            for (int64_t k = 0; k < nf; ++k) face_write((uint32_t)k, 1 - p, seed_of(steps_done + 1));
        }
        double acc[6] = {0, 0, 0, 0, 0, 0};   // FP64 accumulation in record order
        for (const Rec<T> &R : recs)
            for (int j = 0; j < 6; ++j) acc[j] += (double)R.ft[j];
        for (int j = 0; j < 6; ++j) ft[j] = acc[j];
        if (si.sample) { sample_verts(1 - p, s); sample_slices(1 - p); }
        steps_done += 1;
    }
};

// ---- host setup, lattice.rs:229-526 ----
template <class T> Lat<T> *make_lat(const simlab_lattice_setup *su, const simlab_geometry *g, const simlab_probes *pr,
                                    int store, int nthreads, std::string &err) {
    uint64_t n = (uint64_t)su->dims[0] * su->dims[1] * su->dims[2];
    if (n == 0 || n >= 0xffffffffull) { err = "grid of " + std::to_string(n) + " cells is out of range"; return nullptr; }
    for (int k = 0; k < 3; ++k) if (su->dims[k] < 3) { err = "grid must be at least 3 cells in every direction"; return nullptr; }
    for (int a = 0; a < 3; ++a)
        if ((su->faces[2 * a] == SIMLAB_FACE_PERIODIC) != (su->faces[2 * a + 1] == SIMLAB_FACE_PERIODIC)) {
            err = "periodic faces must come in pairs"; return nullptr;
        }
    auto *L = new Lat<T>();
    L->nthreads = nthreads > 0 ? nthreads : 1;
    L->NX = su->dims[0]; L->NY = su->dims[1]; L->NZ = su->dims[2]; L->N = n;
    for (int a = 0; a < 3; ++a) L->per[a] = su->faces[2 * a] == SIMLAB_FACE_PERIODIC;
    L->store = store;
    L->tau0 = (T)(0.5 + 3.0 * su->nu);
    L->smag2 = (T)(su->smagorinsky * su->smagorinsky);
    L->sponge_tau = (T)su->sponge_tau; L->sigma = (T)su->sponge_sigma;
    L->abs_u = su->absorb_velocity != 0; L->total_p = su->total_pressure_inlet != 0;
    L->wall_model = su->wall_model != 0; L->debug_dump = su->debug_dump != 0;
    for (int k = 0; k < 3; ++k) { L->sp_lo[k] = (T)su->sponge[2 * k]; L->sp_hi[k] = (T)su->sponge[2 * k + 1]; }
    L->sponge_on = false;
    for (int k = 0; k < 6; ++k) if (su->sponge[k]) L->sponge_on = true;
    L->u_inf = {(T)su->u_inf[0], (T)su->u_inf[1], (T)su->u_inf[2]};
    L->u_init = {(T)su->u_init[0], (T)su->u_init[1], (T)su->u_init[2]};
    L->grav = {(T)su->body_force[0], (T)su->body_force[1], (T)su->body_force[2]};
    L->center = {(T)su->center[0], (T)su->center[1], (T)su->center[2]};
    L->forcing = su->body_force[0] != 0 || su->body_force[1] != 0 || su->body_force[2] != 0;
    for (int i = 0; i < 19; ++i) L->w[i] = (T)weight_d(i);
    if (store == ST_F32) L->pf.assign(19 * n, 0.0f);
    else if (store == ST_F64) L->pd.assign(19 * n, 0.0);
    else L->ph.assign(19 * n, 0);

    const uint32_t nx = L->NX, ny = L->NY, nz = L->NZ;
    auto face_type = [&](uint32_t x, uint32_t y, uint32_t z) -> int {   // lattice.rs:286-305
        bool on[6] = {x == 0, x == nx - 1, y == 0, y == ny - 1, z == 0, z == nz - 1};
        int ty = -1;
        for (int f = 0; f < 6; ++f)
            if (on[f] && su->faces[f] != SIMLAB_FACE_PERIODIC) {
                int t = su->faces[f] == SIMLAB_FACE_VELOCITY ? 2 : 3;
                ty = (ty == 2) ? 2 : t;
            }
        return ty;
    };
    for (uint32_t z = 0; z < nz; ++z)   // lattice.rs:306-323
        for (uint32_t y = 0; y < ny; ++y) {
            bool edge = y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
            if (edge) {
                for (uint32_t x = 0; x < nx; ++x) if (face_type(x, y, z) >= 0) L->faces.push_back((z * ny + y) * nx + x);
            } else {
                for (uint32_t x : {0u, nx - 1}) if (face_type(x, y, z) >= 0) L->faces.push_back((z * ny + y) * nx + x);
            }
        }
    L->facebuf.assign(L->faces.size() * 10, (T)0);

    // records (lattice.rs:454-482, record_words :202-218)
    std::vector<std::pair<uint32_t, uint32_t>> rec_of;
    double ui[3] = {su->u_init[0], su->u_init[1], su->u_init[2]};
    double usq = 1.5 * (ui[0] * ui[0] + ui[1] * ui[1] + ui[2] * ui[2]);
    // feq_shifted (lattice.rs:191-200), computed in f64 and stored as f32
    // words; the FP64 storage keeps the f64 values (no f32 record words)
    T fe_sh[19];
    for (int i = 0; i < 19; ++i) {
        double cu = CV[i][0] * ui[0] + CV[i][1] * ui[1] + CV[i][2] * ui[2];
        double v = weight_d(i) * (3.0 * cu + 4.5 * cu * cu - usq);
        fe_sh[i] = store == ST_F64 ? (T)v : (T)(float)v;
    }
    auto init_rec = [&](Rec<T> &R, uint32_t cell, uint32_t mask, const float *q) {
        R.mask = mask; R.cell = cell;
        for (int k = 0; k < 9; ++k) R.q[k] = (uint32_t)f16_bits_host(q[2 * k + 1]) | ((uint32_t)f16_bits_host(q[2 * k + 2]) << 16);
        for (int i = 1; i < 19; ++i) R.fpost[i - 1] = fe_sh[i];
        R.mac[0] = (T)1.0;
        for (int k = 0; k < 3; ++k) R.mac[1 + k] = store == ST_F64 ? (T)ui[k] : (T)(float)ui[k];
        for (int k = 0; k < 6; ++k) R.ft[k] = 0;
    };
    std::vector<uint32_t> solid;
    if (g && g->kind == SIMLAB_GEOM_STATIC) {
        L->recs.resize(g->nwall);
        for (uint64_t r = 0; r < g->nwall; ++r) {
            init_rec(L->recs[r], g->walls[r].cell, g->walls[r].mask, g->walls[r].q);
            rec_of.push_back({g->walls[r].cell, (uint32_t)r + 1});
        }
        solid.assign(g->solid, g->solid + g->nsolid);
    } else if (g && g->kind == SIMLAB_GEOM_SPIN) {
        for (int k = 0; k < 3; ++k)
            if (g->box_lo[k] < 1 || (int64_t)g->box_lo[k] + g->box_n[k] > (int64_t)su->dims[k] - 1) {
                err = "the spinning part's swept box leaves the domain"; delete L; return nullptr;
            }
        Geo &G = L->geo;
        G.moving = 1;
        for (int k = 0; k < 3; ++k) {
            G.box_lo[k] = g->box_lo[k]; G.box_n[k] = g->box_n[k];
            G.occ_lo[k] = (float)g->occ_lo[k]; G.occ_n[k] = (int)g->occ_dims[k];
        }
        G.occ_invh = (float)(1.0 / g->occ_h);
        G.occ.assign(g->occ_bits, g->occ_bits + g->nocc_words);
        G.tris.assign(g->tris, g->tris + g->ntri * 12);
        uint64_t nrec = (uint64_t)g->box_n[0] * g->box_n[1] * g->box_n[2];
        L->recs.resize(nrec);
        float ones[27];
        for (float &o : ones) o = 1.0f;
        uint32_t k = 0;
        for (uint32_t z = 0; z < g->box_n[2]; ++z)
            for (uint32_t y = 0; y < g->box_n[1]; ++y)
                for (uint32_t x = 0; x < g->box_n[0]; ++x) {
                    uint64_t c = ((uint64_t)(z + g->box_lo[2]) * ny + (y + g->box_lo[1])) * nx + x + g->box_lo[0];
                    init_rec(L->recs[k], (uint32_t)c, 0, ones);
                    rec_of.push_back({(uint32_t)c, k + 1});
                    ++k;
                }
        L->qscr.assign(nrec * 27, 0);
        L->qclear();
    }
    std::sort(rec_of.begin(), rec_of.end());
    std::sort(solid.begin(), solid.end());
    L->flags.assign(n, 0);   // lattice.rs:489-526 (order of precedence: face, solid, record)
    for (uint64_t c = 0; c < n; ++c) {
        uint32_t x, y, z; L->coords(c, x, y, z);
        int t = face_type(x, y, z);
        if (t >= 0) L->flags[c] = (uint32_t)t;
    }
    for (uint32_t c : solid) if (L->flags[c] == 0) L->flags[c] = 1;
    for (auto &rc : rec_of) if ((L->flags[rc.first] & 3u) == 0) L->flags[rc.first] = rc.second << 2;

    if (pr) {
        for (uint64_t v = 0; v < pr->nvert; ++v) {
            const double *d = &pr->verts[6 * v];
            float row[8] = {(float)d[0], (float)d[1], (float)d[2], 0.0f, (float)d[3], (float)d[4], (float)d[5], 0.0f};
            L->verts.insert(L->verts.end(), row, row + 8);
        }
        L->slicedef = {(int32_t)pr->nslice, 0, 0, 0};
        for (uint64_t s = 0; s < pr->nslice; ++s) {
            const simlab_slice_def &d = pr->slices[s];
            int32_t r[12] = {d.origin[0], d.origin[1], d.origin[2], (int32_t)d.nu, d.u[0], d.u[1], d.u[2], (int32_t)d.nv,
                             d.v[0], d.v[1], d.v[2], 0};
            L->slicedef.insert(L->slicedef.end(), r, r + 12);
            L->npix += d.nu * d.nv;
        }
        L->voff = (float)pr->vertex_offset;
    }
    L->vacc.assign(L->verts.size() / 8 * 2, 0.0f);
    L->sacc.assign((size_t)L->npix * 5, 0.0f);

    // init (shader.rs:623-628): feq(0, U_INIT) in the parity-0 layout
    T fe[19];
    L->feq((T)0, L->u_init, fe);
    for (uint64_t c = 0; c < n; ++c) L->scatter_pre(c, 0, fe, L->seed_of(0));
    return L;
}

thread_local std::string g_err;

struct Handle {
    int store;
    void *lat;
};

template <class T> Lat<T> *as(Handle *h) { return (Lat<T> *)h->lat; }

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

// store: 0 f32, 1 f16 (this library's dither), 2 bf16, 3 f64, 4 f16 with the reference's per-store hash
void *oracle_new(const simlab_lattice_setup *su, const simlab_geometry *g, const simlab_probes *pr, int store, int nthreads) {
    std::string err;
    Handle *h = new Handle{store, nullptr};
    if (store == ST_F64) h->lat = make_lat<double>(su, g, pr, store, nthreads, err);
    else h->lat = make_lat<float>(su, g, pr, store, nthreads, err);
    if (!h->lat) { g_err = err; delete h; return nullptr; }
    return h;
}

void oracle_free(void *hp) {
    Handle *h = (Handle *)hp;
    if (!h) return;
    if (h->store == ST_F64) delete as<double>(h); else delete as<float>(h);
    delete h;
}

#define DISPATCH(h, expr)                                   \
    do {                                                    \
        if ((h)->store == ST_F64) { auto *L = as<double>(h); expr; } \
        else { auto *L = as<float>(h); expr; }              \
    } while (0)

void oracle_set_fused(void *hp, int on) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, L->fused = on != 0);
}

// Only the face update that opens the next step (face_gather + face_write of
// module p = steps_done % 2, shader.rs:630-687), without the step itself and
// without advancing steps_done: the face cells' outgoing populations as the
// main step would read them, for the face-rule tests.
void oracle_face_update(void *hp) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        const int p = (int)(L->steps_done % 2);
        const uint32_t seed = (uint32_t)L->steps_done * 2654435761u;
        for (uint32_t k = 0; k < (uint32_t)L->faces.size(); ++k) L->face_gather(k, p);
        for (uint32_t k = 0; k < (uint32_t)L->faces.size(); ++k) L->face_write(k, p, seed);
    });
}

int oracle_run(void *hp, const simlab_step_in *steps, uint64_t nsteps, double *ft) {
    Handle *h = (Handle *)hp;
    double tmp[6];
    DISPATCH(h, L->first_of_run = true);
    DISPATCH(h, for (uint64_t k = 0; k < nsteps; ++k) L->step(steps[k], ft ? ft + 6 * k : tmp));
    return 0;
}

int oracle_run_still(void *hp, uint64_t nsteps, double *ft) {
    simlab_step_in s;
    std::memset(&s, 0, sizeof s);
    for (int k = 0; k < 3; ++k) { s.rot[k][k] = 1.0; s.rot_prev[k][k] = 1.0; }
    Handle *h = (Handle *)hp;
    double tmp[6];
    DISPATCH(h, L->first_of_run = true);
    DISPATCH(h, for (uint64_t k = 0; k < nsteps; ++k) L->step(s, ft ? ft + 6 * k : tmp));
    return 0;
}

// test hook: only the re-voxelization stage of the next step (a spinning
// part's tri_scatter, voxelize and qclear), without the stream-collide
int oracle_spin_stage(void *hp, const simlab_step_in *si) {
    Handle *h = (Handle *)hp;
    int ok = 1;
    DISPATCH(h, if (L->geo.moving) L->spin_stage(*si); else ok = 0);
    return ok ? 0 : -1;
}

uint64_t oracle_steps_done(void *hp) { Handle *h = (Handle *)hp; uint64_t r = 0; DISPATCH(h, r = L->steps_done); return r; }
uint32_t oracle_nrec(void *hp) { Handle *h = (Handle *)hp; uint32_t r = 0; DISPATCH(h, r = (uint32_t)L->recs.size()); return r; }
uint32_t oracle_nface(void *hp) { Handle *h = (Handle *)hp; uint32_t r = 0; DISPATCH(h, r = (uint32_t)L->faces.size()); return r; }

// lattice.rs:818-835 macros (no debug_dump requirement in the oracle)
int oracle_macros(void *hp, float *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        int p = (int)(L->steps_done % 2);
        for (uint64_t m = 0; m < L->N; ++m) {
            decltype(L->tau0) rho; V3<decltype(L->tau0)> u;
            if (L->cell_macro(m, p, rho, u)) { out[4 * m] = (float)rho; out[4 * m + 1] = (float)u.x; out[4 * m + 2] = (float)u.y; out[4 * m + 3] = (float)u.z; }
            else for (int k = 0; k < 4; ++k) out[4 * m + k] = std::nanf("");
        }
    });
    return 0;
}

// Same in double precision (FP64 oracle readout).
int oracle_macros_f64(void *hp, double *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        int p = (int)(L->steps_done % 2);
        for (uint64_t m = 0; m < L->N; ++m) {
            decltype(L->tau0) rho; V3<decltype(L->tau0)> u;
            if (L->cell_macro(m, p, rho, u)) { out[4 * m] = rho; out[4 * m + 1] = u.x; out[4 * m + 2] = u.y; out[4 * m + 3] = u.z; }
            else for (int k = 0; k < 4; ++k) out[4 * m + k] = std::nan("");
        }
    });
    return 0;
}

// Records packed as the reference's 40 words (shader.rs:62-69).
int oracle_records(void *hp, uint32_t *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        for (size_t r = 0; r < L->recs.size(); ++r) {
            const auto &R = L->recs[r];
            uint32_t *w = out + r * RS;
            std::memset(w, 0, RS * 4);
            w[0] = R.mask; w[1] = R.cell;
            for (int k = 0; k < 9; ++k) w[2 + k] = R.q[k];
            for (int i = 0; i < 18; ++i) w[11 + i] = fbits((float)R.fpost[i]);
            for (int k = 0; k < 4; ++k) w[29 + k] = fbits((float)R.mac[k]);
            for (int k = 0; k < 6; ++k) w[33 + k] = fbits((float)R.ft[k]);
        }
    });
    return 0;
}

// Raw population slots (storage element order: direction-major, cell fastest).
int oracle_get_pops(void *hp, void *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        if (L->store == ST_F32) std::memcpy(out, L->pf.data(), L->pf.size() * 4);
        else if (L->store == ST_F64) std::memcpy(out, L->pd.data(), L->pd.size() * 8);
        else std::memcpy(out, L->ph.data(), L->ph.size() * 2);
    });
    return 0;
}
int oracle_set_pops(void *hp, const void *in) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        if (L->store == ST_F32) std::memcpy(L->pf.data(), in, L->pf.size() * 4);
        else if (L->store == ST_F64) std::memcpy(L->pd.data(), in, L->pd.size() * 8);
        else std::memcpy(L->ph.data(), in, L->ph.size() * 2);
    });
    return 0;
}
int oracle_get_flags(void *hp, uint32_t *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, std::memcpy(out, L->flags.data(), L->flags.size() * 4));
    return 0;
}
int oracle_get_faces(void *hp, uint32_t *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, std::memcpy(out, L->faces.data(), L->faces.size() * 4));
    return 0;
}

// Total mass sum(rho) over colliding cells, FP64 (cell_macro readout).
double oracle_mass(void *hp) {
    Handle *h = (Handle *)hp;
    double m = 0;
    DISPATCH(h, {
        int p = (int)(L->steps_done % 2);
        for (uint64_t c = 0; c < L->N; ++c) {
            decltype(L->tau0) rho; V3<decltype(L->tau0)> u;
            if (L->cell_macro(c, p, rho, u)) m += (double)rho;
        }
    });
    return m;
}

int oracle_vertex_means(void *hp, float *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        size_t nv = L->vacc.size() / 2;
        for (size_t i = 0; i < nv; ++i) out[i] = L->vacc[2 * i + 1] > 0 ? L->vacc[2 * i] / L->vacc[2 * i + 1] : std::nanf("");
    });
    return 0;
}
int oracle_slice_means(void *hp, float *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        for (uint32_t i = 0; i < L->npix; ++i) {
            float c = L->sacc[5 * i + 4];
            for (int k = 0; k < 4; ++k) out[4 * i + k] = c > 0 ? L->sacc[5 * i + k] / c : std::nanf("");
        }
    });
    return 0;
}
int oracle_reset_samples(void *hp) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, { std::fill(L->vacc.begin(), L->vacc.end(), 0.0f); std::fill(L->sacc.begin(), L->sacc.end(), 0.0f); });
    return 0;
}

// ---- known-answer helpers ----
float oracle_sround_ref(float x, uint32_t key, uint32_t seed) { return sround_ref(x, key, seed); }
uint16_t oracle_f16_bits_host(float x) { return f16_bits_host(x); }
uint16_t oracle_f16_rne(float x) { return float_to_half_rne(x); }
uint16_t oracle_f16_rtz(float x) { return float_to_half_rtz(x); }
float oracle_half_to_float(uint16_t h) { return half_to_float(h); }
uint32_t oracle_cell_hash(uint32_t n, uint32_t seed) { return cell_hash(n, seed); }
void oracle_dither16(uint32_t n, uint32_t seed, uint16_t *out) { dither16(n, seed, out); }
uint16_t oracle_enc_f16_sr(float x, uint16_t d) { return enc_f16_sr(x, d); }
uint16_t oracle_enc_bf16_sr(float x, uint16_t d) { return enc_bf16_sr(x, d); }

// AA slot maps (shader.rs:254-339) for a dims lattice: for each cell n and
// direction i, the slot index (dir * N + cell) read by load_in(p) and
// written by store_out(p). out_load/out_store: N x 19 uint64.
void oracle_aa_maps(const uint32_t dims[3], int p, uint64_t *out_load, uint64_t *out_store) {
    Lat<float> L;
    L.NX = dims[0]; L.NY = dims[1]; L.NZ = dims[2]; L.N = (uint64_t)dims[0] * dims[1] * dims[2];
    for (uint64_t n = 0; n < L.N; ++n) {
        uint32_t x, y, z; L.coords(n, x, y, z);
        for (int i = 0; i < 19; ++i) {
            out_load[n * 19 + i] = (p == 0 || i == 0) ? L.slot(i, n) : L.slot(opp(i), L.nb(x, y, z, -CV[i][0], -CV[i][1], -CV[i][2]));
            out_store[n * 19 + i] = (p == 0 || i == 0) ? L.slot(opp(i), n) : L.slot(i, L.nb(x, y, z, CV[i][0], CV[i][1], CV[i][2]));
        }
    }
}

// One collision of a single cell (no walls): inputs f~ (19), outputs the
// post-collision f~ before streaming, for the collision known-answer tests.
void oracle_collide_f64(const double *fin, double tau0, double cs, const double grav[3], double *fout) {
    Lat<double> L;
    L.tau0 = tau0; L.smag2 = cs * cs; L.grav = {grav[0], grav[1], grav[2]};
    L.forcing = grav[0] != 0 || grav[1] != 0 || grav[2] != 0;
    L.sponge_on = false; L.wall_model = false; L.sigma = 0; L.sponge_tau = 0;
    for (int i = 0; i < 19; ++i) L.w[i] = weight_d(i);
    // a 1x1x1 periodic lattice: gather and scatter hit the same cell
    L.NX = L.NY = L.NZ = 1; L.N = 1; L.per[0] = L.per[1] = L.per[2] = true;
    L.store = ST_F64; L.pd.assign(19, 0.0); L.flags.assign(1, 0);
    for (int i = 0; i < 19; ++i) L.pd[i] = fin[i];
    StepU s; std::memset(&s, 0, sizeof s);
    L.main_cell(0, 0, s, false);
    for (int i = 0; i < 19; ++i) fout[i] = L.pd[opp(i)];   // store_out(0) writes (n, opp i)
}

// The Spalding wall model of one wall cell (lattice units): its link mask
// and q per direction (as the record packs them, f16), the cell position
// relative to the spin centre, its velocity, the LES tau, the angular
// velocity and tau0; returns the cell's tau, for the wall-model tests.
double oracle_wall_tau_f64(uint32_t mask, const float *q19, const double pc[3], const double u[3], double tau_les,
                           const double om[3], double tau0) {
    Lat<double> L;
    L.tau0 = tau0;
    Rec<double> r;
    r.mask = mask;
    for (int k = 0; k < 9; ++k)
        r.q[k] = (uint32_t)f16_bits_host(q19[2 * k + 1]) | ((uint32_t)f16_bits_host(q19[2 * k + 2]) << 16);
    return L.wall_tau(r, mask, V3<double>{pc[0], pc[1], pc[2]}, V3<double>{u[0], u[1], u[2]}, tau_les,
                      V3<double>{om[0], om[1], om[2]});
}

// The refill's mean non-equilibrium stress over n neighbour records
// (post-collision f~ of directions 1..18, then rho, u), for its tests.
void oracle_refill_stress_f64(const double *fpost18, const double *mac4, int n, double pis_out[6]) {
    double pis[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < n; ++k) {
        Rec<double> R;
        for (int j = 0; j < 18; ++j) R.fpost[j] = fpost18[18 * k + j];
        for (int j = 0; j < 4; ++j) R.mac[j] = mac4[4 * k + j];
        Lat<double>::add_neighbour_stress(R, pis);
    }
    for (int j = 0; j < 6; ++j) pis_out[j] = n > 0 ? pis[j] / n : 0.0;
}

}  // extern "C"
