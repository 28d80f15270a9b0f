// ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this file; only tests/ and the smoke check do, and only as the
// checker.
//
// A plain CPU restatement of the reference's refined D3Q27 cumulant solver
// (offgridevices/simlab crates/cfd): the WGSL kernels generated by
// crates/cfd/src/mshader.rs and the host bookkeeping and dispatch schedule of
// crates/cfd/src/mlattice.rs, one loop iteration per GPU thread, in the same
// order of operations as the WGSL text (constants rounded to f32 as the
// generated source writes them). Templated on the arithmetic type: float
// for FP32/FP16 storage, double for the FP64 oracle run.
//
// Pinned by tests/test_oracle_refined.py: the Gaussian post-collision state
// against its central moments computed by brute force, the rebuilt
// face/ghost states against their density, momentum and stress, the shear
// wave's viscous decay (tau = 1/2 + 3 nu), the reference's own tests of
// crates/cfd/tests/refined.rs (Poiseuille second order with the force
// balance, the momentum balance round a sphere, a uniform stream crossing two
// refinement levels unchanged), and uniform states that feel no force.

#include "../include/simlab_cfd.h"
#include "half.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <array>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using namespace oracle_half;

// d3q27.rs:1-30: direction i has velocity (i%3-1, (i/3)%3-1, i/9-1); rest 13;
// opposite 26-i; weights 8/27, 2/27, 1/54, 1/216; moving-direction slot
constexpr int Q = 27, REST = 13;
constexpr int cvx(int i) { return i % 3 - 1; }
constexpr int cvy(int i) { return (i / 3) % 3 - 1; }
constexpr int cvz(int i) { return i / 9 - 1; }
constexpr int copp(int i) { return 26 - i; }
constexpr int cslot(int i) { return i < REST ? i : i - 1; }
double weight27(int i) {
    const int s = std::abs(cvx(i)) + std::abs(cvy(i)) + std::abs(cvz(i));
    return s == 0 ? 8.0 / 27.0 : s == 1 ? 2.0 / 27.0 : s == 2 ? 1.0 / 54.0 : 1.0 / 216.0;
}

// mshader.rs:33-48
constexpr uint32_t RS = 24, R_Q = 2, R_FT = 15, MW = 12, RED_CHUNK = 4096;
constexpr uint32_t T_SOLID = 1, T_FACE = 2, T_GHOST = 3, T_RESTRICT = 4;

enum Store { ST_F32 = 0, ST_F16 = 1, ST_F64 = 3 };

std::string g_err;

// a WGSL constant written by `fl` (mshader.rs:91-93): the f64 value as f32
template <class T> inline T fl(double x) { return std::is_same<T, double>::value ? (T)x : (T)(float)x; }

template <class T> struct Mom {
    T rho1;
    T j[3];
    T s[6];
};

template <class T> struct Consts {
    T W[Q];
    T W45[Q];   // f32(4.5 w_i), as fneq writes it
    T W6[Q];    // 6 * W[i] evaluated in T (a const-expression of the generated source)
    T third, two3;
    Consts() {
        for (int i = 0; i < Q; ++i) {
            W[i] = fl<T>(weight27(i));
            W45[i] = fl<T>(4.5 * weight27(i));
            W6[i] = (T)6.0 * W[i];
        }
        third = (T)(1.0 / 3.0);
        two3 = (T)(2.0 / 3.0);
    }
};

// mshader.rs:260-277 moments of a shifted population set, each a sum in
// direction order of the terms with coefficient +-1
template <class T> Mom<T> moments(const T f[Q]) {
    Mom<T> m;
    T r = 0;
    for (int i = 0; i < Q; ++i) r = r + f[i];
    m.rho1 = r;
    const int ab[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int a = 0; a < 3; ++a) {
        T s = 0;
        for (int i = 0; i < Q; ++i) {
            const int c[3] = {cvx(i), cvy(i), cvz(i)};
            if (c[a] == 1) s = s + f[i];
            else if (c[a] == -1) s = s - f[i];
        }
        m.j[a] = s;
    }
    for (int k = 0; k < 6; ++k) {
        T s = 0;
        for (int i = 0; i < Q; ++i) {
            const int c[3] = {cvx(i), cvy(i), cvz(i)};
            const int v = c[ab[k][0]] * c[ab[k][1]];
            if (v == 1) s = s + f[i];
            else if (v == -1) s = s - f[i];
        }
        m.s[k] = s;
    }
    return m;
}

// mshader.rs:279-287 product-form equilibrium, shifted
template <class T> void feq(const Consts<T> &K, T rho, const T u[3], T f[Q]) {
    T a[3][3];
    for (int k = 0; k < 3; ++k) {
        const T uu = u[k] * u[k];
        a[k][0] = (T)0.5 * ((K.third + uu) - u[k]);
        a[k][1] = K.two3 - uu;
        a[k][2] = (T)0.5 * ((K.third + uu) + u[k]);
    }
    for (int i = 0; i < Q; ++i) f[i] = ((rho * a[0][cvx(i) + 1]) * a[1][cvy(i) + 1]) * a[2][cvz(i) + 1] - K.W[i];
}

// mshader.rs:288-301 second-order Hermite non-equilibrium of a stress p
// (xx yy zz xy xz yz)
template <class T> void fneq(const Consts<T> &K, const T p[6], T f[Q]) {
    const T tr = ((p[0] + p[1]) + p[2]) * K.third;
    const int ab[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int i = 0; i < Q; ++i) {
        const int c[3] = {cvx(i), cvy(i), cvz(i)};
        T t = -tr;
        for (int k = 0; k < 6; ++k) {
            const int a = ab[k][0], b = ab[k][1];
            const int m = c[a] * c[b] * (a == b ? 1 : 2);
            if (m != 0) t = t + (T)m * p[k];
        }
        f[i] = K.W45[i] * t;
    }
}

// mshader.rs:307-375 gauss: the shifted populations whose central moments
// about u are rho times those of a Gaussian with covariance sg (xx yy zz xy
// xz yz) and first central moments m1, by the backward chimera transform
// along z, then y, then x
template <class T> void gauss(const Consts<T> &K, T rho, const T u[3], const T sg[6], const T m1[3], T f[Q]) {
    const T a = sg[0], b = sg[1], cc = sg[2], d = sg[3], e = sg[4], g = sg[5];
    T k[3][3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) k[i][j][l] = 0;
    auto set = [&](int i, int j, int l, T v) { k[i][j][l] = v; };
    const T two = 2, eight = 8;
    set(0, 0, 0, 1);
    set(1, 0, 0, m1[0]);
    set(0, 1, 0, m1[1]);
    set(0, 0, 1, m1[2]);
    set(2, 0, 0, a);
    set(0, 2, 0, b);
    set(0, 0, 2, cc);
    set(1, 1, 0, d);
    set(1, 0, 1, e);
    set(0, 1, 1, g);
    set(2, 2, 0, a * b + (two * d) * d);
    set(2, 0, 2, a * cc + (two * e) * e);
    set(0, 2, 2, b * cc + (two * g) * g);
    set(2, 1, 1, a * g + (two * d) * e);
    set(1, 2, 1, b * e + (two * d) * g);
    set(1, 1, 2, cc * d + (two * e) * g);
    set(2, 2, 2, ((a * b) * cc + two * (((a * g) * g + (b * e) * e) + (cc * d) * d)) + ((eight * d) * e) * g);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) k[i][j][l] = rho * k[i][j][l];
    // (k0, k1, k2) about w -> populations at -1, 0, +1
    auto back = [&](T k0, T k1, T k2, T w, T out[3]) {
        const T r1 = k1 + w * k0;
        const T r2 = (k2 + ((T)2.0 * w) * k1) + (w * w) * k0;
        out[0] = (T)0.5 * (r2 - r1);
        out[1] = k0 - r2;
        out[2] = (T)0.5 * (r2 + r1);
    };
    T z[3][3][3], y[3][3][3], x[3][3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) back(k[i][j][0], k[i][j][1], k[i][j][2], u[2], z[i][j]);   // z[i][j][cz]
    for (int i = 0; i < 3; ++i)
        for (int cz = 0; cz < 3; ++cz) {
            T o[3];
            back(z[i][0][cz], z[i][1][cz], z[i][2][cz], u[1], o);
            for (int cy = 0; cy < 3; ++cy) y[i][cz][cy] = o[cy];
        }
    for (int cy = 0; cy < 3; ++cy)
        for (int cz = 0; cz < 3; ++cz) {
            T o[3];
            back(y[0][cz][cy], y[1][cz][cy], y[2][cz][cy], u[0], o);
            for (int cx = 0; cx < 3; ++cx) x[cy][cz][cx] = o[cx];
        }
    for (int i = 0; i < Q; ++i) f[i] = x[cvy(i) + 1][cvz(i) + 1][cvx(i) + 1] - K.W[i];
}

// mshader.rs:466-485 stress and deviator
template <class T> void stress(const Consts<T> &K, const Mom<T> &mo, T rho, const T u[3], T p[6]) {
    const T r3 = mo.rho1 / (T)3.0;
    p[0] = (mo.s[0] - (rho * u[0]) * u[0]) - r3;
    p[1] = (mo.s[1] - (rho * u[1]) * u[1]) - r3;
    p[2] = (mo.s[2] - (rho * u[2]) * u[2]) - r3;
    p[3] = mo.s[3] - (rho * u[0]) * u[1];
    p[4] = mo.s[4] - (rho * u[0]) * u[2];
    p[5] = mo.s[5] - (rho * u[1]) * u[2];
    (void)K;
}
template <class T> void deviator(const T p[6], T d[6]) {
    for (int k = 0; k < 6; ++k) d[k] = p[k];
    const T tr = ((p[0] + p[1]) + p[2]) / (T)3.0;
    d[0] -= tr;
    d[1] -= tr;
    d[2] -= tr;
}
// mshader.rs:486-493 a post-collision state from macroscopic data
template <class T> void rebuild(const Consts<T> &K, T rho, const T u[3], const T p_post[6], T f[Q]) {
    T fe[Q], fn[Q], dv[6];
    feq(K, rho, u, fe);
    deviator(p_post, dv);
    fneq(K, dv, fn);
    for (int i = 0; i < Q; ++i) f[i] = fe[i] + fn[i];
}

struct LevelCfg {
    uint32_t dims[3];
    uint64_t n;
    bool per[3];
    double tau0, smag2, sponge_lo[3], sponge_hi[3], sponge_tau, sigma;
    bool absorb, total_p;
    double u_inf[3], u_init[3], grav[3], center[3];
    bool has_child, has_parent;
    int32_t ch_lo[3];
    uint32_t ch_n[3];
    uint64_t ch_nodes;
    uint32_t pn[3];
    uint64_t pnn;
    uint32_t nface = 0, nghost = 0, nrestrict = 0;
    uint32_t npix = 0;
};

template <class T> struct Level {
    LevelCfg c;
    // per-level constants as the generated source holds them (f32 for T = float)
    T TAU0, SMAG2, SPONGE_TAU, SIGMA, SP_LO[3], SP_HI[3], U_INF[3], U_INIT[3], GRAV[3], CENTER[3];
    bool n_sponge = false, forcing = false;
    std::vector<T> val[2];          // FP32/FP64: [set][i * n + cell]
    std::vector<uint16_t> h16[2];   // FP16 bits
    std::vector<uint32_t> flags, lists, spix;
    std::vector<T> mom_child, mom_up;
    uint32_t src = 0, mslot = 1;
    uint64_t steps = 0;
};

template <class T> struct MLat {
    Consts<T> K;
    int store;
    int nthreads;
    simlab_multi_setup setup;
    std::vector<Level<T>> lv;
    std::vector<uint32_t> recs;    // nrec x RS
    std::vector<double> rft;       // nrec x 6: each record's F and T in the arithmetic type
    uint32_t nrec = 0, nvert = 0, npix = 0;
    std::vector<float> probes;     // nvert x 8
    std::vector<float> sacc;       // npix x 5
    double vertex_offset = 0;
    uint64_t fine_steps = 0;
    std::vector<std::array<double, 6>> hist;
    uint32_t hist_cap = 1024;

    // ----- storage (mshader.rs:182-199)
    inline T ld(const Level<T> &L, uint32_t set, int i, uint64_t cell) const {
        const size_t k = (size_t)i * L.c.n + cell;
        if (store == ST_F16) return (T)(half_to_float(L.h16[set][k]) * (1.0f / 32768.0f));
        return L.val[set][k];
    }
    inline void st(Level<T> &L, uint32_t set, int i, uint64_t cell, T v, uint32_t seed) const {
        const size_t k = (size_t)i * L.c.n + cell;
        if (store == ST_F16) {
            const uint32_t key = (uint32_t)cell * 27u + (uint32_t)i;
            L.h16[set][k] = float_to_half_rne(sround_ref((float)v * 32768.0f, key, seed));
        } else {
            L.val[set][k] = v;
        }
    }
    void own(const Level<T> &L, uint32_t set, uint64_t n, T f[Q]) const {
        for (int i = 0; i < Q; ++i) f[i] = ld(L, set, i, n);
    }
    void put(Level<T> &L, uint32_t set, uint64_t n, const T f[Q], uint32_t seed) const {
        for (int i = 0; i < Q; ++i) st(L, set, i, n, f[i], seed);
    }

    static void coords(const LevelCfg &c, uint64_t n, uint32_t xyz[3]) {
        const uint64_t nxy = (uint64_t)c.dims[0] * c.dims[1];
        xyz[2] = (uint32_t)(n / nxy);
        const uint64_t r = n - (uint64_t)xyz[2] * nxy;
        xyz[1] = (uint32_t)(r / c.dims[0]);
        xyz[0] = (uint32_t)(r - (uint64_t)xyz[1] * c.dims[0]);
    }
    static uint64_t cidx(const LevelCfg &c, const uint32_t x[3]) {
        return ((uint64_t)x[2] * c.dims[1] + x[1]) * c.dims[0] + x[0];
    }

    // mshader.rs:436-450 (sponge ramp)
    T sponge(const Level<T> &L, const uint32_t xyz[3]) const {
        T s = 0;
        for (int k = 0; k < 3; ++k) {
            const T p = (T)xyz[k];
            const T hi = (T)(L.c.dims[k] - 1) - p;
            if (L.SP_LO[k] > 0) s = std::max(s, (T)1.0 - p / L.SP_LO[k]);
            if (L.SP_HI[k] > 0) s = std::max(s, (T)1.0 - hi / L.SP_HI[k]);
        }
        return std::min(std::max(s, (T)0), (T)1);
    }

    // ----- init (mshader.rs:508-514)
    void init(uint32_t seed) {
        for (auto &L : lv) {
            T fe[Q];
            feq(K, (T)1, L.U_INIT, fe);
            for (uint64_t n = 0; n < L.c.n; ++n) {
                put(L, 0, n, fe, seed);
                put(L, 1, n, fe, seed);
            }
        }
    }

    // ----- step (mshader.rs:636-728 STEP, 376-414 walls)
    void step(int l, uint32_t src, uint32_t mw, uint32_t seed, const simlab_step_in &sin) {
        Level<T> &L = lv[l];
        const LevelCfg &c = L.c;
        const uint32_t dst = 1 - src;
        const T om3[3] = {(T)sin.omega[0], (T)sin.omega[1], (T)sin.omega[2]};
        const int64_t N = (int64_t)c.n;
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int64_t nn = 0; nn < N; ++nn) {
            const uint64_t n = (uint64_t)nn;
            const uint32_t fg = L.flags[n];
            if ((fg & 7u) != 0u) continue;
            uint32_t xyz[3];
            coords(c, n, xyz);
            // pull (mshader.rs:219-227): the population moving along c_i comes from n - c_i
            T f[Q];
            for (int i = 0; i < Q; ++i) {
                if (i == REST) {
                    f[i] = ld(L, src, i, n);
                    continue;
                }
                uint32_t s[3];
                const int cc[3] = {cvx(i), cvy(i), cvz(i)};
                for (int k = 0; k < 3; ++k) {
                    const int64_t v = (int64_t)xyz[k] - cc[k];
                    s[k] = (uint32_t)(v < 0 ? (int64_t)c.dims[k] - 1 : (v >= (int64_t)c.dims[k] ? 0 : v));
                }
                f[i] = ld(L, src, i, cidx(c, s));
            }
            const uint32_t rec = fg >> 3;
            T F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
            size_t base = 0;
            if (rec != 0u) {
                base = (size_t)(rec - 1) * RS;
                const uint32_t mask = recs[base];
                T fin[Q];
                std::memcpy(fin, f, sizeof fin);
                const T pc[3] = {(T)xyz[0] - L.CENTER[0], (T)xyz[1] - L.CENTER[1], (T)xyz[2] - L.CENTER[2]};
                for (int i = 0; i < Q; ++i) {
                    if (i == REST || !(mask & (1u << i))) continue;
                    const int o = copp(i);
                    const int sl = cslot(i);
                    const uint32_t word = recs[base + R_Q + sl / 2];
                    const T q = std::max((T)half_to_float((uint16_t)(sl % 2 ? word >> 16 : word & 0xffff)), (T)0.001);
                    const T w = K.W[i];
                    const T fpi = ld(L, src, i, n) + w;
                    const T fpo = ld(L, src, o, n) + w;
                    const T ci[3] = {(T)cvx(i), (T)cvy(i), (T)cvz(i)};
                    const T xw[3] = {pc[0] + q * ci[0], pc[1] + q * ci[1], pc[2] + q * ci[2]};
                    const T uw[3] = {om3[1] * xw[2] - om3[2] * xw[1], om3[2] * xw[0] - om3[0] * xw[2],
                                     om3[0] * xw[1] - om3[1] * xw[0]};
                    const T cu = (ci[0] * uw[0] + ci[1] * uw[1]) + ci[2] * uw[2];
                    T fb;
                    if (q >= (T)0.5) fb = ((fpi + ((T)2.0 * q - (T)1.0) * fpo) - K.W6[i] * cu) / ((T)2.0 * q);
                    else if (!(mask & (1u << o)))
                        fb = (((T)2.0 * q) * fpi + ((T)1.0 - (T)2.0 * q) * (fin[i] + w)) - K.W6[i] * cu;
                    else fb = fpi - K.W6[i] * cu;
                    f[o] = fb - w;
                    T flk[3];
                    for (int k = 0; k < 3; ++k) flk[k] = (ci[k] - uw[k]) * fpi + (ci[k] + uw[k]) * fb;
                    for (int k = 0; k < 3; ++k) F[k] += flk[k];
                    Tq[0] += xw[1] * flk[2] - xw[2] * flk[1];
                    Tq[1] += xw[2] * flk[0] - xw[0] * flk[2];
                    Tq[2] += xw[0] * flk[1] - xw[1] * flk[0];
                }
            }
            const Mom<T> mo = moments(f);
            const T rho = (T)1.0 + mo.rho1;
            const T G[3] = {rho * L.GRAV[0], rho * L.GRAV[1], rho * L.GRAV[2]};
            T u[3];
            for (int k = 0; k < 3; ++k) u[k] = (mo.j[k] + (T)0.5 * G[k]) / rho;
            T p[6];
            stress(K, mo, rho, u, p);
            T tau = L.TAU0;
            if (L.SMAG2 > 0) {
                T dv[6];
                deviator(p, dv);
                const T pp = ((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]) +
                             (T)2.0 * ((dv[3] * dv[3] + dv[4] * dv[4]) + dv[5] * dv[5]);
                tau = (T)0.5 * (L.TAU0 + std::sqrt(L.TAU0 * L.TAU0 + (((T)18.0 * L.SMAG2) * std::sqrt((T)2.0 * pp)) / rho));
            }
            if (L.n_sponge) {
                const T sp = sponge(L, xyz);
                if (sp > 0 && L.SPONGE_TAU > tau) tau += (sp * sp) * (L.SPONGE_TAU - tau);
            }
            const T om = (T)1.0 / tau;
            T dv[6];
            deviator(p, dv);
            T sg[6];
            for (int k = 0; k < 3; ++k) sg[k] = K.third + (((T)1.0 - om) * dv[k]) / rho;
            for (int k = 3; k < 6; ++k) sg[k] = (((T)1.0 - om) * dv[k]) / rho;
            T m1[3] = {0, 0, 0};
            if (L.forcing)
                for (int k = 0; k < 3; ++k) m1[k] = ((T)0.5 * G[k]) / rho;
            T fo[Q];
            gauss(K, rho, u, sg, m1, fo);
            if (L.n_sponge && L.SIGMA > 0) {
                const T sp = sponge(L, xyz);
                if (sp > 0) {
                    const T *ut = c.absorb ? L.U_INF : u;
                    T ft[Q], fc[Q];
                    feq(K, (T)1.0, ut, ft);
                    feq(K, rho, u, fc);
                    const T sgm = (L.SIGMA * sp) * sp;
                    for (int i = 0; i < Q; ++i) fo[i] += sgm * (ft[i] - fc[i]);
                }
            }
            if (rec != 0u) {
                for (int k = 0; k < 3; ++k) {
                    float a = (float)F[k], b = (float)Tq[k];
                    std::memcpy(&recs[base + R_FT + k], &a, 4);
                    std::memcpy(&recs[base + R_FT + 3 + k], &b, 4);
                    rft[(size_t)(rec - 1) * 6 + k] = (double)F[k];
                    rft[(size_t)(rec - 1) * 6 + 3 + k] = (double)Tq[k];
                }
            }
            for (int i = 0; i < Q; ++i) st(L, dst, i, n, fo[i], seed);
            auto put_mom = [&](T *o) {
                o[0] = rho;
                o[1] = u[0];
                o[2] = u[1];
                o[3] = u[2];
                for (int k = 0; k < 6; ++k) o[4 + k] = p[k];
                o[10] = tau;
            };
            if (c.has_child) {
                int64_t lc[3];
                bool in = true, surf = false;
                for (int k = 0; k < 3; ++k) {
                    lc[k] = (int64_t)xyz[k] - c.ch_lo[k];
                    if (lc[k] < 0 || lc[k] >= (int64_t)c.ch_n[k]) in = false;
                    if (lc[k] == 0 || lc[k] == (int64_t)c.ch_n[k] - 1) surf = true;
                }
                if (in && surf) {
                    const size_t o = ((size_t)mw * c.ch_nodes + ((size_t)lc[2] * c.ch_n[1] + lc[1]) * c.ch_n[0] + lc[0]) * MW;
                    put_mom(&L.mom_child[o]);
                }
            }
            if (c.has_parent && !(xyz[0] & 1u) && !(xyz[1] & 1u) && !(xyz[2] & 1u)) {
                const uint32_t h[3] = {xyz[0] / 2, xyz[1] / 2, xyz[2] / 2};
                const size_t o = (((size_t)h[2] * c.pn[1] + h[1]) * c.pn[0] + h[0]) * MW;
                put_mom(&L.mom_up[o]);
            }
        }
    }

    // ----- face_bc (mshader.rs:732-762)
    void face(uint32_t src, uint32_t seed) {
        Level<T> &L = lv[0];
        const LevelCfg &c = L.c;
        for (uint32_t k = 0; k < c.nface; ++k) {
            const uint32_t b = L.lists[k];
            uint32_t xyz[3];
            coords(c, b, xyz);
            for (int a = 0; a < 3; ++a)
                if (!c.per[a]) xyz[a] = std::min(std::max(xyz[a], 1u), c.dims[a] - 2);
            const uint64_t m = cidx(c, xyz);
            T rho = 1, u[3] = {L.U_INF[0], L.U_INF[1], L.U_INF[2]}, pp[6] = {0, 0, 0, 0, 0, 0};
            if ((L.flags[m] & 7u) == 0u) {
                T f[Q];
                own(L, src, m, f);
                const Mom<T> mo = moments(f);
                rho = (T)1.0 + mo.rho1;
                for (int a = 0; a < 3; ++a) u[a] = mo.j[a] / rho;
                stress(K, mo, rho, u, pp);
            }
            const uint32_t ty = (L.flags[b] >> 3) & 3u;
            if (ty == 1u) {
                if (c.total_p) {
                    const T uinf2 = (L.U_INF[0] * L.U_INF[0] + L.U_INF[1] * L.U_INF[1]) + L.U_INF[2] * L.U_INF[2];
                    const T u2 = (u[0] * u[0] + u[1] * u[1]) + u[2] * u[2];
                    rho = (K.third + (T)0.5 * uinf2) / (K.third + (T)0.5 * u2);
                } else {
                    for (int a = 0; a < 3; ++a) u[a] = L.U_INF[a];
                }
            } else {
                rho = 1;
            }
            T f[Q];
            rebuild(K, rho, u, pp, f);
            put(L, src, b, f, seed);
        }
    }

    // ----- ghost_fill (mshader.rs:771-801) on level l from its parent's moments
    void ghost(int l, uint32_t src, uint32_t mold, uint32_t mnew, T blend, uint32_t seed) {
        Level<T> &L = lv[l];
        const LevelCfg &c = L.c;
        const Level<T> &P = lv[l - 1];
        for (uint32_t k = 0; k < c.nghost; ++k) {
            const uint32_t n = L.lists[c.nface + k];
            uint32_t xyz[3];
            coords(c, n, xyz);
            const uint32_t h0[3] = {xyz[0] / 2, xyz[1] / 2, xyz[2] / 2};
            const uint32_t odd[3] = {xyz[0] & 1u, xyz[1] & 1u, xyz[2] & 1u};
            T acc[11];
            for (T &v : acc) v = 0;
            T wsum = 0;
            for (uint32_t cc = 0; cc < 8; ++cc) {
                const uint32_t o[3] = {cc & 1u, (cc >> 1) & 1u, (cc >> 2) & 1u};
                bool skip = false;
                for (int a = 0; a < 3; ++a)
                    if (o[a] == 1u && odd[a] == 0u) skip = true;
                if (skip) continue;
                const uint32_t node[3] = {h0[0] + o[0], h0[1] + o[1], h0[2] + o[2]};
                const T w = (odd[0] == 1u ? (T)0.5 : (T)1.0) * (odd[1] == 1u ? (T)0.5 : (T)1.0) * (odd[2] == 1u ? (T)0.5 : (T)1.0);
                const size_t nodei = ((size_t)node[2] * c.pn[1] + node[1]) * c.pn[0] + node[0];
                const T *a = &P.mom_child[((size_t)mold * c.pnn + nodei) * MW];
                const T *b = &P.mom_child[((size_t)mnew * c.pnn + nodei) * MW];
                for (int j = 0; j < 11; ++j) acc[j] += w * (((T)1.0 - blend) * a[j] + blend * b[j]);
                wsum += w;
            }
            const T rho = acc[0] / wsum;
            const T u[3] = {acc[1] / wsum, acc[2] / wsum, acc[3] / wsum};
            const T tc = acc[10] / wsum;
            const T tf = (T)0.5 + (T)2.0 * (tc - (T)0.5);
            const T sc = tf / ((T)2.0 * tc) * ((T)1.0 - (T)1.0 / tf);
            T p[6];
            for (int j = 0; j < 6; ++j) p[j] = acc[4 + j] / wsum * sc;
            T f[Q];
            rebuild(K, rho, u, p, f);
            put(L, src, n, f, seed);
        }
    }

    // ----- restrict_fill (mshader.rs:803-820) on level l from its child's moments
    void restrict_fill(int l, uint32_t src, uint32_t seed) {
        Level<T> &L = lv[l];
        const LevelCfg &c = L.c;
        const Level<T> &Ch = lv[l + 1];
        for (uint32_t k = 0; k < c.nrestrict; ++k) {
            const uint32_t n = L.lists[c.nface + c.nghost + k];
            uint32_t xyz[3];
            coords(c, n, xyz);
            const uint32_t lc[3] = {(uint32_t)((int64_t)xyz[0] - c.ch_lo[0]), (uint32_t)((int64_t)xyz[1] - c.ch_lo[1]),
                                    (uint32_t)((int64_t)xyz[2] - c.ch_lo[2])};
            const T *m = &Ch.mom_up[(((size_t)lc[2] * c.ch_n[1] + lc[1]) * c.ch_n[0] + lc[0]) * MW];
            const T rho = m[0];
            const T u[3] = {m[1], m[2], m[3]};
            const T tf = m[10];
            const T tc = (T)0.5 + (T)0.5 * (tf - (T)0.5);
            const T sc = (T)2.0 * tc / tf * ((T)1.0 - (T)1.0 / tc);
            T p[6];
            for (int j = 0; j < 6; ++j) p[j] = m[4 + j] * sc;
            T f[Q];
            rebuild(K, rho, u, p, f);
            put(L, src, n, f, seed);
        }
    }

    // ----- macros of a cell (mshader.rs:495-506)
    bool cell_macro(const Level<T> &L, uint32_t set, uint64_t m, T &rho, T u[3]) const {
        if ((L.flags[m] & 7u) == 1u) return false;
        T f[Q];
        own(L, set, m, f);
        const Mom<T> mo = moments(f);
        rho = (T)1.0 + mo.rho1;
        for (int a = 0; a < 3; ++a) u[a] = mo.j[a] / rho - (T)0.5 * L.GRAV[a];
        return true;
    }

    // ----- samplers (mshader.rs:571-611)
    void sample_verts(uint32_t set, const simlab_step_in &sin) {
        const Level<T> &L = lv.back();
        const LevelCfg &c = L.c;
        for (uint32_t v = 0; v < nvert; ++v) {
            const float *pr = &probes[8 * (size_t)v];
            T b[3];
            for (int k = 0; k < 3; ++k) b[k] = (T)pr[k] + (T)(float)vertex_offset * (T)pr[3 + k];
            T s[3];
            for (int k = 0; k < 3; ++k)
                s[k] = L.CENTER[k] + (((T)(float)sin.rot[k][0] * b[0] + (T)(float)sin.rot[k][1] * b[1]) +
                                      (T)(float)sin.rot[k][2] * b[2]);
            T b0[3], fr[3];
            for (int k = 0; k < 3; ++k) {
                b0[k] = std::floor(s[k]);
                fr[k] = s[k] - b0[k];
            }
            T acc = 0, wsum = 0;
            for (uint32_t kk = 0; kk < 8; ++kk) {
                const uint32_t o[3] = {kk & 1u, (kk >> 1) & 1u, (kk >> 2) & 1u};
                int64_t cc[3];
                bool out = false;
                for (int k = 0; k < 3; ++k) {
                    cc[k] = (int64_t)(int32_t)b0[k] + o[k];
                    if (cc[k] < 1 || cc[k] >= (int64_t)c.dims[k] - 1) out = true;
                }
                if (out) continue;
                T w = 1;
                for (int k = 0; k < 3; ++k) w = w * (o[k] == 1u ? fr[k] : (T)1.0 - fr[k]);
                const uint32_t ci[3] = {(uint32_t)cc[0], (uint32_t)cc[1], (uint32_t)cc[2]};
                T rho, u[3];
                if (cell_macro(L, set, cidx(c, ci), rho, u) && w > 0) {
                    acc += w * rho;
                    wsum += w;
                }
            }
            if (wsum > (T)1e-6) {
                probes[8 * (size_t)v + 6] += (float)(acc / wsum);
                probes[8 * (size_t)v + 7] += 1.0f;
            }
        }
    }
    void sample_slices(int l, uint32_t set) {
        const Level<T> &L = lv[l];
        for (uint32_t k = 0; k < L.c.npix; ++k) {
            const uint32_t px = L.spix[2 * k], cell = L.spix[2 * k + 1];
            T rho, u[3];
            if (cell_macro(L, set, cell, rho, u)) {
                float *a = &sacc[5 * (size_t)px];
                a[0] += (float)rho;
                a[1] += (float)u[0];
                a[2] += (float)u[1];
                a[3] += (float)u[2];
                a[4] += 1.0f;
            }
        }
    }

    // forces of a finest step: the records' F and T, summed in FP64 in
    // record order (the reference sums the f32 words in f32, mshader.rs:524-569)
    std::array<double, 6> reduce() const {
        std::array<double, 6> s{0, 0, 0, 0, 0, 0};
        for (uint32_t r = 0; r < nrec; ++r)
            for (int k = 0; k < 6; ++k) s[k] += rft[(size_t)r * 6 + k];
        return s;
    }

    // mlattice.rs:690-705 seed of one dispatch
    uint32_t seed(uint32_t hslot, uint32_t src, uint32_t mw) const {
        return ((uint32_t)fine_steps + hslot) * 2654435761u + (src * 7919u + mw * 104729u);
    }

    // mlattice.rs:786-842 the recursive schedule of one coarse step of level l
    void schedule(int l, const simlab_step_in *fine, uint64_t &fine_i) {
        const int nl = (int)lv.size(), finest = nl - 1;
        simlab_step_in still{};
        for (int k = 0; k < 3; ++k) still.rot[k][k] = still.rot_prev[k][k] = 1.0;
        const simlab_step_in &sin = l == finest ? fine[fine_i] : still;
        const uint32_t lsrc = lv[l].src;
        const uint32_t lmw = 1 - lv[l].mslot;
        const uint32_t hslot = (uint32_t)(fine_i % hist_cap);
        const uint32_t sd = seed(hslot, lsrc, lmw);
        if (l == 0 && lv[0].c.nface > 0) face(lsrc, sd);
        step(l, lsrc, lmw, sd, sin);
        lv[l].src = 1 - lsrc;
        lv[l].mslot = lmw;
        lv[l].steps += 1;
        if (l == finest) {
            hist.push_back(nrec > 0 ? reduce() : std::array<double, 6>{0, 0, 0, 0, 0, 0});
            if (sin.sample) {
                if (nvert > 0) sample_verts(1 - lsrc, sin);
                for (int k = 0; k < nl; ++k)
                    if (lv[k].c.npix > 0) sample_slices(k, lv[k].src);
            }
            fine_i += 1;
        }
        if (l + 1 < nl) {
            const uint32_t mold = 1 - lmw;
            for (int sub = 0; sub < 2; ++sub) {
                const int c = l + 1;
                const uint32_t csrc = lv[c].src;
                ghost(c, csrc, mold, lmw, (T)(0.5f * (float)sub), seed(0, csrc, 0));
                schedule(c, fine, fine_i);
            }
            restrict_fill(l, lv[l].src, seed(0, lv[l].src, 0));
        }
    }

    // mlattice.rs:725-784
    bool run(const simlab_step_in *fine, uint64_t nsteps, double *ft) {
        const uint64_t ss = (uint64_t)1 << (lv.size() - 1);
        if (nsteps % ss != 0) {
            g_err = "step count must be a multiple of " + std::to_string(ss);
            return false;
        }
        const uint64_t max_coarse = std::max<uint64_t>(hist_cap / ss, 1);
        uint64_t done = 0;
        while (done < nsteps) {
            const uint64_t nc = std::min((nsteps - done) / ss, max_coarse);
            hist.clear();
            uint64_t fine_i = 0;
            for (uint64_t k = 0; k < nc; ++k) schedule(0, fine + done, fine_i);
            for (uint64_t k = 0; k < fine_i; ++k)
                if (ft)
                    for (int j = 0; j < 6; ++j) ft[(done + k) * 6 + j] = hist[k][j];
            fine_steps += fine_i;
            done += fine_i;
        }
        return true;
    }
};

// mlattice.rs:165-181 rec_words
void rec_words(uint32_t cell, uint32_t mask, const float q[27], uint32_t *w) {
    for (uint32_t k = 0; k < RS; ++k) w[k] = 0;
    w[0] = mask;
    w[1] = cell;
    float qs[26];
    for (float &v : qs) v = 1.0f;
    for (int i = 0; i < Q; ++i)
        if (i != REST) qs[cslot(i)] = q[i];
    for (int k = 0; k < 13; ++k)
        w[R_Q + k] = (uint32_t)f16_bits_host(qs[2 * k]) | ((uint32_t)f16_bits_host(qs[2 * k + 1]) << 16);
}

uint32_t level_dim(const simlab_multi_setup &s, int l, int k) {
    return l == 0 ? s.dims0[k] : 2 * (s.boxes[l - 1].n[k] - 1) + 1;
}

// mlattice.rs:185-690 (the parts that decide what the kernels compute)
template <class T> MLat<T> *make(const simlab_multi_setup *su, const simlab_geometry *g, const simlab_probes *pr, int store,
                                 int nthreads, std::string &err) {
    auto *M = new MLat<T>();
    M->store = store;
    M->nthreads = nthreads > 0 ? nthreads : 1;
    M->setup = *su;
    const simlab_multi_setup &s = *su;
    const int nl = (int)s.nboxes + 1;
    const int finest = nl - 1;
    const bool periodic0[3] = {s.faces[0] == SIMLAB_FACE_PERIODIC, s.faces[2] == SIMLAB_FACE_PERIODIC,
                               s.faces[4] == SIMLAB_FACE_PERIODIC};
    if (nl > 1 && (periodic0[0] || periodic0[1] || periodic0[2])) {
        err = "periodic faces only work without refinement";
        delete M;
        return nullptr;
    }
    for (int l = 1; l < nl; ++l) {
        const simlab_level_box &b = s.boxes[l - 1];
        for (int k = 0; k < 3; ++k) {
            const int64_t hi = (int64_t)b.lo[k] + b.n[k] - 1;
            const int margin = l == 1 ? 2 : 3;
            const int64_t pd = level_dim(s, l - 1, k);
            if (b.lo[k] < margin || hi > pd - 1 - margin || b.n[k] < 8) {
                err = "refinement box " + std::to_string(l) + " does not fit inside its parent";
                delete M;
                return nullptr;
            }
        }
    }
    // records (finest level)
    std::vector<std::pair<uint32_t, uint32_t>> rec_of;
    std::vector<uint32_t> solid;
    if (g && g->kind == SIMLAB_GEOM_STATIC) {
        M->nrec = (uint32_t)g->nwall;
        M->recs.assign((size_t)M->nrec * RS, 0);
        M->rft.assign((size_t)M->nrec * 6, 0.0);
        for (uint64_t r = 0; r < g->nwall; ++r) {
            rec_words(g->walls[r].cell, g->walls[r].mask, g->walls[r].q, &M->recs[r * RS]);
            rec_of.push_back({g->walls[r].cell, (uint32_t)r + 1});
        }
        solid.assign(g->solid, g->solid + g->nsolid);
    }
    std::sort(rec_of.begin(), rec_of.end());
    std::sort(solid.begin(), solid.end());
    // probes
    if (pr && pr->nvert) {
        M->nvert = (uint32_t)pr->nvert;
        M->probes.assign((size_t)M->nvert * 8, 0.0f);
        for (uint64_t v = 0; v < pr->nvert; ++v)
            for (int k = 0; k < 6; ++k) M->probes[8 * v + k] = (float)pr->verts[6 * v + k];
    }
    M->vertex_offset = pr ? pr->vertex_offset : 0.0;
    // origins and scales of the levels in finest node coordinates
    std::vector<std::array<double, 3>> org(nl, {0, 0, 0});
    std::vector<double> scale(nl, 1.0);
    scale[0] = (double)((uint64_t)1 << finest);
    for (int l = 1; l < nl; ++l) {
        scale[l] = scale[l - 1] / 2.0;
        for (int k = 0; k < 3; ++k) org[l][k] = org[l - 1][k] + (double)s.boxes[l - 1].lo[k] * scale[l - 1];
    }
    // slice pixels: each from the finest level whose interior holds it
    std::vector<std::vector<uint32_t>> pix(nl);
    uint32_t pbase = 0;
    for (uint64_t si = 0; pr && si < pr->nslice; ++si) {
        const simlab_slice_def &sd = pr->slices[si];
        for (uint32_t iv = 0; iv < sd.nv; ++iv)
            for (uint32_t iu = 0; iu < sd.nu; ++iu) {
                double p[3];
                for (int k = 0; k < 3; ++k)
                    p[k] = (double)(int32_t)(sd.origin[k] + (int32_t)iu * sd.u[k] + (int32_t)iv * sd.v[k]);
                for (int l = nl - 1; l >= 0; --l) {
                    bool ok = true;
                    uint32_t cell[3];
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
    M->sacc.assign((size_t)M->npix * 5, 0.0f);
    M->lv.resize(nl);
    for (int l = 0; l < nl; ++l) {
        Level<T> &L = M->lv[l];
        LevelCfg &c = L.c;
        for (int k = 0; k < 3; ++k) c.dims[k] = level_dim(s, l, k);
        const uint32_t nx = c.dims[0], ny = c.dims[1], nz = c.dims[2];
        c.n = (uint64_t)nx * ny * nz;
        if (c.n >= 0xffffffffull) {
            err = "level " + std::to_string(l) + " has too many cells";
            delete M;
            return nullptr;
        }
        c.has_child = l + 1 < nl;
        c.has_parent = l > 0;
        if (c.has_child)
            for (int k = 0; k < 3; ++k) {
                c.ch_lo[k] = s.boxes[l].lo[k];
                c.ch_n[k] = s.boxes[l].n[k];
            }
        else
            for (int k = 0; k < 3; ++k) {
                c.ch_lo[k] = 0;
                c.ch_n[k] = 1;
            }
        c.ch_nodes = (uint64_t)c.ch_n[0] * c.ch_n[1] * c.ch_n[2];
        for (int k = 0; k < 3; ++k) c.pn[k] = (c.dims[k] - 1) / 2 + 1;
        c.pnn = (uint64_t)c.pn[0] * c.pn[1] * c.pn[2];
        std::vector<uint32_t> faces, ghosts, restricted;
        L.flags.assign(c.n, 0);
        for (uint64_t cc = 0; cc < c.n; ++cc) {
            const uint32_t x = (uint32_t)(cc % nx), y = (uint32_t)((cc / nx) % ny), z = (uint32_t)(cc / ((uint64_t)nx * ny));
            int ty = 0;
            if (l == 0) {
                const bool on[6] = {x == 0, x == nx - 1, y == 0, y == ny - 1, z == 0, z == nz - 1};
                for (int f = 0; f < 6; ++f)
                    if (on[f] && s.faces[f] != SIMLAB_FACE_PERIODIC) {
                        const int t = s.faces[f] == SIMLAB_FACE_VELOCITY ? 1 : 2;
                        ty = ty == 1 ? 1 : t;
                    }
            }
            if (ty) {
                L.flags[cc] = T_FACE | ((uint32_t)ty << 3);
                faces.push_back((uint32_t)cc);
            } else if (l > 0 && (x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1)) {
                L.flags[cc] = T_GHOST;
                ghosts.push_back((uint32_t)cc);
            } else if (c.has_child) {
                const int32_t p[3] = {(int32_t)x, (int32_t)y, (int32_t)z};
                bool in = true;
                for (int k = 0; k < 3; ++k)
                    if (!(p[k] >= c.ch_lo[k] + 2 && p[k] <= c.ch_lo[k] + (int32_t)c.ch_n[k] - 3)) in = false;
                if (in) {
                    L.flags[cc] = T_RESTRICT;
                    restricted.push_back((uint32_t)cc);
                }
            }
        }
        if (l == finest) {
            for (uint32_t sc : solid)
                if (sc < c.n && L.flags[sc] == 0) L.flags[sc] = T_SOLID;
            for (auto &rf : rec_of)
                if (rf.first < c.n && ((L.flags[rf.first] & 7u) == 0u || (L.flags[rf.first] & 7u) == T_SOLID))
                    L.flags[rf.first] = (L.flags[rf.first] & 7u) | (rf.second << 3);
        }
        c.nface = (uint32_t)faces.size();
        c.nghost = (uint32_t)ghosts.size();
        c.nrestrict = (uint32_t)restricted.size();
        L.lists = faces;
        L.lists.insert(L.lists.end(), ghosts.begin(), ghosts.end());
        L.lists.insert(L.lists.end(), restricted.begin(), restricted.end());
        L.spix = pix[l];
        c.npix = (uint32_t)(pix[l].size() / 2);
        const double tau_l = 0.5 + 3.0 * s.nu0 * (double)((uint64_t)1 << l);
        auto init_m = [&](uint64_t nodes, std::vector<T> &v) {
            v.assign(nodes * MW, (T)0);
            for (uint64_t k = 0; k < nodes; ++k) {
                v[k * MW + 0] = (T)1.0;
                for (int a = 0; a < 3; ++a) v[k * MW + 1 + a] = (T)(float)s.u_init[a];
                v[k * MW + 10] = (T)(float)tau_l;
                if (std::is_same<T, double>::value) {
                    for (int a = 0; a < 3; ++a) v[k * MW + 1 + a] = (T)s.u_init[a];
                    v[k * MW + 10] = (T)tau_l;
                }
            }
        };
        init_m(c.has_child ? 2 * c.ch_nodes : 0, L.mom_child);
        init_m(l > 0 ? c.pnn : 0, L.mom_up);
        for (int k = 0; k < 3; ++k) c.per[k] = l == 0 ? periodic0[k] : false;
        const double scale_l = (double)((uint64_t)1 << (finest - l));
        for (int k = 0; k < 3; ++k) {
            c.sponge_lo[k] = l == 0 ? (double)s.sponge[2 * k] : 0.0;
            c.sponge_hi[k] = l == 0 ? (double)s.sponge[2 * k + 1] : 0.0;
            c.u_inf[k] = s.u_inf[k];
            c.u_init[k] = s.u_init[k];
            c.grav[k] = s.body_force0[k] / (double)((uint64_t)1 << l);
            c.center[k] = l == finest ? s.center[k] : (s.center[k] - org[l][k]) / scale_l;
        }
        c.tau0 = tau_l;
        c.smag2 = s.smagorinsky * s.smagorinsky;
        c.sponge_tau = l == 0 ? s.sponge_tau : 0.0;
        c.sigma = l == 0 ? s.sponge_sigma : 0.0;
        c.absorb = s.absorb_velocity != 0;
        c.total_p = s.total_pressure_inlet != 0;
        L.TAU0 = fl<T>(c.tau0);
        L.SMAG2 = fl<T>(c.smag2);
        L.SPONGE_TAU = fl<T>(c.sponge_tau);
        L.SIGMA = fl<T>(c.sigma);
        for (int k = 0; k < 3; ++k) {
            L.SP_LO[k] = fl<T>(c.sponge_lo[k]);
            L.SP_HI[k] = fl<T>(c.sponge_hi[k]);
            L.U_INF[k] = fl<T>(c.u_inf[k]);
            L.U_INIT[k] = fl<T>(c.u_init[k]);
            L.GRAV[k] = fl<T>(c.grav[k]);
            L.CENTER[k] = fl<T>(c.center[k]);
            if (c.sponge_lo[k] > 0 || c.sponge_hi[k] > 0) L.n_sponge = true;
            if (c.grav[k] != 0.0) L.forcing = true;
        }
        if (store == ST_F16) {
            L.h16[0].assign((size_t)Q * c.n, 0);
            L.h16[1].assign((size_t)Q * c.n, 0);
        } else {
            L.val[0].assign((size_t)Q * c.n, (T)0);
            L.val[1].assign((size_t)Q * c.n, (T)0);
        }
    }
    M->init(0u);
    return M;
}

struct Handle {
    int store;
    void *lat;
};
template <class T> MLat<T> *as(Handle *h) { return (MLat<T> *)h->lat; }

#define DISPATCH(h, ...)                                                    \
    do {                                                                    \
        if ((h)->store == ST_F64) { auto *L = as<double>(h); __VA_ARGS__; } \
        else { auto *L = as<float>(h); __VA_ARGS__; }                       \
    } while (0)

}  // namespace

extern "C" {

const char *moracle_last_error(void) { return g_err.c_str(); }

// store: 0 f32, 1 f16 (the reference's per-store rounding), 3 f64
void *moracle_new(const simlab_multi_setup *su, const simlab_geometry *g, const simlab_probes *pr, int store,
                  int nthreads) {
    std::string err;
    Handle *h = new Handle{store, nullptr};
    if (store == ST_F64) h->lat = make<double>(su, g, pr, store, nthreads, err);
    else h->lat = make<float>(su, g, pr, store, nthreads, err);
    if (!h->lat) {
        g_err = err;
        delete h;
        return nullptr;
    }
    return h;
}

void moracle_free(void *hp) {
    Handle *h = (Handle *)hp;
    if (!h) return;
    if (h->store == ST_F64) delete as<double>(h);
    else delete as<float>(h);
    delete h;
}

int moracle_run(void *hp, const simlab_step_in *steps, uint64_t n, double *ft) {
    Handle *h = (Handle *)hp;
    int ok = 0;
    DISPATCH(h, ok = L->run(steps, n, ft));
    return ok ? 0 : 1;
}

int moracle_run_still(void *hp, uint64_t n, double *ft) {
    std::vector<simlab_step_in> st(n);
    for (auto &s : st) {
        std::memset(&s, 0, sizeof s);
        for (int k = 0; k < 3; ++k) s.rot[k][k] = s.rot_prev[k][k] = 1.0;
    }
    return moracle_run(hp, st.data(), n, ft);
}

uint32_t moracle_levels(void *hp) {
    Handle *h = (Handle *)hp;
    uint32_t v = 0;
    DISPATCH(h, v = (uint32_t)L->lv.size());
    return v;
}

void moracle_level_dims(void *hp, uint32_t *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, for (size_t l = 0; l < L->lv.size(); ++l) for (int k = 0; k < 3; ++k) out[3 * l + k] = L->lv[l].c.dims[k]);
}

uint32_t moracle_nrec(void *hp) {
    Handle *h = (Handle *)hp;
    uint32_t v = 0;
    DISPATCH(h, v = L->nrec);
    return v;
}
uint32_t moracle_npix(void *hp) {
    Handle *h = (Handle *)hp;
    uint32_t v = 0;
    DISPATCH(h, v = L->npix);
    return v;
}

// (rho, u) of every cell of level l from its source set; NaN where not a
// stepped fluid cell (mshader.rs:613-622 dump_macro)
void moracle_macros(void *hp, uint32_t l, float *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        const auto &Lv = L->lv[l];
        for (uint64_t n = 0; n < Lv.c.n; ++n) {
            decltype(L->K.third) rho, u[3];
            const bool ok = L->cell_macro(Lv, Lv.src, n, rho, u);
            if (ok && (Lv.flags[n] & 7u) == 0u) {
                out[4 * n] = (float)rho;
                for (int k = 0; k < 3; ++k) out[4 * n + 1 + k] = (float)u[k];
            } else {
                for (int k = 0; k < 4; ++k) out[4 * n + k] = NAN;
            }
        }
    });
}

void moracle_records(void *hp, uint32_t *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, std::memcpy(out, L->recs.data(), L->recs.size() * 4));
}

void moracle_get_pops(void *hp, uint32_t l, double *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        const auto &Lv = L->lv[l];
        for (int i = 0; i < Q; ++i)
            for (uint64_t n = 0; n < Lv.c.n; ++n) out[(size_t)i * Lv.c.n + n] = (double)L->ld(Lv, Lv.src, i, n);
    });
}

// sets the populations the next step of level l reads (deviations f - w,
// direction i of cell n at i * cells + n), for initial states of the tests
void moracle_set_pops(void *hp, uint32_t l, const double *in) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        auto &Lv = L->lv[l];
        for (int i = 0; i < Q; ++i)
            for (uint64_t n = 0; n < Lv.c.n; ++n)
                L->st(Lv, Lv.src, i, n, (decltype(L->K.third))in[(size_t)i * Lv.c.n + n], 0u);
    });
}

// the flags of level l (cells x u32)
void moracle_flags(void *hp, uint32_t l, uint32_t *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, std::memcpy(out, L->lv[l].flags.data(), L->lv[l].flags.size() * 4));
}

void moracle_reset_samples(void *hp) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, {
        for (uint32_t v = 0; v < L->nvert; ++v) L->probes[8 * v + 6] = L->probes[8 * v + 7] = 0.0f;
        std::fill(L->sacc.begin(), L->sacc.end(), 0.0f);
    });
}
void moracle_vertex_means(void *hp, float *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, for (uint32_t v = 0; v < L->nvert; ++v) {
        const float c = L->probes[8 * v + 7];
        out[v] = c > 0 ? L->probes[8 * v + 6] / c : NAN;
    });
}
void moracle_slice_means(void *hp, float *out) {
    Handle *h = (Handle *)hp;
    DISPATCH(h, for (uint32_t p = 0; p < L->npix; ++p) {
        const float c = L->sacc[5 * p + 4];
        for (int k = 0; k < 4; ++k) out[4 * p + k] = c > 0 ? L->sacc[5 * p + k] / c : NAN;
    });
}

// pieces of the collision in FP64, for the closed-form pins
void moracle_gauss_f64(double rho, const double *u, const double *sg, const double *m1, double *out) {
    static const Consts<double> K;
    gauss(K, rho, u, sg, m1, out);
}
void moracle_feq_f64(double rho, const double *u, double *out) {
    static const Consts<double> K;
    feq(K, rho, u, out);
}
void moracle_rebuild_f64(double rho, const double *u, const double *p, double *out) {
    static const Consts<double> K;
    rebuild(K, rho, u, p, out);
}
// one collision of a cell with pulled populations f (shifted), as the step
// does for a fluid cell without walls, forcing or sponge
void moracle_collide_f64(const double *f, double tau0, double smag2, double *out) {
    static const Consts<double> K;
    const Mom<double> mo = moments(f);
    const double rho = 1.0 + mo.rho1;
    double u[3], p[6], dv[6], sg[6];
    for (int k = 0; k < 3; ++k) u[k] = mo.j[k] / rho;
    stress(K, mo, rho, u, p);
    double tau = tau0;
    deviator(p, dv);
    if (smag2 > 0) {
        const double pp = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2] + 2.0 * (dv[3] * dv[3] + dv[4] * dv[4] + dv[5] * dv[5]);
        tau = 0.5 * (tau0 + std::sqrt(tau0 * tau0 + 18.0 * smag2 * std::sqrt(2.0 * pp) / rho));
    }
    const double om = 1.0 / tau;
    for (int k = 0; k < 3; ++k) sg[k] = 1.0 / 3.0 + (1.0 - om) * dv[k] / rho;
    for (int k = 3; k < 6; ++k) sg[k] = (1.0 - om) * dv[k] / rho;
    const double m1[3] = {0, 0, 0};
    gauss(K, rho, u, sg, m1, out);
}

}  // extern "C"
