// gfx950 kernels of the D3Q19 wind-tunnel lattice.
//
// Semantics follow the reference's WGSL (crates/cfd/src/shader.rs); the
// data layout and launch structure are this library's own:
//  * populations: SoA, slot (i, n) = element i * ncell + n, stored as
//    deviations f~ = f - w_i (FP32 / FP64 / scaled FP16 / BF16);
//  * AA pattern in place (shader.rs:5-13): a step of parity 0 reads (n, i)
//    and writes (n, opp i); parity 1 reads (n - c_i, opp i) and writes
//    (n + c_i, i). Every slot is touched by one thread per step;
//  * one thread per cell, a block = whole x-rows (bx >= nx lanes, by rows),
//    so the +-1 x-offset accesses of parity 1 stay inside one block (one CU
//    and one XCD L2) and every wave reads/writes whole cache lines;
//  * cells with a wall record are taken out of the bulk blocks and processed
//    by "wall blocks" (scheduled first in the same launch) that walk the
//    record list, so bulk waves never diverge on wall logic. The wall blocks
//    also reduce the momentum-exchange force and torque in FP64 (wave
//    shuffles -> LDS -> per-block partial -> last block sums the partials in
//    block order, deterministic), replacing reduce1/reduce2 (shader.rs:689-734).
//  * FP16/BF16 stores round stochastically: one hash per cell, xorshift32
//    words give 16 dither bits per stored direction, added below the kept
//    mantissa bits before a round-toward-zero conversion.

#include "kernels.h"

#include <cstdint>
#include <type_traits>
#include <utility>

namespace simlab {
namespace {

// ---------------------------------------------------------------- velocity set
// shader.rs:22-60 (same order): 0 rest, 1-6 axes, 7-18 diagonals in +/- pairs
constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

template <class T> __device__ __forceinline__ constexpr T wgt(int i) {
    return i == 0 ? (T)(1.0 / 3.0) : (i <= 6 ? (T)(1.0 / 18.0) : (T)(1.0 / 36.0));
}

template <int... Is, class F>
__device__ __forceinline__ void sfor_impl(std::integer_sequence<int, Is...>, F &&f) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void sfor(F &&f) {
    sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

// ---------------------------------------------------------------- arithmetic
__device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_(double x) { return 1.0 / x; }
__device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float div_(float a, float b) { return a / b; }
__device__ __forceinline__ double div_(double a, double b) { return a / b; }

// ---------------------------------------------------------------- storage codecs
struct CF32 { using E = float; using T = float; static constexpr bool SR = false; };
struct CF16 { using E = uint16_t; using T = float; static constexpr bool SR = true; };
struct CBF16 { using E = uint16_t; using T = float; static constexpr bool SR = true; };
struct CF64 { using E = double; using T = double; static constexpr bool SR = false; };

constexpr float FS = 32768.0f;   // shader.rs:73 FP16_SCALE
constexpr float INV_S = 1.0f / 32768.0f;

typedef __fp16 half2_t __attribute__((ext_vector_type(2)));

template <class C> __device__ __forceinline__ typename C::T dec(typename C::E e) {
    if constexpr (std::is_same_v<C, CF16>) {
        return (float)__builtin_bit_cast(_Float16, e) * INV_S;
    } else if constexpr (std::is_same_v<C, CBF16>) {
        return __uint_as_float((uint32_t)e << 16);
    } else {
        return e;
    }
}

// d16: 16 dither bits (ignored by the plain storages)
template <class C> __device__ __forceinline__ typename C::E enc(typename C::T v, uint32_t d16) {
    if constexpr (std::is_same_v<C, CF16>) {
        float s = __builtin_amdgcn_fmed3f(v * FS, -65000.0f, 65000.0f);
        float t = __uint_as_float(__float_as_uint(s) + (d16 >> 3));
        half2_t h = __builtin_amdgcn_cvt_pkrtz(t, t);
        return __builtin_bit_cast(uint16_t, h[0]);
    } else if constexpr (std::is_same_v<C, CBF16>) {
        return (uint16_t)((__float_as_uint(v) + d16) >> 16);
    } else {
        (void)d16;
        return v;
    }
}

// per-cell dither words: the reference's sround mixing (shader.rs:511-512)
// applied once to the cell index, then xorshift32 words
__device__ __forceinline__ uint32_t cell_hash(uint32_t n, uint32_t seed) {
    uint32_t h = (n * 0x9E3779B1u) ^ (seed * 0x85EBCA77u + 0x632BE5ABu);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}
struct Dither {
    uint32_t w[10];
    __device__ __forceinline__ void init(uint32_t n, uint32_t seed) {
        uint32_t x = cell_hash(n, seed);
#pragma unroll
        for (int k = 0; k < 10; ++k) {
            w[k] = x;
            x ^= x << 13;
            x ^= x >> 17;
            x ^= x << 5;
        }
    }
    template <int J> __device__ __forceinline__ uint32_t d16() const {
        return (J & 1) ? (w[J >> 1] >> 16) : (w[J >> 1] & 0xffffu);
    }
};

template <class C> __device__ __forceinline__ typename C::E *popdir(const DevLattice &L, int i) {
    return (typename C::E *)L.pops + (size_t)i * L.ncell;
}

// ---------------------------------------------------------------- neighbours
// Offsets to the +-1 neighbour along each axis, wrapping at the domain edge
// (shader.rs:247-252: the generated kernels always wrap).
struct Nbr {
    uint32_t n, x, y, z;
    int32_t xp, xm, yp, ym, zp, zm;
    __device__ __forceinline__ void make(const DevLattice &L, uint32_t x_, uint32_t y_, uint32_t z_, uint32_t n_) {
        n = n_; x = x_; y = y_; z = z_;
        const int32_t nx = (int32_t)L.nx, nxy = (int32_t)(L.nx * L.ny);
        xp = (x == L.nx - 1) ? -(nx - 1) : 1;
        xm = (x == 0) ? (nx - 1) : -1;
        yp = (y == L.ny - 1) ? -(int32_t)((L.ny - 1) * L.nx) : nx;
        ym = (y == 0) ? (int32_t)((L.ny - 1) * L.nx) : -nx;
        zp = (z == L.nz - 1) ? -(int32_t)((L.nz - 1) * L.nx * L.ny) : nxy;
        zm = (z == 0) ? (int32_t)((L.nz - 1) * L.nx * L.ny) : -nxy;
    }
    template <int cx, int cy, int cz> __device__ __forceinline__ uint32_t at() const {
        int32_t o = 0;
        if constexpr (cx > 0) o += xp;
        if constexpr (cx < 0) o += xm;
        if constexpr (cy > 0) o += yp;
        if constexpr (cy < 0) o += ym;
        if constexpr (cz > 0) o += zp;
        if constexpr (cz < 0) o += zm;
        return n + (uint32_t)o;
    }
};

__device__ __forceinline__ void coords(const DevLattice &L, uint32_t n, uint32_t &x, uint32_t &y, uint32_t &z) {
    const uint32_t nxy = L.nx * L.ny;
    z = n / nxy;
    const uint32_t r = n - z * nxy;
    y = r / L.nx;
    x = r - y * L.nx;
}

// ---------------------------------------------------------------- AA access
// Named by the parity Q of the step whose inputs (pre) or outputs (post) are
// meant: load_in(p) = gather_pre<p>, store_out(p) = scatter_post<p>,
// load_prev_out(module p) = gather_post<1-p>, store_prev_out(module p) =
// scatter_post<1-p, guarded>, store_in(p) = scatter_pre<p> (shader.rs:254-339).
template <class C, int Q> __device__ __forceinline__ void gather_pre(const DevLattice &L, const Nbr &nb, typename C::T f[19]) {
    sfor<19>([&](auto I) {
        constexpr int i = I;
        if constexpr (Q == 0 || i == 0) {
            f[i] = dec<C>(popdir<C>(L, i)[nb.n]);
        } else {
            f[i] = dec<C>(popdir<C>(L, OPP[i])[nb.template at<-CX[i], -CY[i], -CZ[i]>()]);
        }
    });
}

template <class C, int Q> __device__ __forceinline__ void gather_post(const DevLattice &L, const Nbr &nb, typename C::T f[19]) {
    sfor<19>([&](auto I) {
        constexpr int i = I;
        if constexpr (Q == 0 || i == 0) {
            f[i] = dec<C>(popdir<C>(L, OPP[i])[nb.n]);
        } else {
            f[i] = dec<C>(popdir<C>(L, i)[nb.template at<CX[i], CY[i], CZ[i]>()]);
        }
    });
}

// guard (shader.rs:287-297): skip targets that leave a non-periodic domain
template <int i> __device__ __forceinline__ bool keeps_inside(const DevLattice &L, const Nbr &nb) {
    bool ok = true;
    if constexpr (CX[i] == 1) ok = ok && (L.per[0] || nb.x < L.nx - 1);
    if constexpr (CX[i] == -1) ok = ok && (L.per[0] || nb.x > 0);
    if constexpr (CY[i] == 1) ok = ok && (L.per[1] || nb.y < L.ny - 1);
    if constexpr (CY[i] == -1) ok = ok && (L.per[1] || nb.y > 0);
    if constexpr (CZ[i] == 1) ok = ok && (L.per[2] || nb.z < L.nz - 1);
    if constexpr (CZ[i] == -1) ok = ok && (L.per[2] || nb.z > 0);
    return ok;
}

template <class C, int Q, bool GUARD>
__device__ __forceinline__ void scatter_post(const DevLattice &L, const Nbr &nb, const typename C::T f[19], uint32_t seed) {
    Dither d;
    if constexpr (C::SR) d.init(nb.n, seed);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        if constexpr (Q == 0 || i == 0) {
            constexpr int j = OPP[i];   // slot direction
            popdir<C>(L, j)[nb.n] = enc<C>(f[i], C::SR ? d.template d16<j>() : 0u);
        } else {
            bool ok = true;
            if constexpr (GUARD) ok = keeps_inside<i>(L, nb);
            if (ok) popdir<C>(L, i)[nb.template at<CX[i], CY[i], CZ[i]>()] = enc<C>(f[i], C::SR ? d.template d16<i>() : 0u);
        }
    });
}

template <class C> __device__ __forceinline__ void scatter_pre0(const DevLattice &L, uint32_t n, const typename C::T f[19], uint32_t seed) {
    Dither d;
    if constexpr (C::SR) d.init(n, seed);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        popdir<C>(L, i)[n] = enc<C>(f[i], C::SR ? d.template d16<i>() : 0u);
    });
}

// ---------------------------------------------------------------- moments
// shader.rs:341-365, summed over +-pairs (a = f_i + f_opp, b = f_i - f_opp)
template <class T> struct Mom {
    T rho1, jx, jy, jz, sxx, syy, szz, sxy, sxz, syz;
};
template <class T> __device__ __forceinline__ Mom<T> moments(const T f[19]) {
    const T a1 = f[1] + f[2], b1 = f[1] - f[2];
    const T a3 = f[3] + f[4], b3 = f[3] - f[4];
    const T a5 = f[5] + f[6], b5 = f[5] - f[6];
    const T a7 = f[7] + f[8], b7 = f[7] - f[8];
    const T a9 = f[9] + f[10], b9 = f[9] - f[10];
    const T a11 = f[11] + f[12], b11 = f[11] - f[12];
    const T a13 = f[13] + f[14], b13 = f[13] - f[14];
    const T a15 = f[15] + f[16], b15 = f[15] - f[16];
    const T a17 = f[17] + f[18], b17 = f[17] - f[18];
    Mom<T> m;
    m.sxx = a1 + a7 + a9 + a13 + a15;
    m.syy = a3 + a7 + a11 + a13 + a17;
    m.szz = a5 + a9 + a11 + a15 + a17;
    m.rho1 = f[0] + (a1 + a3 + a5) + (a7 + a9 + a11) + (a13 + a15 + a17);
    m.jx = b1 + b7 + b9 + b13 + b15;
    m.jy = b3 + b7 + b11 - b13 + b17;
    m.jz = b5 + b9 + b11 - b15 - b17;
    m.sxy = a7 - a13;
    m.sxz = a9 - a15;
    m.syz = a11 - a17;
    return m;
}

// c_i . v for direction i
template <int i, class T> __device__ __forceinline__ T cdot(T vx, T vy, T vz) {
    T r = (T)0;
    if constexpr (CX[i] == 1) r = r + vx;
    if constexpr (CX[i] == -1) r = r - vx;
    if constexpr (CY[i] == 1) r = r + vy;
    if constexpr (CY[i] == -1) r = r - vy;
    if constexpr (CZ[i] == 1) r = r + vz;
    if constexpr (CZ[i] == -1) r = r - vz;
    return r;
}

// sum_ab c_a c_b P_ab (off-diagonal twice) for direction i
template <int i, class T> __device__ __forceinline__ T cpc(const T p[6]) {
    T r = (T)0;
    if constexpr (CX[i] != 0) r = r + p[0];
    if constexpr (CY[i] != 0) r = r + p[1];
    if constexpr (CZ[i] != 0) r = r + p[2];
    if constexpr (CX[i] * CY[i] == 1) r = r + (T)2 * p[3];
    if constexpr (CX[i] * CY[i] == -1) r = r - (T)2 * p[3];
    if constexpr (CX[i] * CZ[i] == 1) r = r + (T)2 * p[4];
    if constexpr (CX[i] * CZ[i] == -1) r = r - (T)2 * p[4];
    if constexpr (CY[i] * CZ[i] == 1) r = r + (T)2 * p[5];
    if constexpr (CY[i] * CZ[i] == -1) r = r - (T)2 * p[5];
    return r;
}

// feq_i - w_i (shader.rs:367-377)
template <class T> __device__ __forceinline__ void feq_arr(T rho1, T ux, T uy, T uz, T fe[19]) {
    const T rho = (T)1 + rho1;
    const T usq = (T)1.5 * (ux * ux + uy * uy + uz * uz);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        const T cu = cdot<i>(ux, uy, uz);
        fe[i] = wgt<T>(i) * (rho1 + rho * ((T)3 * cu + (T)4.5 * cu * cu - usq));
    });
}
// shader.rs:379-396
template <class T> __device__ __forceinline__ void fneq_arr(const T p[6], T fq[19]) {
    const T tr = (p[0] + p[1] + p[2]) * ((T)1 / (T)3);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        fq[i] = (T)4.5 * wgt<T>(i) * (cpc<i>(p) - tr);
    });
}

template <class T> __device__ __forceinline__ T sponge_at(const DevLattice &L, uint32_t x, uint32_t y, uint32_t z) {
    // shader.rs:528-538
    const T p[3] = {(T)x, (T)y, (T)z};
    const T hi[3] = {(T)(L.nx - 1) - p[0], (T)(L.ny - 1) - p[1], (T)(L.nz - 1) - p[2]};
    T s = (T)0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T lo = (T)L.sp_lo[k], hw = (T)L.sp_hi[k];
        if (lo > (T)0) s = fmax(s, (T)1 - p[k] / lo);
        if (hw > (T)0) s = fmax(s, (T)1 - hi[k] / hw);
    }
    return fmin(fmax(s, (T)0), (T)1);
}

// ---------------------------------------------------------------- collision
// shader.rs:815-847 (+ Guo forcing :442-455). On exit f holds the
// post-collision f~ (before streaming), rho and u_out as the record saves.
template <class T, bool FORCE, bool SPONGE, class Hook>
__device__ __forceinline__ void collide(const DevLattice &L, uint32_t x, uint32_t y, uint32_t z, T f[19], T &rho_out,
                                        T u_out[3], Hook &&tau_hook) {
    const Mom<T> m = moments(f);
    const T rho = (T)1 + m.rho1;
    const T irho = rcp_(rho);
    T ux = m.jx * irho, uy = m.jy * irho, uz = m.jz * irho;
    const T gx = (T)L.grav[0], gy = (T)L.grav[1], gz = (T)L.grav[2];
    if constexpr (FORCE) {
        ux = ux + (T)0.5 * gx;
        uy = uy + (T)0.5 * gy;
        uz = uz + (T)0.5 * gz;
    }
    const T r3 = m.rho1 * ((T)1 / (T)3);
    T p[6];
    p[0] = m.sxx - rho * ux * ux - r3;
    p[1] = m.syy - rho * uy * uy - r3;
    p[2] = m.szz - rho * uz * uz - r3;
    p[3] = m.sxy - rho * ux * uy;
    p[4] = m.sxz - rho * ux * uz;
    p[5] = m.syz - rho * uy * uz;
    // Smagorinsky from |Pi_neq| (closed form)
    const T pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + (T)2 * (p[3] * p[3] + p[4] * p[4] + p[5] * p[5]);
    const T qn = sqrt_((T)2 * pp);
    const T tau0 = (T)L.tau0;
    T tau = (T)0.5 * (tau0 + sqrt_(tau0 * tau0 + (T)18 * (T)L.smag2 * qn * irho));
    tau = tau_hook(ux, uy, uz, tau);
    T sp = (T)0;
    if constexpr (SPONGE) {
        sp = sponge_at<T>(L, x, y, z);
        const T stau = (T)L.sponge_tau;
        if (sp > (T)0 && stau > tau) tau = tau + sp * sp * (stau - tau);
    }
    const T om = rcp_(tau);
    const T k1 = (T)1 - om;
    if constexpr (SPONGE) {
        T fe[19], fq[19];
        feq_arr<T>(m.rho1, ux, uy, uz, fe);
        fneq_arr<T>(p, fq);
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = fe[i] + k1 * fq[i];
        if (sp > (T)0 && (T)L.sigma > (T)0) {
            const T tx = L.abs_u ? (T)L.u_inf[0] : ux, ty = L.abs_u ? (T)L.u_inf[1] : uy, tz = L.abs_u ? (T)L.u_inf[2] : uz;
            T ft[19];
            feq_arr<T>((T)0, tx, ty, tz, ft);
            const T sg = (T)L.sigma * sp * sp;
#pragma unroll
            for (int i = 0; i < 19; ++i) f[i] = f[i] + sg * (ft[i] - fe[i]);
        }
    } else {
        // feq + (1-om) fneq, evaluated per +-pair: the even part is shared
        const T usq = (T)1.5 * (ux * ux + uy * uy + uz * uz);
        const T tr = (p[0] + p[1] + p[2]) * ((T)1 / (T)3);
        f[0] = wgt<T>(0) * (m.rho1 - rho * usq) + k1 * ((T)4.5 * wgt<T>(0)) * (-tr);
        sfor<9>([&](auto K) {
            constexpr int i = 2 * K + 1;
            constexpr int o = i + 1;
            const T cu = cdot<i>(ux, uy, uz);
            const T ev = wgt<T>(i) * (m.rho1 + rho * ((T)4.5 * cu * cu - usq)) + k1 * ((T)4.5 * wgt<T>(i)) * (cpc<i>(p) - tr);
            const T od = ((T)3 * wgt<T>(i)) * rho * cu;
            f[i] = ev + od;
            f[o] = ev - od;
        });
    }
    if constexpr (FORCE) {
        const T Gx = rho * gx, Gy = rho * gy, Gz = rho * gz;
        const T uG = ux * Gx + uy * Gy + uz * Gz;
        const T fo = (T)1 - (T)0.5 * om;
        sfor<19>([&](auto I) {
            constexpr int i = I;
            const T cu = cdot<i>(ux, uy, uz);
            const T cg = cdot<i>(Gx, Gy, Gz);
            f[i] = f[i] + wgt<T>(i) * (fo * ((T)3 * (cg - uG) + (T)9 * cu * cg) - (T)1.5 * ((T)1 - om) * cg);
        });
    }
    rho_out = rho;
    u_out[0] = ux;
    u_out[1] = uy;
    u_out[2] = uz;
}

struct NoHook {
    template <class T> __device__ __forceinline__ T operator()(T, T, T, T tau) const { return tau; }
};

__device__ __forceinline__ float q_of(const DevLattice &L, uint32_t r, int i) {
    const uint32_t w = L.rq[(size_t)((i - 1) >> 1) * L.nrec + r];
    const uint16_t h = (uint16_t)(((i - 1) & 1) ? (w >> 16) : (w & 0xffffu));
    return (float)__builtin_bit_cast(_Float16, h);
}

// shader.rs:546-597 Spalding wall model
template <class T>
__device__ T wall_tau(const DevLattice &L, uint32_t r, uint32_t mask, T pcx, T pcy, T pcz, T ux, T uy, T uz, T tau_les,
                      const float om[3]) {
    T nx = 0, ny = 0, nz = 0;
    for (int i = 1; i < 19; ++i) {
        if (mask & (1u << i)) {
            const T q = (T)q_of(L, r, i);
            const T wv = (i < 7 ? (T)1 / (T)18 : (T)1 / (T)36) * ((T)1.5 - q);
            nx = nx + (T)CX[i] * wv;
            ny = ny + (T)CY[i] * wv;
            nz = nz + (T)CZ[i] * wv;
        }
    }
    const T nl = sqrt_(nx * nx + ny * ny + nz * nz);
    if (nl < (T)1e-6) return tau_les;
    nx = nx / nl; ny = ny / nl; nz = nz / nl;
    T y = 2;
    for (int i = 1; i < 19; ++i) {
        if (mask & (1u << i)) {
            const T cn = (T)CX[i] * nx + (T)CY[i] * ny + (T)CZ[i] * nz;
            if (cn > (T)0.3) y = fmin(y, fmax((T)q_of(L, r, i), (T)0.02) * cn);
        }
    }
    if (y > (T)1.9) return tau_les;
    const T wx = pcx + y * nx, wy = pcy + y * ny, wz = pcz + y * nz;
    const T ox = (T)om[0], oy = (T)om[1], oz = (T)om[2];
    const T uwx = oy * wz - oz * wy, uwy = oz * wx - ox * wz, uwz = ox * wy - oy * wx;
    const T rx = ux - uwx, ry = uy - uwy, rz = uz - uwz;
    const T rn = rx * nx + ry * ny + rz * nz;
    const T tx = rx - rn * nx, ty = ry - rn * ny, tz = rz - rn * nz;
    const T ut = sqrt_(tx * tx + ty * ty + tz * tz);
    const T nu0 = ((T)L.tau0 - (T)0.5) / (T)3;
    if (ut < (T)1e-7) return tau_les;
    const T kap = (T)0.41;
    const T eb = exp(-kap * (T)5.2);
    T ust = sqrt_(nu0 * ut / y);
    for (int it = 0; it < 12; ++it) {
        const T up = ut / ust;
        const T yp = y * ust / nu0;
        const T ku = fmin(kap * up, (T)30);
        const T ek = exp(ku);
        const T g = up + eb * (ek - (T)1 - ku - (T)0.5 * ku * ku - ku * ku * ku / (T)6) - yp;
        const T dgdup = (T)1 + eb * kap * (ek - (T)1 - ku - (T)0.5 * ku * ku);
        const T dg = dgdup * (-up / ust) - y / nu0;
        const T step = g / dg;
        ust = fmin(fmax(ust - step, (T)0.1 * ust), (T)10 * ust);
        if (fabs(step) < (T)1e-5 * ust) break;
    }
    const T nuw = ust * ust * y / ut;
    return fmax((T)0.5 + (T)3 * fmax(nuw, nu0), tau_les);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of 6 doubles (fixed order); result valid in thread 0.
__device__ __forceinline__ void block_sum6(double v[6], double (*sh)[6], uint32_t tid, uint32_t bs) {
#pragma unroll
    for (int j = 0; j < 6; ++j) v[j] = wave_sum(v[j]);
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j) sh[wave][j] = v[j];
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t nw = (bs + 63) / 64;
        for (int j = 0; j < 6; ++j) v[j] = 0.0;
        for (uint32_t w = 0; w < nw; ++w)
            for (int j = 0; j < 6; ++j) v[j] += sh[w][j];
    }
}

// ---------------------------------------------------------------- wall blocks
// One thread per wall record: Bouzidi / moving-wall links with momentum
// exchange (shader.rs:401-436, 808-814), the collision, the record save
// (:848-860), then the FP64 force/torque reduction of all records.
template <class C, int P, bool FORCE, bool SPONGE>
__device__ __forceinline__ void wall_path(const DevLattice &L, const StepArgs &a, uint32_t wb, uint32_t tid, uint32_t bs) {
    using T = typename C::T;
    __shared__ double sh[16][6];
    __shared__ int last;
    if (wb >= L.wall_blocks) return;   // whole block
    const uint32_t r = wb * bs + tid;
    T F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
    T *rfpost = (T *)L.rfpost;
    T *rmac = (T *)L.rmac;
    T *rft = (T *)L.rft;
    const size_t NR = L.nrec;
    if (r < L.nrec) {
        const uint32_t n = L.rcell[r];
        const uint32_t fg = L.flags[n];
        if ((fg & 3u) == 0 && (fg >> 2) == r + 1) {
            uint32_t x, y, z;
            coords(L, n, x, y, z);
            Nbr nb;
            nb.make(L, x, y, z, n);
            T f[19];
            gather_pre<C, P>(L, nb, f);
            const uint32_t mask = L.rmask[r];
            const T pcx = (T)x - (T)L.center[0], pcy = (T)y - (T)L.center[1], pcz = (T)z - (T)L.center[2];
            const T omx = (T)a.omega[0], omy = (T)a.omega[1], omz = (T)a.omega[2];
            if (mask) {
                T fin[19];
#pragma unroll
                for (int i = 0; i < 19; ++i) fin[i] = f[i];
                sfor<18>([&](auto I) {
                    constexpr int i = I + 1;
                    constexpr int o = OPP[i];
                    if (mask & (1u << i)) {
                        const T wi = wgt<T>(i);
                        const T q = fmax((T)q_of(L, r, i), (T)0.001);
                        const T fpi = rfpost[(size_t)(i - 1) * NR + r] + wi;
                        const T fpo = rfpost[(size_t)(o - 1) * NR + r] + wi;
                        const T cx = (T)CX[i], cy = (T)CY[i], cz = (T)CZ[i];
                        const T xwx = pcx + q * cx, xwy = pcy + q * cy, xwz = pcz + q * cz;
                        const T uwx = omy * xwz - omz * xwy, uwy = omz * xwx - omx * xwz, uwz = omx * xwy - omy * xwx;
                        const T cu = cx * uwx + cy * uwy + cz * uwz;
                        T fb;
                        if (q >= (T)0.5) fb = div_(fpi + ((T)2 * q - (T)1) * fpo - (T)6 * wi * cu, (T)2 * q);
                        else if (!(mask & (1u << o))) fb = (T)2 * q * fpi + ((T)1 - (T)2 * q) * (fin[i] + wi) - (T)6 * wi * cu;
                        else fb = fpi - (T)6 * wi * cu;
                        f[o] = fb - wi;
                        const T flx = (cx - uwx) * fpi + (cx + uwx) * fb;
                        const T fly = (cy - uwy) * fpi + (cy + uwy) * fb;
                        const T flz = (cz - uwz) * fpi + (cz + uwz) * fb;
                        F[0] = F[0] + flx;
                        F[1] = F[1] + fly;
                        F[2] = F[2] + flz;
                        Tq[0] = Tq[0] + (xwy * flz - xwz * fly);
                        Tq[1] = Tq[1] + (xwz * flx - xwx * flz);
                        Tq[2] = Tq[2] + (xwx * fly - xwy * flx);
                    }
                });
            }
            T rho, u[3];
            const float omf[3] = {a.omega[0], a.omega[1], a.omega[2]};
            auto hook = [&](T ux, T uy, T uz, T tau) -> T {
                if (L.wall_model && mask != 0) return wall_tau<T>(L, r, mask, pcx, pcy, pcz, ux, uy, uz, tau, omf);
                return tau;
            };
            collide<T, FORCE, SPONGE>(L, x, y, z, f, rho, u, hook);
#pragma unroll
            for (int i = 1; i < 19; ++i) rfpost[(size_t)(i - 1) * NR + r] = f[i];
            rmac[r] = rho;
            rmac[NR + r] = u[0];
            rmac[2 * NR + r] = u[1];
            rmac[3 * NR + r] = u[2];
            if (L.moving) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    F[k] = F[k] + rft[k * NR + r];
                    Tq[k] = Tq[k] + rft[(3 + k) * NR + r];
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                rft[k * NR + r] = F[k];
                rft[(3 + k) * NR + r] = Tq[k];
            }
            scatter_post<C, P, false>(L, nb, f, a.seed);
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                F[k] = rft[k * NR + r];
                Tq[k] = rft[(3 + k) * NR + r];
            }
        }
    }
    double v[6] = {(double)F[0], (double)F[1], (double)F[2], (double)Tq[0], (double)Tq[1], (double)Tq[2]};
    block_sum6(v, sh, tid, bs);
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j) __hip_atomic_store(&L.partials[(size_t)wb * 8 + j], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const uint32_t prev = atomicAdd(L.counter, 1u);
        last = (prev == L.wall_blocks - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = tid; b < L.wall_blocks; b += bs)
#pragma unroll
        for (int j = 0; j < 6; ++j) s[j] += __hip_atomic_load(&L.partials[(size_t)b * 8 + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    block_sum6(s, sh, tid, bs);
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j) L.hist[(size_t)a.slot * HIST_STRIDE + j] = s[j];
        __hip_atomic_store(L.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- stream-collide
// Grid: (xblocks * yblocks, wall_rows + nz); block (bx, by). Rows
// blockIdx.y < wall_rows are wall blocks; the rest are bulk tiles of `by`
// x-rows (split in xblocks pieces when nx > bx) at one z.
template <class C, int P, bool FORCE, bool SPONGE>
__global__ void __launch_bounds__(1024) k_stream_collide(DevLattice L, StepArgs a) {
    using T = typename C::T;
    const uint32_t bs = blockDim.x * blockDim.y;
    const uint32_t tid = threadIdx.x + threadIdx.y * blockDim.x;
    if (blockIdx.y < L.wall_rows) {
        wall_path<C, P, FORCE, SPONGE>(L, a, blockIdx.y * gridDim.x + blockIdx.x, tid, bs);
        return;
    }
    const uint32_t xb = blockIdx.x % L.xblocks, yb = blockIdx.x / L.xblocks;
    const uint32_t x = xb * L.bx + threadIdx.x;
    const uint32_t y = yb * L.by + threadIdx.y;
    const uint32_t z = blockIdx.y - L.wall_rows;
    if (x >= L.nx || y >= L.ny) return;
    const uint32_t n = (z * L.ny + y) * L.nx + x;
    if (L.flags[n] != 0u) return;   // solid, face or wall-record cell
    Nbr nb;
    nb.make(L, x, y, z, n);
    T f[19];
    gather_pre<C, P>(L, nb, f);
    T rho, u[3];
    collide<T, FORCE, SPONGE>(L, x, y, z, f, rho, u, NoHook{});
    scatter_post<C, P, false>(L, nb, f, a.seed);
}

// ---------------------------------------------------------------- faces
// face_gather + face_write fused (shader.rs:630-687). The gather reads only
// the previous-step outputs of interior cells and the write only touches
// face-cell output slots, so one thread can do both.
template <class C, int P> __global__ void __launch_bounds__(256) k_faces(DevLattice L, uint32_t seed) {
    using T = typename C::T;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L.nface) return;
    const uint32_t b = L.faces[k];
    uint32_t x, y, z;
    coords(L, b, x, y, z);
    uint32_t mx = x, my = y, mz = z;
    if (!L.per[0]) mx = min(max(mx, 1u), L.nx - 2);
    if (!L.per[1]) my = min(max(my, 1u), L.ny - 2);
    if (!L.per[2]) mz = min(max(mz, 1u), L.nz - 2);
    const uint32_t m = (mz * L.ny + my) * L.nx + mx;
    T rho = (T)1;
    T ux = (T)L.u_inf[0], uy = (T)L.u_inf[1], uz = (T)L.u_inf[2];
    T p[6] = {0, 0, 0, 0, 0, 0};
    if ((L.flags[m] & 3u) == 0) {
        Nbr nm;
        nm.make(L, mx, my, mz, m);
        T f[19];
        gather_post<C, 1 - P>(L, nm, f);
        const Mom<T> mo = moments(f);
        rho = (T)1 + mo.rho1;
        ux = div_(mo.jx, rho) - (T)0.5 * (T)L.grav[0];
        uy = div_(mo.jy, rho) - (T)0.5 * (T)L.grav[1];
        uz = div_(mo.jz, rho) - (T)0.5 * (T)L.grav[2];
        const T r3 = div_(mo.rho1, (T)3);
        p[0] = mo.sxx - rho * ux * ux - r3;
        p[1] = mo.syy - rho * uy * uy - r3;
        p[2] = mo.szz - rho * uz * uz - r3;
        p[3] = mo.sxy - rho * ux * uy;
        p[4] = mo.sxz - rho * ux * uz;
        p[5] = mo.syz - rho * uy * uz;
    }
    const uint32_t ty = L.flags[b] & 3u;
    if (ty == 2u) {
        if (L.total_p) {
            const T ui2 = (T)(L.u_inf[0] * L.u_inf[0] + L.u_inf[1] * L.u_inf[1] + L.u_inf[2] * L.u_inf[2]);
            rho = div_((T)1 / (T)3 + (T)0.5 * ui2, (T)1 / (T)3 + (T)0.5 * (ux * ux + uy * uy + uz * uz));
        } else {
            ux = (T)L.u_inf[0];
            uy = (T)L.u_inf[1];
            uz = (T)L.u_inf[2];
        }
    } else {
        rho = (T)1;
    }
    T fe[19], fq[19];
    feq_arr<T>(rho - (T)1, ux, uy, uz, fe);
    fneq_arr<T>(p, fq);
#pragma unroll
    for (int i = 0; i < 19; ++i) fe[i] = fe[i] + fq[i];
    Nbr nbb;
    nbb.make(L, x, y, z, b);
    scatter_post<C, 1 - P, true>(L, nbb, fe, seed);
}

// ---------------------------------------------------------------- init
template <class C> __global__ void __launch_bounds__(256) k_init(DevLattice L) {
    using T = typename C::T;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= L.ncell) return;
    T fe[19];
    feq_arr<T>((T)0, (T)L.u_init[0], (T)L.u_init[1], (T)L.u_init[2], fe);
    scatter_pre0<C>(L, n, fe, 0u);   // seed of step 0 (lattice.rs:676)
}

// ---------------------------------------------------------------- cell_macro
// shader.rs:599-621 with module parity P (= steps_done % 2)
template <class C, int P>
__device__ __forceinline__ bool cell_macro(const DevLattice &L, uint32_t m, typename C::T &rho, typename C::T u[3]) {
    using T = typename C::T;
    const uint32_t fg = L.flags[m];
    if (fg & 3u) return false;
    const uint32_t rec = fg >> 2;
    if (rec != 0) {
        const T *rmac = (const T *)L.rmac;
        const size_t NR = L.nrec;
        rho = rmac[rec - 1];
        u[0] = rmac[NR + rec - 1];
        u[1] = rmac[2 * NR + rec - 1];
        u[2] = rmac[3 * NR + rec - 1];
        return true;
    }
    uint32_t x, y, z;
    coords(L, m, x, y, z);
    Nbr nb;
    nb.make(L, x, y, z, m);
    T f[19];
    gather_post<C, 1 - P>(L, nb, f);
    const Mom<T> mo = moments(f);
    rho = (T)1 + mo.rho1;
    u[0] = div_(mo.jx, rho) - (T)0.5 * (T)L.grav[0];
    u[1] = div_(mo.jy, rho) - (T)0.5 * (T)L.grav[1];
    u[2] = div_(mo.jz, rho) - (T)0.5 * (T)L.grav[2];
    return true;
}

template <class C, int P> __global__ void __launch_bounds__(256) k_macros(DevLattice L, float4 *out) {
    using T = typename C::T;
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L.ncell) return;
    T rho, u[3];
    if (cell_macro<C, P>(L, m, rho, u)) out[m] = make_float4((float)rho, (float)u[0], (float)u[1], (float)u[2]);
    else out[m] = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""));
}

// mass: per-block partial sums (grid-stride), then one block sums them in order
template <class C, int P> __global__ void __launch_bounds__(256) k_mass_partial(DevLattice L, double *partials) {
    using T = typename C::T;
    __shared__ double sh[4];
    double s = 0.0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < L.ncell; m += gridDim.x * blockDim.x) {
        T rho, u[3];
        if (cell_macro<C, P>(L, m, rho, u)) s += (double)rho;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ void __launch_bounds__(256) k_sum_final(const double *partials, uint32_t n, double *out) {
    __shared__ double sh[4];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = sh[0] + sh[1] + sh[2] + sh[3];
}

// ---------------------------------------------------------------- probes
// shader.rs:736-784 (run with the module of the other parity after a step)
template <class C, int P> __global__ void __launch_bounds__(256) k_sample_verts(DevLattice L, StepArgs a) {
    using T = typename C::T;
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= L.nvert) return;
    const float4 pb = L.verts[2 * v], nn = L.verts[2 * v + 1];
    const float bx = pb.x + L.vertex_offset * nn.x, by = pb.y + L.vertex_offset * nn.y, bz = pb.z + L.vertex_offset * nn.z;
    // CENTER + from_body(b)
    const float sx = (float)L.center[0] + (a.r[0][0] * bx + a.r[1][0] * by + a.r[2][0] * bz);
    const float sy = (float)L.center[1] + (a.r[0][1] * bx + a.r[1][1] * by + a.r[2][1] * bz);
    const float sz = (float)L.center[2] + (a.r[0][2] * bx + a.r[1][2] * by + a.r[2][2] * bz);
    const float b0x = floorf(sx), b0y = floorf(sy), b0z = floorf(sz);
    const float frx = sx - b0x, fry = sy - b0y, frz = sz - b0z;
    float acc = 0.0f, wsum = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const int cx = (int)b0x + ox, cy = (int)b0y + oy, cz = (int)b0z + oz;
        if (cx < 1 || cy < 1 || cz < 1 || cx >= (int)L.nx - 1 || cy >= (int)L.ny - 1 || cz >= (int)L.nz - 1) continue;
        const float w = (ox ? frx : 1.0f - frx) * (oy ? fry : 1.0f - fry) * (oz ? frz : 1.0f - frz);
        T rho, u[3];
        const bool ok = cell_macro<C, P>(L, ((uint32_t)cz * L.ny + (uint32_t)cy) * L.nx + (uint32_t)cx, rho, u);
        if (ok && w > 0.0f) {
            acc += w * (float)rho;
            wsum += w;
        }
    }
    if (wsum > 1e-6f) {
        L.vacc[2 * v] += acc / wsum;
        L.vacc[2 * v + 1] += 1.0f;
    }
}

template <class C, int P> __global__ void __launch_bounds__(256) k_sample_slices(DevLattice L) {
    using T = typename C::T;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L.npix) return;
    const int nsl = L.slicedef[0].x;
    uint32_t acc = 0;
    for (int s = 0; s < nsl; ++s) {
        const int4 o = L.slicedef[1 + 3 * s], du = L.slicedef[2 + 3 * s], dv = L.slicedef[3 + 3 * s];
        const uint32_t nu = (uint32_t)o.w, nv = (uint32_t)du.w;
        if (k < acc + nu * nv) {
            const uint32_t l = k - acc, iv = l / nu, iu = l - iv * nu;
            const int cx = o.x + (int)iu * du.x + (int)iv * dv.x;
            const int cy = o.y + (int)iu * du.y + (int)iv * dv.y;
            const int cz = o.z + (int)iu * du.z + (int)iv * dv.z;
            T rho, u[3];
            if (cell_macro<C, P>(L, ((uint32_t)cz * L.ny + (uint32_t)cy) * L.nx + (uint32_t)cx, rho, u)) {
                float *b = &L.sacc[5 * (size_t)k];
                b[0] += (float)rho;
                b[1] += (float)u[0];
                b[2] += (float)u[1];
                b[3] += (float)u[2];
                b[4] += 1.0f;
            }
            return;
        }
        acc += nu * nv;
    }
}

// ---------------------------------------------------------------- dispatch helpers
inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

template <class C, int P>
hipError_t step_launch(const DevLattice &L, const StepArgs &a, hipStream_t s) {
    dim3 grid(L.xblocks * L.yblocks, L.wall_rows + L.nz), block(L.bx, L.by);
    const bool force = L.forcing != 0, sponge = L.sponge != 0;
    if (!force && !sponge) k_stream_collide<C, P, false, false><<<grid, block, 0, s>>>(L, a);
    else if (force && !sponge) k_stream_collide<C, P, true, false><<<grid, block, 0, s>>>(L, a);
    else if (!force && sponge) k_stream_collide<C, P, false, true><<<grid, block, 0, s>>>(L, a);
    else k_stream_collide<C, P, true, true><<<grid, block, 0, s>>>(L, a);
    return hipGetLastError();
}

template <class C> hipError_t step_launch_p(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    return p == 0 ? step_launch<C, 0>(L, a, s) : step_launch<C, 1>(L, a, s);
}

}  // namespace

hipError_t launch_init(const DevLattice &L, hipStream_t s) {
    const uint32_t nb = cdiv(L.ncell, 256);
    switch (L.storage) {
        case SIMLAB_STORAGE_FP32: k_init<CF32><<<nb, 256, 0, s>>>(L); break;
        case SIMLAB_STORAGE_FP16: k_init<CF16><<<nb, 256, 0, s>>>(L); break;
        case SIMLAB_STORAGE_BF16: k_init<CBF16><<<nb, 256, 0, s>>>(L); break;
        default: k_init<CF64><<<nb, 256, 0, s>>>(L); break;
    }
    return hipGetLastError();
}

hipError_t launch_faces(const DevLattice &L, int p, uint32_t seed, hipStream_t s) {
    if (L.nface == 0) return hipSuccess;
    const uint32_t nb = cdiv(L.nface, 256);
#define FACES(C)                                                    \
    do {                                                            \
        if (p == 0) k_faces<C, 0><<<nb, 256, 0, s>>>(L, seed);      \
        else k_faces<C, 1><<<nb, 256, 0, s>>>(L, seed);             \
    } while (0)
    switch (L.storage) {
        case SIMLAB_STORAGE_FP32: FACES(CF32); break;
        case SIMLAB_STORAGE_FP16: FACES(CF16); break;
        case SIMLAB_STORAGE_BF16: FACES(CBF16); break;
        default: FACES(CF64); break;
    }
#undef FACES
    return hipGetLastError();
}

hipError_t launch_stream_collide(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    switch (L.storage) {
        case SIMLAB_STORAGE_FP32: return step_launch_p<CF32>(L, p, a, s);
        case SIMLAB_STORAGE_FP16: return step_launch_p<CF16>(L, p, a, s);
        case SIMLAB_STORAGE_BF16: return step_launch_p<CBF16>(L, p, a, s);
        default: return step_launch_p<CF64>(L, p, a, s);
    }
}

hipError_t launch_macros(const DevLattice &L, int p, float4 *out, hipStream_t s) {
    const uint32_t nb = cdiv(L.ncell, 256);
#define MAC(C)                                                      \
    do {                                                            \
        if (p == 0) k_macros<C, 0><<<nb, 256, 0, s>>>(L, out);      \
        else k_macros<C, 1><<<nb, 256, 0, s>>>(L, out);             \
    } while (0)
    switch (L.storage) {
        case SIMLAB_STORAGE_FP32: MAC(CF32); break;
        case SIMLAB_STORAGE_FP16: MAC(CF16); break;
        case SIMLAB_STORAGE_BF16: MAC(CBF16); break;
        default: MAC(CF64); break;
    }
#undef MAC
    return hipGetLastError();
}

hipError_t launch_mass(const DevLattice &L, int p, double *partials, uint32_t nblocks, double *out, hipStream_t s) {
#define MASS(C)                                                                \
    do {                                                                       \
        if (p == 0) k_mass_partial<C, 0><<<nblocks, 256, 0, s>>>(L, partials); \
        else k_mass_partial<C, 1><<<nblocks, 256, 0, s>>>(L, partials);        \
    } while (0)
    switch (L.storage) {
        case SIMLAB_STORAGE_FP32: MASS(CF32); break;
        case SIMLAB_STORAGE_FP16: MASS(CF16); break;
        case SIMLAB_STORAGE_BF16: MASS(CBF16); break;
        default: MASS(CF64); break;
    }
#undef MASS
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_sum_final<<<1, 256, 0, s>>>(partials, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_samplers(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    // samplers use the module of the other parity (lattice.rs:733-737)
    const int q = 1 - p;
#define SAMP(C)                                                                                   \
    do {                                                                                          \
        if (L.nvert) {                                                                            \
            if (q == 0) k_sample_verts<C, 0><<<cdiv(L.nvert, 256), 256, 0, s>>>(L, a);            \
            else k_sample_verts<C, 1><<<cdiv(L.nvert, 256), 256, 0, s>>>(L, a);                   \
        }                                                                                         \
        if (L.npix) {                                                                             \
            if (q == 0) k_sample_slices<C, 0><<<cdiv(L.npix, 256), 256, 0, s>>>(L);               \
            else k_sample_slices<C, 1><<<cdiv(L.npix, 256), 256, 0, s>>>(L);                      \
        }                                                                                         \
    } while (0)
    switch (L.storage) {
        case SIMLAB_STORAGE_FP32: SAMP(CF32); break;
        case SIMLAB_STORAGE_FP16: SAMP(CF16); break;
        case SIMLAB_STORAGE_BF16: SAMP(CBF16); break;
        default: SAMP(CF64); break;
    }
#undef SAMP
    return hipGetLastError();
}

}  // namespace simlab
