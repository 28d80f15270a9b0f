// gfx950 kernels of the refined D3Q27 cumulant solver.
//
// Semantics follow the reference's generated WGSL (crates/cfd/src/mshader.rs,
// d3q27.rs); layout and launch structure are this library's own:
//  * pull streaming with two population sets per level (A, B), as the
//    reference: a step reads the neighbours' post-collision populations
//    from the source set, applies the wall conditions, collides and writes
//    the cell's own post-collision populations to the other set;
//  * each set keeps the cells in groups of 64 consecutive indices with a
//    group's 27 directions one after the other, so a wave's 27 stores of one
//    group fill whole cache lines and its loads of a neighbour group are one
//    contiguous span per direction;
//  * one thread per cell, 256-thread blocks over consecutive cells;
//  * collision: the cumulant collision with every higher-order rate 1 (the
//    post-collision central moments are those of a Gaussian with the relaxed
//    second moments), built by the backward chimera transform along z, y, x;
//  * FP16 storage rounds with the reference's per-store stochastic rounding
//    (mshader.rs:418-426): the same hash of cell * 27 + direction and seed;
//  * forces: the wall cells' F and T go to their records; one reduction
//    kernel per finest step sums them per block in FP64, and one kernel per
//    chunk of steps adds the blocks in a fixed order.

#include "mkernels.h"

#include <cstdint>
#include <type_traits>
#include <utility>

namespace simlab {
namespace m27 {
namespace {

// ---------------------------------------------------------------- velocity set (d3q27.rs)
__host__ __device__ constexpr int cx(int i) { return i % 3 - 1; }
__host__ __device__ constexpr int cy(int i) { return (i / 3) % 3 - 1; }
__host__ __device__ constexpr int cz(int i) { return i / 9 - 1; }
__host__ __device__ constexpr int opp(int i) { return 26 - i; }
__host__ __device__ constexpr int mslot(int i) { return i < (int)REST ? i : i - 1; }
__host__ __device__ constexpr double wgt(int i) {
    return (cx(i) != 0) + (cy(i) != 0) + (cz(i) != 0) == 0   ? 8.0 / 27.0
           : (cx(i) != 0) + (cy(i) != 0) + (cz(i) != 0) == 1 ? 2.0 / 27.0
           : (cx(i) != 0) + (cy(i) != 0) + (cz(i) != 0) == 2 ? 1.0 / 54.0
                                                             : 1.0 / 216.0;
}

template <int... Is, class F>
__device__ __forceinline__ void sfor_impl(std::integer_sequence<int, Is...>, F &&f) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void sfor(F &&f) {
    sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

// ---------------------------------------------------------------- storages
struct SF32 { using E = float; using T = float; };
struct SF16 { using E = uint16_t; using T = float; };
struct SF64 { using E = double; using T = double; };

constexpr float FS = 32768.0f;   // mshader.rs:42 FP16_SCALE
constexpr float INV_S = 1.0f / 32768.0f;

// mshader.rs:418-426 sround: the reference's per-store hash and dither
__device__ __forceinline__ float sround(float x, uint32_t key, uint32_t seed) {
    uint32_t h = (key * 0x9E3779B1u) ^ (seed * 0x85EBCA77u + 0x632BE5ABu);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    const float r = (float)(h >> 8) * (1.0f / 16777216.0f) - 0.5f;
    const int e = (int)((__float_as_uint(x) >> 23) & 0xffu) - 127;
    // 2^(max(e, -14) - 10): always a normal float, built from its exponent bits
    const float ulp = __uint_as_float((uint32_t)(max(e, -14) - 10 + 127) << 23);
    return fminf(fmaxf(x + r * ulp, -65000.0f), 65000.0f);
}

template <class S> __device__ __forceinline__ typename S::T dec(typename S::E e) {
    if constexpr (std::is_same_v<S, SF16>) return (float)__builtin_bit_cast(_Float16, e) * INV_S;
    else return e;
}
template <class S> __device__ __forceinline__ typename S::E enc(typename S::T v, uint32_t key, uint32_t seed) {
    if constexpr (std::is_same_v<S, SF16>) {
        const _Float16 h = (_Float16)sround(v * FS, key, seed);   // round to nearest even
        return __builtin_bit_cast(uint16_t, h);
    } else {
        (void)key;
        (void)seed;
        return v;
    }
}

template <class S> __device__ __forceinline__ typename S::E *set_ptr(const DevLevel &L, uint32_t set) {
    return (typename S::E *)L.pops + (size_t)set * set_elems(L.n);
}

// ---------------------------------------------------------------- arithmetic
template <class T> struct K {
    static __device__ __forceinline__ T third() { return (T)(1.0 / 3.0); }
    static __device__ __forceinline__ T two3() { return (T)(2.0 / 3.0); }
};
template <class T, int I> __device__ __forceinline__ constexpr T W() { return (T)wgt(I); }
__device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }

template <class T> struct Mom {
    T rho1, j[3], s[6];
};

// mshader.rs:260-277
template <class T> __device__ __forceinline__ Mom<T> moments(const T f[Q]) {
    Mom<T> m;
    T r = 0, jx = 0, jy = 0, jz = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
    sfor<Q>([&](auto I) {
        constexpr int i = I;
        const T v = f[i];
        r += v;
        if constexpr (cx(i) == 1) jx += v;
        if constexpr (cx(i) == -1) jx -= v;
        if constexpr (cy(i) == 1) jy += v;
        if constexpr (cy(i) == -1) jy -= v;
        if constexpr (cz(i) == 1) jz += v;
        if constexpr (cz(i) == -1) jz -= v;
        if constexpr (cx(i) != 0) sxx += v;
        if constexpr (cy(i) != 0) syy += v;
        if constexpr (cz(i) != 0) szz += v;
        if constexpr (cx(i) * cy(i) == 1) sxy += v;
        if constexpr (cx(i) * cy(i) == -1) sxy -= v;
        if constexpr (cx(i) * cz(i) == 1) sxz += v;
        if constexpr (cx(i) * cz(i) == -1) sxz -= v;
        if constexpr (cy(i) * cz(i) == 1) syz += v;
        if constexpr (cy(i) * cz(i) == -1) syz -= v;
    });
    m.rho1 = r;
    m.j[0] = jx;
    m.j[1] = jy;
    m.j[2] = jz;
    m.s[0] = sxx;
    m.s[1] = syy;
    m.s[2] = szz;
    m.s[3] = sxy;
    m.s[4] = sxz;
    m.s[5] = syz;
    return m;
}

// mshader.rs:279-287 product-form equilibrium, shifted; added (times s) to f
template <class T, bool ADD> __device__ __forceinline__ void feq(T rho, const T u[3], T f[Q], T s = (T)1) {
    T a[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T uu = u[k] * u[k];
        a[k][0] = (T)0.5 * ((K<T>::third() + uu) - u[k]);
        a[k][1] = K<T>::two3() - uu;
        a[k][2] = (T)0.5 * ((K<T>::third() + uu) + u[k]);
    }
    sfor<Q>([&](auto I) {
        constexpr int i = I;
        const T v = rho * a[0][cx(i) + 1] * a[1][cy(i) + 1] * a[2][cz(i) + 1] - W<T, i>();
        if constexpr (ADD) f[i] += s * v;
        else f[i] = v;
    });
}

// mshader.rs:288-301 + 486-493 rebuild: feq(rho, u) + fneq(deviator(p))
template <class T> __device__ __forceinline__ void rebuild(T rho, const T u[3], const T p[6], T f[Q]) {
    feq<T, false>(rho, u, f);
    const T tr0 = (p[0] + p[1] + p[2]) / (T)3.0;
    const T d[6] = {p[0] - tr0, p[1] - tr0, p[2] - tr0, p[3], p[4], p[5]};
    const T tr = (d[0] + d[1] + d[2]) * K<T>::third();
    sfor<Q>([&](auto I) {
        constexpr int i = I;
        T t = -tr;
        if constexpr (cx(i) != 0) t += d[0];
        if constexpr (cy(i) != 0) t += d[1];
        if constexpr (cz(i) != 0) t += d[2];
        if constexpr (cx(i) * cy(i) != 0) t += (T)(2 * cx(i) * cy(i)) * d[3];
        if constexpr (cx(i) * cz(i) != 0) t += (T)(2 * cx(i) * cz(i)) * d[4];
        if constexpr (cy(i) * cz(i) != 0) t += (T)(2 * cy(i) * cz(i)) * d[5];
        f[i] += (T)(4.5 * wgt(i)) * t;
    });
}

// mshader.rs:307-375 the shifted populations with the central moments of a
// Gaussian (covariance sg = xx yy zz xy xz yz, first moments m1) times rho
template <class T> __device__ __forceinline__ void gauss(T rho, const T u[3], const T sg[6], const T m1[3], T f[Q]) {
    const T a = sg[0], b = sg[1], cc = sg[2], d = sg[3], e = sg[4], g = sg[5];
    // k[ix][iy][iz]: central moment of powers (ix, iy, iz), times rho; the
    // transform then runs in place: after z, k[ix][iy][cz]; after y,
    // k[ix][cy][cz]; after x, k[cx][cy][cz] = the population of (cx, cy, cz)
    T k[3][3][3] = {};
    k[0][0][0] = rho;
    k[1][0][0] = rho * m1[0];
    k[0][1][0] = rho * m1[1];
    k[0][0][1] = rho * m1[2];
    k[2][0][0] = rho * a;
    k[0][2][0] = rho * b;
    k[0][0][2] = rho * cc;
    k[1][1][0] = rho * d;
    k[1][0][1] = rho * e;
    k[0][1][1] = rho * g;
    k[2][2][0] = rho * (a * b + (T)2.0 * d * d);
    k[2][0][2] = rho * (a * cc + (T)2.0 * e * e);
    k[0][2][2] = rho * (b * cc + (T)2.0 * g * g);
    k[2][1][1] = rho * (a * g + (T)2.0 * d * e);
    k[1][2][1] = rho * (b * e + (T)2.0 * d * g);
    k[1][1][2] = rho * (cc * d + (T)2.0 * e * g);
    k[2][2][2] = rho * (a * b * cc + (T)2.0 * (a * g * g + b * e * e + cc * d * d) + (T)8.0 * d * e * g);
    // (k0, k1, k2) about w -> populations at -1, 0, +1
    auto back = [](T &k0, T &k1, T &k2, T w) {
        const T r1 = k1 + w * k0;
        const T r2 = k2 + (T)2.0 * w * k1 + w * w * k0;
        const T m = (T)0.5 * (r2 - r1), o = k0 - r2, p = (T)0.5 * (r2 + r1);
        k0 = m;
        k1 = o;
        k2 = p;
    };
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) back(k[i][j][0], k[i][j][1], k[i][j][2], u[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int z = 0; z < 3; ++z) back(k[i][0][z], k[i][1][z], k[i][2][z], u[1]);
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int z = 0; z < 3; ++z) back(k[0][y][z], k[1][y][z], k[2][y][z], u[0]);
    sfor<Q>([&](auto I) {
        constexpr int i = I;
        f[i] = k[cx(i) + 1][cy(i) + 1][cz(i) + 1] - W<T, i>();
    });
}

// mshader.rs:466-477
template <class T> __device__ __forceinline__ void stress(const Mom<T> &mo, T rho, const T u[3], T p[6]) {
    const T r3 = mo.rho1 * K<T>::third();
    p[0] = mo.s[0] - rho * u[0] * u[0] - r3;
    p[1] = mo.s[1] - rho * u[1] * u[1] - r3;
    p[2] = mo.s[2] - rho * u[2] * u[2] - r3;
    p[3] = mo.s[3] - rho * u[0] * u[1];
    p[4] = mo.s[4] - rho * u[0] * u[2];
    p[5] = mo.s[5] - rho * u[1] * u[2];
}

__device__ __forceinline__ void coords(const DevLevel &L, uint32_t n, uint32_t &x, uint32_t &y, uint32_t &z) {
    const uint32_t nxy = L.nx * L.ny;
    z = n / nxy;
    const uint32_t r = n - z * nxy;
    y = r / L.nx;
    x = r - y * L.nx;
}

// mshader.rs:436-450
template <class T> __device__ __forceinline__ T sponge(const DevLevel &L, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t xyz[3] = {x, y, z};
    const uint32_t dm[3] = {L.nx, L.ny, L.nz};
    T s = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T p = (T)xyz[k];
        const T hi = (T)(dm[k] - 1u) - p;
        const T lo = (T)L.sp_lo[k], hw = (T)L.sp_hi[k];
        if (lo > (T)0) s = max(s, (T)1.0 - p / lo);
        if (hw > (T)0) s = max(s, (T)1.0 - hi / hw);
    }
    return min(max(s, (T)0), (T)1);
}

template <class S> __device__ __forceinline__ void load_own(const DevLevel &L, uint32_t set, uint32_t n, typename S::T f[Q]) {
    const typename S::E *p = set_ptr<S>(L, set);
    sfor<Q>([&](auto I) { f[I] = dec<S>(p[slot(I, n)]); });
}
template <class S> __device__ __forceinline__ void store_own(const DevLevel &L, uint32_t set, uint32_t n,
                                                            const typename S::T f[Q], uint32_t seed) {
    typename S::E *p = set_ptr<S>(L, set);
    sfor<Q>([&](auto I) { p[slot(I, n)] = enc<S>(f[I], n * 27u + (uint32_t)I, seed); });
}

template <class T> __device__ __forceinline__ void write_mom(T *o, T rho, const T u[3], const T p[6], T tau) {
    o[0] = rho;
    o[1] = u[0];
    o[2] = u[1];
    o[3] = u[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) o[4 + k] = p[k];
    o[10] = tau;
}

// ---------------------------------------------------------------- kernels
template <class S> __global__ __launch_bounds__(256) void k_init(DevLevel L) {
    using T = typename S::T;
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= L.n) return;
    T f[Q];
    const T u[3] = {(T)L.u_init[0], (T)L.u_init[1], (T)L.u_init[2]};
    feq<T, false>((T)1, u, f);
    store_own<S>(L, 0, n, f, 0u);
    store_own<S>(L, 1, n, f, 0u);
}

// mshader.rs:636-728 (STEP) with the walls of mshader.rs:376-414
template <class S> __global__ __launch_bounds__(256) void k_step(DevLevel L, StepArgs a) {
    using T = typename S::T;
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= L.n) return;
    const uint32_t fg = L.flags[n];
    if ((fg & 7u) != 0u) return;
    uint32_t x, y, z;
    coords(L, n, x, y, z);
    // neighbour offsets, wrapping at the level's edges (mshader.rs:103-106)
    const int32_t nx = (int32_t)L.nx, nxy = (int32_t)(L.nx * L.ny);
    const int32_t xp = x == L.nx - 1 ? -(nx - 1) : 1, xm = x == 0 ? nx - 1 : -1;
    const int32_t yp = y == L.ny - 1 ? -(int32_t)((L.ny - 1) * L.nx) : nx, ym = y == 0 ? (int32_t)((L.ny - 1) * L.nx) : -nx;
    const int32_t zp = z == L.nz - 1 ? -(int32_t)((L.nz - 1) * L.nx * L.ny) : nxy;
    const int32_t zm = z == 0 ? (int32_t)((L.nz - 1) * L.nx * L.ny) : -nxy;
    const typename S::E *src = set_ptr<S>(L, a.src);
    T f[Q];
    // pull: the population moving along c_i comes from n - c_i (mshader.rs:219-227)
    sfor<Q>([&](auto I) {
        constexpr int i = I;
        int32_t o = 0;
        if constexpr (cx(i) == 1) o += xm;
        if constexpr (cx(i) == -1) o += xp;
        if constexpr (cy(i) == 1) o += ym;
        if constexpr (cy(i) == -1) o += yp;
        if constexpr (cz(i) == 1) o += zm;
        if constexpr (cz(i) == -1) o += zp;
        f[i] = dec<S>(src[slot(i, n + (uint32_t)o)]);
    });
    const uint32_t rec = fg >> 3;
    T F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
    if (rec != 0u) {
        const uint32_t r = rec - 1u;
        const uint32_t mask = L.rmask[r];
        T fin[Q];
        sfor<Q>([&](auto I) { fin[I] = f[I]; });
        const T pc[3] = {(T)x - (T)L.center[0], (T)y - (T)L.center[1], (T)z - (T)L.center[2]};
        const T om[3] = {(T)a.omega[0], (T)a.omega[1], (T)a.omega[2]};
        sfor<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (i != (int)REST) {
                if (mask & (1u << i)) {
                    constexpr int o = opp(i), sl = mslot(i);
                    const uint32_t word = L.rq[(size_t)(sl / 2) * L.nrec + r];
                    const uint16_t hb = (uint16_t)((sl & 1) ? (word >> 16) : (word & 0xffffu));
                    const T q = max((T)(float)__builtin_bit_cast(_Float16, hb), (T)0.001);
                    constexpr T w = W<T, i>();
                    const T fpi = dec<S>(src[slot(i, n)]) + w;
                    const T fpo = dec<S>(src[slot(o, n)]) + w;
                    constexpr T ci[3] = {(T)cx(i), (T)cy(i), (T)cz(i)};
                    const T xw[3] = {pc[0] + q * ci[0], pc[1] + q * ci[1], pc[2] + q * ci[2]};
                    const T uw[3] = {om[1] * xw[2] - om[2] * xw[1], om[2] * xw[0] - om[0] * xw[2],
                                     om[0] * xw[1] - om[1] * xw[0]};
                    const T cu = ci[0] * uw[0] + ci[1] * uw[1] + ci[2] * uw[2];
                    const T w6 = (T)6.0 * w;
                    T fb;
                    if (q >= (T)0.5) fb = (fpi + ((T)2.0 * q - (T)1.0) * fpo - w6 * cu) / ((T)2.0 * q);
                    else if (!(mask & (1u << o))) fb = (T)2.0 * q * fpi + ((T)1.0 - (T)2.0 * q) * (fin[i] + w) - w6 * cu;
                    else fb = fpi - w6 * cu;
                    f[o] = fb - w;
                    T flk[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) flk[k] = (ci[k] - uw[k]) * fpi + (ci[k] + uw[k]) * fb;
#pragma unroll
                    for (int k = 0; k < 3; ++k) F[k] += flk[k];
                    Tq[0] += xw[1] * flk[2] - xw[2] * flk[1];
                    Tq[1] += xw[2] * flk[0] - xw[0] * flk[2];
                    Tq[2] += xw[0] * flk[1] - xw[1] * flk[0];
                }
            }
        });
    }
    const Mom<T> mo = moments<T>(f);
    const T rho = (T)1.0 + mo.rho1;
    // one reciprocal of rho for its eleven divisions (an IEEE division is
    // about ten instructions, and this step is bound by its arithmetic in
    // 16-bit storage): the same values to about an ulp
    const T irho = (T)1.0 / rho;
    const T G[3] = {rho * (T)L.grav[0], rho * (T)L.grav[1], rho * (T)L.grav[2]};
    T u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) u[k] = (mo.j[k] + (T)0.5 * G[k]) * irho;
    T p[6];
    stress<T>(mo, rho, u, p);
    const T tr = (p[0] + p[1] + p[2]) * K<T>::third();
    const T dv[6] = {p[0] - tr, p[1] - tr, p[2] - tr, p[3], p[4], p[5]};
    T tau = (T)L.tau0;
    if (L.smag2 > 0.0) {
        const T pp = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2] + (T)2.0 * (dv[3] * dv[3] + dv[4] * dv[4] + dv[5] * dv[5]);
        const T t0 = (T)L.tau0;
        tau = (T)0.5 * (t0 + sqrt_(t0 * t0 + (T)18.0 * (T)L.smag2 * sqrt_((T)2.0 * pp) * irho));
    }
    T sp = 0;
    if (L.n_sponge) {
        sp = sponge<T>(L, x, y, z);
        const T st = (T)L.sponge_tau;
        if (sp > (T)0 && st > tau) tau += sp * sp * (st - tau);
    }
    const T omg = (T)1.0 / tau;
    const T cr = ((T)1.0 - omg) * irho;
    T sg[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) sg[k] = K<T>::third() + cr * dv[k];
#pragma unroll
    for (int k = 3; k < 6; ++k) sg[k] = cr * dv[k];
    T m1[3] = {0, 0, 0};
    if (L.forcing) {
#pragma unroll
        for (int k = 0; k < 3; ++k) m1[k] = (T)0.5 * G[k] * irho;
    }
    gauss<T>(rho, u, sg, m1, f);
    if (L.n_sponge && L.sigma > 0.0 && sp > (T)0) {
        // absorbing layer: toward feq(1, u_t) from feq(rho, u)
        const T sgm = (T)L.sigma * sp * sp;
        const T ut[3] = {L.absorb ? (T)L.u_inf[0] : u[0], L.absorb ? (T)L.u_inf[1] : u[1], L.absorb ? (T)L.u_inf[2] : u[2]};
        feq<T, true>((T)1.0, ut, f, sgm);
        feq<T, true>(rho, u, f, -sgm);
    }
    if (rec != 0u) {
        const uint32_t r = rec - 1u;
        T *ft = (T *)L.rft;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ft[(size_t)k * L.nrec + r] = F[k];
            ft[(size_t)(3 + k) * L.nrec + r] = Tq[k];
        }
    }
    store_own<S>(L, 1u - a.src, n, f, a.seed);
    if (L.has_child) {
        const int32_t l0 = (int32_t)x - L.ch_lo[0], l1 = (int32_t)y - L.ch_lo[1], l2 = (int32_t)z - L.ch_lo[2];
        if (l0 >= 0 && l1 >= 0 && l2 >= 0 && l0 < (int32_t)L.ch_n[0] && l1 < (int32_t)L.ch_n[1] && l2 < (int32_t)L.ch_n[2]) {
            if (l0 == 0 || l1 == 0 || l2 == 0 || l0 == (int32_t)L.ch_n[0] - 1 || l1 == (int32_t)L.ch_n[1] - 1 ||
                l2 == (int32_t)L.ch_n[2] - 1) {
                const size_t o = ((size_t)a.mw * L.ch_nodes + ((size_t)l2 * L.ch_n[1] + l1) * L.ch_n[0] + l0) * MW;
                write_mom<T>((T *)L.mom_child + o, rho, u, p, tau);
            }
        }
    }
    if (L.has_parent && !(x & 1u) && !(y & 1u) && !(z & 1u)) {
        const size_t o = (((size_t)(z / 2) * L.pn[1] + y / 2) * L.pn[0] + x / 2) * MW;
        write_mom<T>((T *)L.mom_up + o, rho, u, p, tau);
    }
}

// mshader.rs:732-762 domain faces of the coarsest level
template <class S> __global__ __launch_bounds__(256) void k_face(DevLevel L, StepArgs a) {
    using T = typename S::T;
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= L.nface) return;
    const uint32_t b = L.lists[k];
    uint32_t x, y, z;
    coords(L, b, x, y, z);
    if (!L.per[0]) x = min(max(x, 1u), L.nx - 2u);
    if (!L.per[1]) y = min(max(y, 1u), L.ny - 2u);
    if (!L.per[2]) z = min(max(z, 1u), L.nz - 2u);
    const uint32_t m = (z * L.ny + y) * L.nx + x;
    T rho = 1, u[3] = {(T)L.u_inf[0], (T)L.u_inf[1], (T)L.u_inf[2]}, pp[6] = {0, 0, 0, 0, 0, 0};
    T f[Q];
    if ((L.flags[m] & 7u) == 0u) {
        load_own<S>(L, a.src, m, f);
        const Mom<T> mo = moments<T>(f);
        rho = (T)1.0 + mo.rho1;
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = mo.j[c] / rho;
        stress<T>(mo, rho, u, pp);
    }
    const uint32_t ty = (L.flags[b] >> 3) & 3u;
    if (ty == 1u) {
        if (L.total_p) {
            const T ui[3] = {(T)L.u_inf[0], (T)L.u_inf[1], (T)L.u_inf[2]};
            const T uinf2 = ui[0] * ui[0] + ui[1] * ui[1] + ui[2] * ui[2];
            const T u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            rho = (K<T>::third() + (T)0.5 * uinf2) / (K<T>::third() + (T)0.5 * u2);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) u[c] = (T)L.u_inf[c];
        }
    } else {
        rho = 1;
    }
    rebuild<T>(rho, u, pp, f);
    store_own<S>(L, a.src, b, f, a.seed);
}

// mshader.rs:771-801 ghosts of a finer level from its parent's surface
// moments, linear in space and time, the stress rescaled to this level
template <class S> __global__ __launch_bounds__(256) void k_ghost(DevLevel L, StepArgs a) {
    using T = typename S::T;
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= L.nghost) return;
    const uint32_t n = L.lists[L.nface + k];
    uint32_t x, y, z;
    coords(L, n, x, y, z);
    const uint32_t h0[3] = {x / 2u, y / 2u, z / 2u};
    const uint32_t odd[3] = {x & 1u, y & 1u, z & 1u};
    const T *mp = (const T *)L.mom_parent;
    T acc[11];
#pragma unroll
    for (int j = 0; j < 11; ++j) acc[j] = 0;
    T wsum = 0;
    const T blend = (T)a.blend;
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) {
        const uint32_t o[3] = {c & 1u, (c >> 1) & 1u, (c >> 2) & 1u};
        if ((o[0] & ~odd[0]) | (o[1] & ~odd[1]) | (o[2] & ~odd[2])) continue;
        const size_t node = ((size_t)(h0[2] + o[2]) * L.pn[1] + (h0[1] + o[1])) * L.pn[0] + (h0[0] + o[0]);
        const T w = (odd[0] ? (T)0.5 : (T)1.0) * (odd[1] ? (T)0.5 : (T)1.0) * (odd[2] ? (T)0.5 : (T)1.0);
        const T *ma = mp + ((size_t)a.mold * L.pnn + node) * MW;
        const T *mb = mp + ((size_t)a.mnew * L.pnn + node) * MW;
#pragma unroll
        for (int j = 0; j < 11; ++j) acc[j] += w * (((T)1.0 - blend) * ma[j] + blend * mb[j]);
        wsum += w;
    }
    const T rho = acc[0] / wsum;
    const T u[3] = {acc[1] / wsum, acc[2] / wsum, acc[3] / wsum};
    const T tc = acc[10] / wsum;
    const T tf = (T)0.5 + (T)2.0 * (tc - (T)0.5);
    // the same physical stress at half the time step, Pi_f = Pi_c tf / (2 tc),
    // then its post-collision part (1 - 1/tf)
    const T sc = tf / ((T)2.0 * tc) * ((T)1.0 - (T)1.0 / tf);
    T p[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) p[j] = acc[4 + j] / wsum * sc;
    T f[Q];
    rebuild<T>(rho, u, p, f);
    store_own<S>(L, a.src, n, f, a.seed);
}

// mshader.rs:803-820 coarse cells inside the child box from the child's
// coincident nodes
template <class S> __global__ __launch_bounds__(256) void k_restrict(DevLevel L, StepArgs a) {
    using T = typename S::T;
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= L.nrestrict) return;
    const uint32_t n = L.lists[L.nface + L.nghost + k];
    uint32_t x, y, z;
    coords(L, n, x, y, z);
    const uint32_t l0 = x - (uint32_t)L.ch_lo[0], l1 = y - (uint32_t)L.ch_lo[1], l2 = z - (uint32_t)L.ch_lo[2];
    const T *m = (const T *)L.mom_child_up + (((size_t)l2 * L.ch_n[1] + l1) * L.ch_n[0] + l0) * MW;
    const T rho = m[0];
    const T u[3] = {m[1], m[2], m[3]};
    const T tf = m[10];
    const T tc = (T)0.5 + (T)0.5 * (tf - (T)0.5);
    const T sc = (T)2.0 * tc / tf * ((T)1.0 - (T)1.0 / tc);
    T p[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) p[j] = m[4 + j] * sc;
    T f[Q];
    rebuild<T>(rho, u, p, f);
    store_own<S>(L, a.src, n, f, a.seed);
}

// records' F and T of one step: block sums in FP64 (fixed order: a thread's
// records in order, then the wave tree, then the waves in order)
template <class T> __global__ __launch_bounds__(256) void k_reduce1(const T *ft, uint32_t nrec, double *partials) {
    __shared__ double sh[4][6];
    double s[6] = {0, 0, 0, 0, 0, 0};
    const uint32_t start = blockIdx.x * RED_CHUNK;
    const uint32_t end = min(start + RED_CHUNK, nrec);
    for (uint32_t r = start + threadIdx.x; r < end; r += 256u)
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] += (double)ft[(size_t)k * nrec + r];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off, 64);
    const uint32_t wave = threadIdx.x / 64u, lane = threadIdx.x % 64u;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) sh[wave][k] = s[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = (int)threadIdx.x;
        partials[(size_t)blockIdx.x * 6 + k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}

__global__ void k_reduce2(const double *partials, uint32_t nblk, uint32_t nsteps, double *hist) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsteps * 6u) return;
    const uint32_t step = t / 6u, k = t % 6u;
    double s = 0;
    for (uint32_t b = 0; b < nblk; ++b) s += partials[((size_t)step * nblk + b) * 6 + k];
    hist[(size_t)step * 8 + k] = s;
}

// (rho, u) of a cell from the source set (mshader.rs:495-506); false for solid
template <class S> __device__ __forceinline__ bool cell_macro(const DevLevel &L, uint32_t set, uint32_t m,
                                                              typename S::T &rho, typename S::T u[3]) {
    using T = typename S::T;
    if ((L.flags[m] & 7u) == 1u) return false;
    T f[Q];
    load_own<S>(L, set, m, f);
    const Mom<T> mo = moments<T>(f);
    rho = (T)1.0 + mo.rho1;
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c] = mo.j[c] / rho - (T)0.5 * (T)L.grav[c];
    return true;
}

// mshader.rs:613-622
template <class S> __global__ __launch_bounds__(256) void k_dump(DevLevel L, uint32_t src) {
    using T = typename S::T;
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= L.n) return;
    T rho, u[3];
    const bool ok = cell_macro<S>(L, src, n, rho, u);
    if (ok && (L.flags[n] & 7u) == 0u) L.dump[n] = make_float4((float)rho, (float)u[0], (float)u[1], (float)u[2]);
    else L.dump[n] = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""));
}

// mshader.rs:571-596 surface pressure probes (finest level)
template <class S> __global__ __launch_bounds__(256) void k_sample_verts(DevLevel L, StepArgs a, float *probes,
                                                                         uint32_t nvert, float voff) {
    using T = typename S::T;
    const uint32_t v = blockIdx.x * 256u + threadIdx.x;
    if (v >= nvert) return;
    float *pr = probes + 8 * (size_t)v;
    const T b[3] = {(T)pr[0] + (T)voff * (T)pr[3], (T)pr[1] + (T)voff * (T)pr[4], (T)pr[2] + (T)voff * (T)pr[5]};
    T s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        s[k] = (T)L.center[k] + ((T)(float)a.rot[k][0] * b[0] + (T)(float)a.rot[k][1] * b[1] + (T)(float)a.rot[k][2] * b[2]);
    T acc = 0, wsum = 0;
    const int32_t dm[3] = {(int32_t)L.nx, (int32_t)L.ny, (int32_t)L.nz};
    for (uint32_t kk = 0; kk < 8; ++kk) {
        const uint32_t o[3] = {kk & 1u, (kk >> 1) & 1u, (kk >> 2) & 1u};
        int32_t c[3];
        bool out = false;
        T w = 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const T b0 = floor(s[k]);
            const T fr = s[k] - b0;
            c[k] = (int32_t)b0 + (int32_t)o[k];
            if (c[k] < 1 || c[k] >= dm[k] - 1) out = true;
            w *= o[k] ? fr : (T)1.0 - fr;
        }
        if (out) continue;
        T rho, u[3];
        const uint32_t m = ((uint32_t)c[2] * L.ny + (uint32_t)c[1]) * L.nx + (uint32_t)c[0];
        if (cell_macro<S>(L, a.src, m, rho, u) && w > (T)0) {
            acc += w * rho;
            wsum += w;
        }
    }
    if (wsum > (T)1e-6) {
        pr[6] += (float)(acc / wsum);
        pr[7] += 1.0f;
    }
}

// mshader.rs:598-611 slice pixels a level owns
template <class S> __global__ __launch_bounds__(256) void k_sample_slices(DevLevel L, uint32_t src) {
    using T = typename S::T;
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= L.npix) return;
    const uint32_t px = L.spix[2 * (size_t)k], cell = L.spix[2 * (size_t)k + 1];
    T rho, u[3];
    if (cell_macro<S>(L, src, cell, rho, u)) {
        float *s = L.sacc + 5 * (size_t)px;
        s[0] += (float)rho;
        s[1] += (float)u[0];
        s[2] += (float)u[1];
        s[3] += (float)u[2];
        s[4] += 1.0f;
    }
}

template <class S> __global__ __launch_bounds__(256) void k_get_pops(DevLevel L, uint32_t src, double *out) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= L.n) return;
    const typename S::E *p = set_ptr<S>(L, src);
#pragma unroll
    for (uint32_t i = 0; i < Q; ++i) out[(size_t)i * L.n + n] = (double)dec<S>(p[slot(i, n)]);
}

inline dim3 blocks(uint64_t threads) { return dim3((uint32_t)((threads + 255) / 256)); }

template <template <class> class KF, class... A> hipError_t by_storage(int32_t storage, uint64_t threads, hipStream_t st,
                                                                       A... args) {
    if (threads == 0) return hipSuccess;
    switch (storage) {
        case SIMLAB_STORAGE_FP32: hipLaunchKernelGGL(KF<SF32>::fn, blocks(threads), dim3(256), 0, st, args...); break;
        case SIMLAB_STORAGE_FP16: hipLaunchKernelGGL(KF<SF16>::fn, blocks(threads), dim3(256), 0, st, args...); break;
        case SIMLAB_STORAGE_FP64: hipLaunchKernelGGL(KF<SF64>::fn, blocks(threads), dim3(256), 0, st, args...); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

#define SIMLAB_KF(name)                                     \
    template <class S> struct F_##name {                    \
        static constexpr auto fn = &name<S>;                \
    };
SIMLAB_KF(k_init)
SIMLAB_KF(k_step)
SIMLAB_KF(k_face)
SIMLAB_KF(k_ghost)
SIMLAB_KF(k_restrict)
SIMLAB_KF(k_dump)
SIMLAB_KF(k_sample_verts)
SIMLAB_KF(k_sample_slices)
SIMLAB_KF(k_get_pops)
#undef SIMLAB_KF

}  // namespace

hipError_t m_init(const DevLevel &L, hipStream_t st) { return by_storage<F_k_init>(L.storage, L.n, st, L); }
hipError_t m_step(const DevLevel &L, const StepArgs &a, hipStream_t st) {
    return by_storage<F_k_step>(L.storage, L.n, st, L, a);
}
hipError_t m_face(const DevLevel &L, const StepArgs &a, hipStream_t st) {
    return by_storage<F_k_face>(L.storage, L.nface, st, L, a);
}
hipError_t m_ghost(const DevLevel &L, const StepArgs &a, hipStream_t st) {
    return by_storage<F_k_ghost>(L.storage, L.nghost, st, L, a);
}
hipError_t m_restrict(const DevLevel &L, const StepArgs &a, hipStream_t st) {
    return by_storage<F_k_restrict>(L.storage, L.nrestrict, st, L, a);
}
hipError_t m_dump(const DevLevel &L, uint32_t src, hipStream_t st) { return by_storage<F_k_dump>(L.storage, L.n, st, L, src); }
hipError_t m_sample_verts(const DevLevel &L, const StepArgs &a, float *probes, uint32_t nvert, double voff, hipStream_t st) {
    return by_storage<F_k_sample_verts>(L.storage, nvert, st, L, a, probes, nvert, (float)voff);
}
hipError_t m_sample_slices(const DevLevel &L, uint32_t src, hipStream_t st) {
    return by_storage<F_k_sample_slices>(L.storage, L.npix, st, L, src);
}
hipError_t m_get_pops(const DevLevel &L, uint32_t src, double *out, hipStream_t st) {
    return by_storage<F_k_get_pops>(L.storage, L.n, st, L, src, out);
}

hipError_t m_reduce1(const DevLevel &L, double *partials, hipStream_t st) {
    if (L.nrec == 0) return hipSuccess;
    const dim3 g((L.nrec + RED_CHUNK - 1) / RED_CHUNK);
    if (L.storage == SIMLAB_STORAGE_FP64)
        hipLaunchKernelGGL(k_reduce1<double>, g, dim3(256), 0, st, (const double *)L.rft, L.nrec, partials);
    else
        hipLaunchKernelGGL(k_reduce1<float>, g, dim3(256), 0, st, (const float *)L.rft, L.nrec, partials);
    return hipGetLastError();
}

hipError_t m_reduce2(const double *partials, uint32_t nblk, uint32_t nsteps, double *hist, hipStream_t st) {
    if (nsteps == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce2, blocks((uint64_t)nsteps * 6), dim3(256), 0, st, partials, nblk, nsteps, hist);
    return hipGetLastError();
}

}  // namespace m27
}  // namespace simlab
