// gfx950 kernels of the D3Q19 wind-tunnel lattice.
//
// Semantics follow the reference's WGSL (crates/cfd/src/shader.rs); the
// data layout and launch structure are this library's own:
//  * populations: cells in groups of 64 consecutive indices, each group
//    holding its 19 directions one after the other (slot (i, n) = element
//    (n / 64) * 19 * 64 + i * 64 + n % 64), stored as deviations
//    f~ = f - w_i (FP32 / FP64 / scaled FP16 / BF16). A wave's 19 loads
//    of one cell group fall in one contiguous span (2.4 KB FP16), which
//    HBM serves far better than 19 (parity 1: 38) separate streams;
//  * AA pattern in place (shader.rs:5-13): a step of parity 0 reads (n, i)
//    and writes (n, opp i); parity 1 reads (n - c_i, opp i) and writes
//    (n + c_i, i). Every slot is touched by one thread per step;
//  * one thread per cell; the bulk kernel's block is `by` whole x-rows, so
//    the +-1 x-offset accesses of parity 1 stay inside one block (one CU
//    and one XCD L2) and waves read and write whole cache lines;
//  * two launches per step: k_special (wall-record cells, cells that feed
//    face cells, and the FP64 momentum-exchange reduction that replaces
//    reduce1/reduce2, shader.rs:689-734) and k_bulk (every other fluid
//    cell). Keeping the wall logic out of k_bulk keeps its register count
//    (and so its occupancy) that of the plain collision;
//  * the face update (shader.rs:630-687) is fused into the step before: the
//    interior cell that a face cell copies its state from writes the face
//    cell's populations right after its own collision;
//  * FP16/BF16 stores round stochastically: one hash per cell, xorshift32
//    words give 16 dither bits per stored direction, added below the kept
//    mantissa bits before a round-toward-zero conversion.

#include "kernels.h"

#ifndef SIMLAB_KSTORE
#error "SIMLAB_KSTORE: the storage compiled here (0-3), or -1 for the storage-independent part"
#endif

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

// flags word: bits 0-1 cell type (0 fluid, 1 solid, 2 velocity face, 3 open
// face), bits 2-30 wall record index + 1, or FLAG_HALFWAY with a link mask
constexpr uint32_t REC_MASK = 0x1FFFFFFFu;
// wall record index + 1 of a flags word (0: none)
__device__ __forceinline__ uint32_t rec_of(uint32_t fg) { return (fg & FLAG_HALFWAY) ? 0u : ((fg >> 2) & REC_MASK); }

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
        // no clamp: the round-toward-zero conversion saturates finite values
        // beyond the half range at +-65504
        float t = __uint_as_float(__float_as_uint(v * FS) + (d16 >> 3));
        half2_t h = __builtin_amdgcn_cvt_pkrtz(t, t);
        return __builtin_bit_cast(uint16_t, h[0]);
    } else if constexpr (std::is_same_v<C, CBF16>) {
        return (uint16_t)((__float_as_uint(v) + d16) >> 16);
    } else {
        (void)d16;
        return v;
    }
}

// FP16 in its stored scale (x 2^15, the bulk step's scaled arithmetic; the
// other storages as dec / enc): decoding skips the multiply by 2^-15,
// encoding the one by 2^15. Both scalings are exact, so a step computed in
// the scaled units stores the same bits.
template <class C> __device__ __forceinline__ typename C::T dec_raw(typename C::E e) {
    if constexpr (std::is_same_v<C, CF16>) return (float)__builtin_bit_cast(_Float16, e);
    else return dec<C>(e);
}
template <class C> __device__ __forceinline__ typename C::E enc_raw(typename C::T v, uint32_t d16) {
    if constexpr (std::is_same_v<C, CF16>) {
        float t = __uint_as_float(__float_as_uint(v) + (d16 >> 3));
        half2_t h = __builtin_amdgcn_cvt_pkrtz(t, t);
        return __builtin_bit_cast(uint16_t, h[0]);
    } else {
        return enc<C>(v, d16);
    }
}
// the population scale of a step in scaled units
template <class C, bool RAW> constexpr float pop_scale() { return (RAW && std::is_same_v<C, CF16>) ? FS : 1.0f; }

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

// population slot of direction i of cell n
template <class C> __device__ __forceinline__ typename C::E &pop(const DevLattice &L, int i, uint32_t n) {
    return ((typename C::E *)L.pops)[pop_index((uint32_t)i, n)];
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
template <class C, int Q, bool RAW = false>
__device__ __forceinline__ void gather_pre(const DevLattice &L, const Nbr &nb, typename C::T f[19]) {
    sfor<19>([&](auto I) {
        constexpr int i = I;
        if constexpr (Q == 0 || i == 0) {
            const typename C::E e = pop<C>(L, i, nb.n);
            f[i] = RAW ? dec_raw<C>(e) : dec<C>(e);
        } else {
            const typename C::E e = pop<C>(L, OPP[i], nb.template at<-CX[i], -CY[i], -CZ[i]>());
            f[i] = RAW ? dec_raw<C>(e) : dec<C>(e);
        }
    });
}

template <class C, int Q> __device__ __forceinline__ void gather_post(const DevLattice &L, const Nbr &nb, typename C::T f[19]) {
    sfor<19>([&](auto I) {
        constexpr int i = I;
        if constexpr (Q == 0 || i == 0) {
            f[i] = dec<C>(pop<C>(L, OPP[i], nb.n));
        } else {
            f[i] = dec<C>(pop<C>(L, i, nb.template at<CX[i], CY[i], CZ[i]>()));
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

template <class C, int Q, bool GUARD, bool RT = false, bool RAW = false>
__device__ __forceinline__ void scatter_post(const DevLattice &L, const Nbr &nb, typename C::T f[19], uint32_t seed) {
    // RT: f[i] becomes the value as stored (decoded back), for face gathers
    Dither d;
    if constexpr (C::SR) d.init(nb.n, seed);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        if constexpr (Q == 0 || i == 0) {
            constexpr int j = OPP[i];   // slot direction
            const uint32_t dj = C::SR ? d.template d16<j>() : 0u;
            const typename C::E e = RAW ? enc_raw<C>(f[i], dj) : enc<C>(f[i], dj);
            pop<C>(L, j, nb.n) = e;
            if constexpr (RT) f[i] = dec<C>(e);
        } else {
            bool ok = true;
            if constexpr (GUARD) ok = keeps_inside<i>(L, nb);
            const uint32_t di = C::SR ? d.template d16<i>() : 0u;
            const typename C::E e = RAW ? enc_raw<C>(f[i], di) : enc<C>(f[i], di);
            if (ok) pop<C>(L, i, nb.template at<CX[i], CY[i], CZ[i]>()) = e;
            if constexpr (RT) f[i] = dec<C>(e);
        }
    });
}

template <class C> __device__ __forceinline__ void scatter_pre0(const DevLattice &L, uint32_t n, const typename C::T f[19], uint32_t seed) {
    Dither d;
    if constexpr (C::SR) d.init(n, seed);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        pop<C>(L, i, n) = enc<C>(f[i], C::SR ? d.template d16<i>() : 0u);
    });
}

// ---------------------------------------------------------------- parity 1 on aligned rows
// With nx a multiple of POP_GROUP every row starts a group, so the slot of
// cell (x + dx, y + dy, z + dz) is element plane(z + dz) + row(y) + ex(dx) +
// sy(dy) + i * 64: ex one of three per-lane values (the cell and its
// x-neighbours, wrapped within the row), the plane, row and sy wave-uniform
// (a wave's 64 cells share y and z). The parity-1 step's 19 loads and 19
// stores then take three offset registers, one buffer base per z-plane and
// scalar offsets instead of a 64-bit address per access. The slots, values
// and dither are those of gather_pre<1> and scatter_post<1, false>.
struct RowAddr {
    uint32_t e[3];                // byte offset of slot 0 of x - 1, x, x + 1 (wrapped within the row) from the row's start
    int32_t sy[3];                // element shift of a row step of -1, 0, +1 (wrapped) less lo; wave-uniform, >= 0
    __amdgpu_buffer_rsrc_t r[3];  // the buffer from the lowest row the wave touches in the plane z - 1, z, z + 1 (wrapped)
};
// element of slot 0 of the cell x of a row, from the row's first slot
__device__ __forceinline__ uint32_t row_elem(uint32_t x) { return (x / POP_GROUP) * (19 * POP_GROUP) + (x % POP_GROUP); }
template <class C> __device__ __forceinline__ RowAddr row_addr(const DevLattice &L, uint32_t x, uint32_t yl, uint32_t zl) {
    constexpr uint32_t ES = sizeof(typename C::E);
    static_assert(ES == 2 || ES == 4, "16- or 32-bit storage");
    // every lane of the wave holds the same y and z (whole wave active here)
    const uint32_t y = __builtin_amdgcn_readfirstlane(yl), z = __builtin_amdgcn_readfirstlane(zl);
    RowAddr r;
    r.e[0] = ES * row_elem(x == 0 ? L.nx - 1 : x - 1);
    r.e[1] = ES * row_elem(x);
    r.e[2] = ES * row_elem(x == L.nx - 1 ? 0 : x + 1);
    const int32_t gr = (int32_t)(L.nx / POP_GROUP) * (19 * (int32_t)POP_GROUP);   // one row
    const int32_t s0 = y == 0 ? gr * (int32_t)(L.ny - 1) : -gr, s2 = y == L.ny - 1 ? -gr * (int32_t)(L.ny - 1) : gr;
    const int32_t lo = min(min(s0, s2), 0);
    // (held in scalar registers: they are the accesses' scalar offsets)
    r.sy[0] = __builtin_amdgcn_readfirstlane(s0 - lo);
    r.sy[1] = __builtin_amdgcn_readfirstlane(-lo);
    r.sy[2] = __builtin_amdgcn_readfirstlane(s2 - lo);
    // GCN/CDNA raw buffer: 0x00020000 in the last word (32-bit data format),
    // no range limit; every offset from these bases is >= 0 and within one
    // plane (below 2^31 bytes, checked when the lattice is built)
    const uint32_t zs[3] = {z == 0 ? L.nz - 1 : z - 1, z, z == L.nz - 1 ? 0 : z + 1};
    for (int k = 0; k < 3; ++k) {
        const int64_t row0 = ((int64_t)zs[k] * L.ny + y) * gr + lo;
        r.r[k] = __builtin_amdgcn_make_buffer_rsrc((char *)L.pops + (int64_t)ES * row0, 0, -1, 0x00020000);
    }
    return r;
}
template <class C>
__device__ __forceinline__ typename C::E row_load(__amdgpu_buffer_rsrc_t r, uint32_t vo, int32_t so) {
    if constexpr (sizeof(typename C::E) == 2) return __builtin_amdgcn_raw_buffer_load_b16(r, (int32_t)vo, so, 0);
    else return __builtin_bit_cast(typename C::E, __builtin_amdgcn_raw_buffer_load_b32(r, (int32_t)vo, so, 0));
}
template <class C>
__device__ __forceinline__ void row_store(__amdgpu_buffer_rsrc_t r, typename C::E v, uint32_t vo, int32_t so) {
    if constexpr (sizeof(typename C::E) == 2) __builtin_amdgcn_raw_buffer_store_b16(v, r, (int32_t)vo, so, 0);
    else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int32_t)vo, so, 0);
}
// input i of a parity-1 step: slot OPP[i] of cell n - c_i
template <class C, bool RAW = false> __device__ __forceinline__ void gather_pre_rows(const RowAddr &r, typename C::T f[19]) {
    constexpr uint32_t ES = sizeof(typename C::E);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        // the slot's constant offset goes with the lane offset, where it
        // folds into the instruction's immediate offset
        const typename C::E e = row_load<C>(r.r[1 - CZ[i]], r.e[1 - CX[i]] + ES * OPP[i] * POP_GROUP, ES * r.sy[1 - CY[i]]);
        f[i] = RAW ? dec_raw<C>(e) : dec<C>(e);
    });
}
// output i of a parity-1 step: slot i of cell n + c_i (slot 0 of n itself)
template <class C, bool RAW = false>
__device__ __forceinline__ void scatter_post_rows(const RowAddr &r, uint32_t n, const typename C::T f[19], uint32_t seed) {
    constexpr uint32_t ES = sizeof(typename C::E);
    Dither d;
    if constexpr (C::SR) d.init(n, seed);
    sfor<19>([&](auto I) {
        constexpr int i = I;
        const uint32_t di = C::SR ? d.template d16<i>() : 0u;
        row_store<C>(r.r[1 + CZ[i]], RAW ? enc_raw<C>(f[i], di) : enc<C>(f[i], di), r.e[1 + CX[i]] + ES * i * POP_GROUP,
                     ES * r.sy[1 + CY[i]]);
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
    // shader.rs:528-538, 1 - d / w as (w - d) * (1 / w): no division per cell
    const T p[3] = {(T)x, (T)y, (T)z};
    const T hi[3] = {(T)(L.nx - 1) - p[0], (T)(L.ny - 1) - p[1], (T)(L.nz - 1) - p[2]};
    T s = (T)0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T lo = (T)L.sp_lo[k], hw = (T)L.sp_hi[k];
        if (lo > (T)0) s = fmax(s, (lo - p[k]) * (T)L.sp_ilo[k]);
        if (hw > (T)0) s = fmax(s, (hw - hi[k]) * (T)L.sp_ihi[k]);
    }
    return fmin(fmax(s, (T)0), (T)1);
}

template <class T> struct FaceSrc {
    T rho, ux, uy, uz, p[6];
};

// rho, u, Pi_neq of post-collision populations (shader.rs:645-653)
template <class T> __device__ __forceinline__ FaceSrc<T> face_source(const DevLattice &L, const T f[19]) {
    FaceSrc<T> s;
    const Mom<T> mo = moments(f);
    s.rho = (T)1 + mo.rho1;
    s.ux = div_(mo.jx, s.rho) - (T)0.5 * (T)L.grav[0];
    s.uy = div_(mo.jy, s.rho) - (T)0.5 * (T)L.grav[1];
    s.uz = div_(mo.jz, s.rho) - (T)0.5 * (T)L.grav[2];
    const T r3 = div_(mo.rho1, (T)3);
    s.p[0] = mo.sxx - s.rho * s.ux * s.ux - r3;
    s.p[1] = mo.syy - s.rho * s.uy * s.uy - r3;
    s.p[2] = mo.szz - s.rho * s.uz * s.uz - r3;
    s.p[3] = mo.sxy - s.rho * s.ux * s.uy;
    s.p[4] = mo.sxz - s.rho * s.ux * s.uz;
    s.p[5] = mo.syz - s.rho * s.uy * s.uz;
    return s;
}

// ---------------------------------------------------------------- collision
// shader.rs:815-847 (+ Guo forcing :442-455). On exit f holds the
// post-collision f~ (before streaming), rho and u_out as the record saves.
// SCALE: f in and out in units of 1 / SCALE (a power of two; the moments
// are scaled back, the outputs' constants scaled up: the same bits, SCALE
// times over)
template <class T, bool FORCE, bool SPONGE, class Hook, int SCALE_LOG2 = 0>
__device__ __forceinline__ void collide(const DevLattice &L, uint32_t x, uint32_t y, uint32_t z, T f[19], T &rho_out,
                                        T u_out[3], Hook &&tau_hook, T ppost[6]) {
    static_assert(SCALE_LOG2 == 0 || !SPONGE, "the absorbing layer runs in true units");
    constexpr T OS = (T)(1u << SCALE_LOG2);
    Mom<T> m = moments(f);
    if constexpr (SCALE_LOG2 != 0) {
        // by ldexp, not a multiply: a multiply could fuse into the stress
        // terms' multiply-adds below and round them differently
        static_assert(std::is_same_v<T, float>, "FP32 arithmetic");
        auto sc = [](float v) { return __builtin_ldexpf(v, -SCALE_LOG2); };
        m.rho1 = sc(m.rho1); m.jx = sc(m.jx); m.jy = sc(m.jy); m.jz = sc(m.jz);
        m.sxx = sc(m.sxx); m.syy = sc(m.syy); m.szz = sc(m.szz);
        m.sxy = sc(m.sxy); m.sxz = sc(m.sxz); m.syz = sc(m.syz);
    }
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
    // feq + (1-om) fneq, evaluated per +-pair: the even part is shared
    const T usq = (T)1.5 * (ux * ux + uy * uy + uz * uz);
    const T tr = (p[0] + p[1] + p[2]) * ((T)1 / (T)3);
    f[0] = (OS * wgt<T>(0)) * (m.rho1 - rho * usq) + k1 * ((T)4.5 * OS * wgt<T>(0)) * (-tr);
    sfor<9>([&](auto K) {
        constexpr int i = 2 * K + 1;
        constexpr int o = i + 1;
        const T cu = cdot<i>(ux, uy, uz);
        const T ev = (OS * wgt<T>(i)) * (m.rho1 + rho * ((T)4.5 * cu * cu - usq)) +
                     k1 * ((T)4.5 * OS * wgt<T>(i)) * (cpc<i>(p) - tr);
        const T od = ((T)3 * OS * wgt<T>(i)) * rho * cu;
        f[i] = ev + od;
        f[o] = ev - od;
    });
    if constexpr (SPONGE) {
        // absorbing layer: f += sigma sp^2 (feq(1, t) - feq(rho, u)), t the
        // far-field velocity or the cell's own, per +-pair as above
        if (sp > (T)0 && (T)L.sigma > (T)0) {
            const T tx = L.abs_u ? (T)L.u_inf[0] : ux, ty = L.abs_u ? (T)L.u_inf[1] : uy, tz = L.abs_u ? (T)L.u_inf[2] : uz;
            const T tsq = (T)1.5 * (tx * tx + ty * ty + tz * tz);
            const T sg = (T)L.sigma * sp * sp;
            f[0] = f[0] + sg * (wgt<T>(0) * (rho * usq - tsq - m.rho1));
            sfor<9>([&](auto K) {
                constexpr int i = 2 * K + 1;
                constexpr int o = i + 1;
                const T cu = cdot<i>(ux, uy, uz);
                const T ct = cdot<i>(tx, ty, tz);
                const T ev = wgt<T>(i) * (((T)4.5 * ct * ct - tsq) - m.rho1 - rho * ((T)4.5 * cu * cu - usq));
                const T od = ((T)3 * wgt<T>(i)) * (ct - rho * cu);
                f[i] = f[i] + sg * (ev + od);
                f[o] = f[o] + sg * (ev - od);
            });
        }
    }
    if constexpr (FORCE) {
        const T Gx = rho * gx, Gy = rho * gy, Gz = rho * gz;
        const T uG = ux * Gx + uy * Gy + uz * Gz;
        const T fo = (T)1 - (T)0.5 * om;
        sfor<19>([&](auto I) {
            constexpr int i = I;
            const T cu = cdot<i>(ux, uy, uz);
            const T cg = cdot<i>(Gx, Gy, Gz);
            f[i] = f[i] + (OS * wgt<T>(i)) * (fo * ((T)3 * (cg - uG) + (T)9 * cu * cg) - (T)1.5 * ((T)1 - om) * cg);
        });
    }
    rho_out = rho;
    u_out[0] = ux;
    u_out[1] = uy;
    u_out[2] = uz;
    // non-equilibrium stress of the post-collision populations (what the
    // next face_gather computes from them): (1 - 1/tau) Pi, plus the second
    // moment of the Guo term (1 - 1/(2 tau)) (u_a G_b + u_b G_a). Exact
    // without the absorbing layer; sponged lattices gather from f instead.
#pragma unroll
    for (int k = 0; k < 6; ++k) ppost[k] = k1 * p[k];
    if constexpr (FORCE) {
        const T fo = (T)1 - (T)0.5 * om;
        const T Gx = rho * gx, Gy = rho * gy, Gz = rho * gz;
        ppost[0] = ppost[0] + fo * (T)2 * ux * Gx;
        ppost[1] = ppost[1] + fo * (T)2 * uy * Gy;
        ppost[2] = ppost[2] + fo * (T)2 * uz * Gz;
        ppost[3] = ppost[3] + fo * (ux * Gy + uy * Gx);
        ppost[4] = ppost[4] + fo * (ux * Gz + uz * Gx);
        ppost[5] = ppost[5] + fo * (uy * Gz + uz * Gy);
    }
}

// the face-source state after collide() (rho, u unchanged by the collision
// unless the absorbing layer acts)
template <class T, bool SPONGE>
__device__ __forceinline__ FaceSrc<T> post_state(const DevLattice &L, const T f[19], T rho, const T u[3], const T ppost[6], bool src) {
    FaceSrc<T> s;
    if constexpr (SPONGE) {
        if (src) return face_source(L, f);
        s.rho = (T)1; s.ux = s.uy = s.uz = (T)0;
#pragma unroll
        for (int k = 0; k < 6; ++k) s.p[k] = (T)0;
        return s;
    } else {
        s.rho = rho;
        s.ux = u[0];
        s.uy = u[1];
        s.uz = u[2];
#pragma unroll
        for (int k = 0; k < 6; ++k) s.p[k] = ppost[k];
        return s;
    }
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

// ---------------------------------------------------------------- faces
// face_gather + face_write (shader.rs:630-687). A face cell's outgoing
// populations are rebuilt from the post-collision state of its inner
// neighbour m (clamped on the non-periodic axes): rho, u and Pi_neq, with
// the inlet/outlet rules applied. Only the populations some colliding cell
// reads are written: direction i of face cell b is live when b + c_i is an
// interior cell (1..n-2 on every non-periodic axis). The other slots the
// reference also writes are read by nobody.
// The update for step k+1 is done by m right after its own collision in
// step k: the live face slots are read and written by no other thread of
// step k. Stores of face cells round with the reference's per-store hash
// (key = cell * 19 + slot direction, shader.rs:510-512).

// velocity set for a direction index known only at run time (2-bit fields)
constexpr uint64_t pack_c(const int (&c)[19]) {
    uint64_t v = 0;
    for (int i = 0; i < 19; ++i) v |= (uint64_t)(c[i] + 1) << (2 * i);
    return v;
}
constexpr uint64_t PCX = pack_c(CX), PCY = pack_c(CY), PCZ = pack_c(CZ);
__device__ __forceinline__ int rcx(int i) { return (int)((PCX >> (2 * i)) & 3u) - 1; }
__device__ __forceinline__ int rcy(int i) { return (int)((PCY >> (2 * i)) & 3u) - 1; }
__device__ __forceinline__ int rcz(int i) { return (int)((PCZ >> (2 * i)) & 3u) - 1; }
__device__ __forceinline__ int ropp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }
template <class T> __device__ __forceinline__ T rwgt(int i) {
    return i == 0 ? (T)(1.0 / 3.0) : (i <= 6 ? (T)(1.0 / 18.0) : (T)(1.0 / 36.0));
}

// the face cell's macroscopic state from its source's (shader.rs:670-681)
template <class T> __device__ __forceinline__ void face_rule(const DevLattice &L, uint32_t ty, FaceSrc<T> &s) {
#pragma clang fp contract(off)
    if (ty == 2u) {
        if (L.total_p) {
            const T ui2 = (T)(L.u_inf[0] * L.u_inf[0] + L.u_inf[1] * L.u_inf[1] + L.u_inf[2] * L.u_inf[2]);
            s.rho = div_((T)1 / (T)3 + (T)0.5 * ui2, (T)1 / (T)3 + (T)0.5 * (s.ux * s.ux + s.uy * s.uy + s.uz * s.uz));
        } else {
            s.ux = (T)L.u_inf[0];
            s.uy = (T)L.u_inf[1];
            s.uz = (T)L.u_inf[2];
        }
    } else {
        s.rho = (T)1;
    }
}

// feq_i(rho, u) - w_i + fneq_i(Pi) for one direction (shader.rs:367-396, 682-685).
// No fused multiply-adds: a face cell's value is then the same whichever
// lane or loop writes it (a lane per direction, or one lane looping over
// directions with the weights and velocities folded to constants).
template <class T> __device__ __forceinline__ T face_value(int i, const FaceSrc<T> &s) {
#pragma clang fp contract(off)
    const T cx = (T)rcx(i), cy = (T)rcy(i), cz = (T)rcz(i), w = rwgt<T>(i);
    const T cu = cx * s.ux + cy * s.uy + cz * s.uz;
    const T usq = (T)1.5 * (s.ux * s.ux + s.uy * s.uy + s.uz * s.uz);
    const T rho1 = s.rho - (T)1;
    const T fe = w * (rho1 + s.rho * ((T)3 * cu + (T)4.5 * cu * cu - usq));
    const T tr = (s.p[0] + s.p[1] + s.p[2]) * ((T)1 / (T)3);
    const T q = cx * cx * s.p[0] + cy * cy * s.p[1] + cz * cz * s.p[2] +
                (T)2 * (cx * cy * s.p[3] + cx * cz * s.p[4] + cy * cz * s.p[5]);
    return fe + (T)4.5 * w * (q - tr);
}

// is direction i of face cell (bx, by, bz) read by a colliding cell?
__device__ __forceinline__ bool face_live(const DevLattice &L, int bx, int by, int bz, int i) {
    const int tx = bx + rcx(i), ty = by + rcy(i), tz = bz + rcz(i);
    return (L.per[0] || (tx >= 1 && tx <= (int)L.nx - 2)) && (L.per[1] || (ty >= 1 && ty <= (int)L.ny - 2)) &&
           (L.per[2] || (tz >= 1 && tz <= (int)L.nz - 2));
}

// store direction i of face cell b into its output slot of parity Q
// (store_prev_out, shader.rs:273-305), dither from the per-store hash
template <class C, int Q>
__device__ __forceinline__ void face_store(const DevLattice &L, int bx, int by, int bz, int i, typename C::T v, uint32_t seed) {
    const uint32_t b = ((uint32_t)bz * L.ny + (uint32_t)by) * L.nx + (uint32_t)bx;
    uint32_t cell, dir;
    if (Q == 0 || i == 0) {
        cell = b;
        dir = (uint32_t)ropp(i);
    } else {
        int tx = bx + rcx(i), ty = by + rcy(i), tz = bz + rcz(i);
        tx = tx < 0 ? tx + (int)L.nx : (tx >= (int)L.nx ? tx - (int)L.nx : tx);
        ty = ty < 0 ? ty + (int)L.ny : (ty >= (int)L.ny ? ty - (int)L.ny : ty);
        tz = tz < 0 ? tz + (int)L.nz : (tz >= (int)L.nz ? tz - (int)L.nz : tz);
        cell = ((uint32_t)tz * L.ny + (uint32_t)ty) * L.nx + (uint32_t)tx;
        dir = (uint32_t)i;
    }
    const uint32_t d16 = C::SR ? (cell_hash(b * 19u + dir, seed) >> 16) : 0u;
    pop<C>(L, (int)dir, cell) = enc<C>(v, d16);
}

// Thread-serial face update of one face cell (standalone kernel)
template <class C, int Q>
__device__ __forceinline__ void face_cell_write(const DevLattice &L, int bx, int by, int bz, FaceSrc<typename C::T> s, uint32_t seed) {
    const uint32_t b = ((uint32_t)bz * L.ny + (uint32_t)by) * L.nx + (uint32_t)bx;
    face_rule(L, L.flags[b] & 3u, s);
    for (int i = 0; i < 19; ++i)
        if (face_live(L, bx, by, bz, i)) face_store<C, Q>(L, bx, by, bz, i, face_value(i, s), seed);
}

// Does a colliding cell at (x, y, z) feed face cells? (it is the clamp of
// some face cell: coordinate 1 or n-2 on a non-periodic axis)
__device__ __forceinline__ bool is_face_source(const DevLattice &L, uint32_t x, uint32_t y, uint32_t z) {
    return (!L.per[0] && (x == 1 || x == L.nx - 2)) || (!L.per[1] && (y == 1 || y == L.ny - 2)) ||
           (!L.per[2] && (z == 1 || z == L.nz - 2));
}
// the face cells: every cell on a face of a non-periodic axis (lattice.rs:283-324)
__device__ __forceinline__ bool is_face_cell(const DevLattice &L, uint32_t x, uint32_t y, uint32_t z) {
    return (!L.per[0] && (x == 0 || x == L.nx - 1)) || (!L.per[1] && (y == 0 || y == L.ny - 1)) ||
           (!L.per[2] && (z == 0 || z == L.nz - 1));
}

// value of lane l (uniform across the wave) for every lane
__device__ __forceinline__ int read_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float read_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(u & 0xffffffffull), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}


// Wave-cooperative face update after a step of parity Q: for every lane
// that is a face source, each face cell it feeds is written by 19 lanes of
// the wave, one direction each. All lanes of the wave must call this.
template <class C, int Q>
__device__ __forceinline__ void face_writes_wave(const DevLattice &L, bool src, uint32_t x, uint32_t y, uint32_t z,
                                                 const FaceSrc<typename C::T> &s, uint32_t seed) {
    using T = typename C::T;
    uint64_t bal = __ballot(src);
    if (__popcll(bal) > 4) {
        // many sources in this wave (a row along a face, or a run of wall
        // records there): each writes its own face cells
        if (src) {
            const bool lo[3] = {!L.per[0] && x == 1, !L.per[1] && y == 1, !L.per[2] && z == 1};
            const bool hi[3] = {!L.per[0] && x == L.nx - 2, !L.per[1] && y == L.ny - 2, !L.per[2] && z == L.nz - 2};
            for (int dz = -1; dz <= 1; ++dz) {
                if ((dz < 0 && !lo[2]) || (dz > 0 && !hi[2])) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    if ((dy < 0 && !lo[1]) || (dy > 0 && !hi[1])) continue;
                    for (int dx = -1; dx <= 1; ++dx) {
                        if ((dx < 0 && !lo[0]) || (dx > 0 && !hi[0])) continue;
                        if (dx == 0 && dy == 0 && dz == 0) continue;
                        face_cell_write<C, Q>(L, (int)x + dx, (int)y + dy, (int)z + dz, s, seed);
                    }
                }
            }
        }
        return;
    }
    const int lane = (int)__lane_id();
    while (bal) {
        // the source lane is the same for the whole wave: read its values
        // with v_readlane (no LDS round trip)
        const int l = __builtin_ctzll(bal);
        bal &= bal - 1;
        const int sx = read_lane((int)x, l), sy = read_lane((int)y, l), sz = read_lane((int)z, l);
        FaceSrc<T> m;
        m.rho = read_lane(s.rho, l);
        m.ux = read_lane(s.ux, l);
        m.uy = read_lane(s.uy, l);
        m.uz = read_lane(s.uz, l);
#pragma unroll
        for (int k = 0; k < 6; ++k) m.p[k] = read_lane(s.p[k], l);
        const bool lo[3] = {!L.per[0] && sx == 1, !L.per[1] && sy == 1, !L.per[2] && sz == 1};
        const bool hi[3] = {!L.per[0] && sx == (int)L.nx - 2, !L.per[1] && sy == (int)L.ny - 2, !L.per[2] && sz == (int)L.nz - 2};
        for (int dz = -1; dz <= 1; ++dz) {
            if ((dz < 0 && !lo[2]) || (dz > 0 && !hi[2])) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                if ((dy < 0 && !lo[1]) || (dy > 0 && !hi[1])) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dx < 0 && !lo[0]) || (dx > 0 && !hi[0])) continue;
                    if (dx == 0 && dy == 0 && dz == 0) continue;
                    const int bx = sx + dx, by = sy + dy, bz = sz + dz;
                    FaceSrc<T> fb = m;
                    // the face cell's type (its flag & 3, set from its
                    // position): velocity if any of its faces is one
                    const bool vel = (dx < 0 && L.vel_face[0]) || (dx > 0 && L.vel_face[1]) ||
                                     (dy < 0 && L.vel_face[2]) || (dy > 0 && L.vel_face[3]) ||
                                     (dz < 0 && L.vel_face[4]) || (dz > 0 && L.vel_face[5]);
                    face_rule(L, vel ? 2u : 3u, fb);
                    if (lane < 19 && face_live(L, bx, by, bz, lane))
                        face_store<C, Q>(L, bx, by, bz, lane, face_value(lane, fb), seed);
                }
            }
        }
    }
}

// Standalone face update before a step of parity P (the reference's
// face_gather + face_write in module P): gathers the stored post-collision
// state of each face cell's inner neighbour. Used at the start of every
// `run`, and before every step where the update cannot be fused.
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
    FaceSrc<T> s;
    if ((L.flags[m] & 3u) == 0) {
        Nbr nm;
        nm.make(L, mx, my, mz, m);
        T f[19];
        gather_post<C, 1 - P>(L, nm, f);
        s = face_source(L, f);
    } else {
        s.rho = (T)1;
        s.ux = (T)L.u_inf[0];
        s.uy = (T)L.u_inf[1];
        s.uz = (T)L.u_inf[2];
#pragma unroll
        for (int j = 0; j < 6; ++j) s.p[j] = (T)0;
    }
    face_cell_write<C, 1 - P>(L, (int)x, (int)y, (int)z, s, seed);
}

// Halfway bounce-back without a record (cells flagged FLAG_HALFWAY): the
// population a cell sent toward the wall last step is read back from the
// slot it was written to (load_prev_out for direction i: (n, opp i) after a
// parity-0 step, (n + c_i, i) after a parity-1 step). The wall neighbour is
// not a face cell and does not collide, so no other thread touches that slot
// in this step. With q = 1/2 the Bouzidi rule (shader.rs:419) reduces to
// f_opp = f_post,i - 6 w (c . u_w), and the momentum exchange is the same
// (c - u_w) f_post,i + (c + u_w) f_b (shader.rs:423-424).
// wall velocity (omega x x_w) at the halfway point of link i, and c . u_w
template <class T> struct HalfwayLink {
    T cx, cy, cz, xwx, xwy, xwz, uwx, uwy, uwz, cu;
};
template <int i, class T>
__device__ __forceinline__ HalfwayLink<T> halfway_link(const DevLattice &L, const StepArgs &a, const Nbr &nb) {
    HalfwayLink<T> h;
    h.cx = (T)CX[i];
    h.cy = (T)CY[i];
    h.cz = (T)CZ[i];
    h.xwx = (T)nb.x - (T)L.center[0] + (T)0.5 * h.cx;
    h.xwy = (T)nb.y - (T)L.center[1] + (T)0.5 * h.cy;
    h.xwz = (T)nb.z - (T)L.center[2] + (T)0.5 * h.cz;
    const T omx = (T)a.omega[0], omy = (T)a.omega[1], omz = (T)a.omega[2];
    h.uwx = omy * h.xwz - omz * h.xwy;
    h.uwy = omz * h.xwx - omx * h.xwz;
    h.uwz = omx * h.xwy - omy * h.xwx;
    h.cu = h.cx * h.uwx + h.cy * h.uwy + h.cz * h.uwz;
    return h;
}

// f_post,i of the previous step, read from where it was stored
template <class C, int P, int i>
__device__ __forceinline__ typename C::T halfway_fpi(const DevLattice &L, const Nbr &nb) {
    if constexpr (P == 0) return dec<C>(pop<C>(L, i, nb.template at<CX[i], CY[i], CZ[i]>()));
    else return dec<C>(pop<C>(L, OPP[i], nb.n));
}

// All 18 slots are loaded, linked or not (each is inside the buffer: Nbr
// wraps at the domain ends), so the loads go out together instead of one
// round trip per link; the unlinked ones are dropped.
template <class C, int P>
__device__ __forceinline__ void halfway_fp(const DevLattice &L, const Nbr &nb, typename C::T fp[19]) {
    sfor<18>([&](auto I) {
        constexpr int i = I + 1;
        fp[i] = halfway_fpi<C, P, i>(L, nb);
    });
}

// OS: f in units of 1 / OS (the scaled step)
template <class C, int P, int SCALE_LOG2 = 0>
__device__ __forceinline__ void halfway_links(const DevLattice &L, const StepArgs &a, const Nbr &nb, uint32_t m,
                                              typename C::T f[19]) {
    using T = typename C::T;
    constexpr T OS = (T)(1u << SCALE_LOG2);
    T fp[19];
    halfway_fp<C, P>(L, nb, fp);
    sfor<18>([&](auto I) {
        constexpr int i = I + 1;
        if (m & (1u << i)) {
            const T wi = wgt<T>(i);
            const HalfwayLink<T> h = halfway_link<i, T>(L, a, nb);
            const T fpi = fp[i] + wi;
            const T v = (fpi - (T)6 * wi * h.cu) - wi;
            f[OPP[i]] = SCALE_LOG2 ? v * OS : v;
        }
    });
}

// the momentum exchange of those links (the slots read above are unchanged
// by this step's own stores)
template <class C, int P>
__device__ __forceinline__ void halfway_force(const DevLattice &L, const StepArgs &a, const Nbr &nb, uint32_t m,
                                              typename C::T F[3], typename C::T Tq[3]) {
    using T = typename C::T;
    T fp[19];
    halfway_fp<C, P>(L, nb, fp);
    sfor<18>([&](auto I) {
        constexpr int i = I + 1;
        if (m & (1u << i)) {
            const T wi = wgt<T>(i);
            const HalfwayLink<T> h = halfway_link<i, T>(L, a, nb);
            const T fpi = fp[i] + wi;
            const T fb = fpi - (T)6 * wi * h.cu;
            const T flx = (h.cx - h.uwx) * fpi + (h.cx + h.uwx) * fb;
            const T fly = (h.cy - h.uwy) * fpi + (h.cy + h.uwy) * fb;
            const T flz = (h.cz - h.uwz) * fpi + (h.cz + h.uwz) * fb;
            F[0] = F[0] + flx;
            F[1] = F[1] + fly;
            F[2] = F[2] + flz;
            Tq[0] = Tq[0] + (h.xwy * flz - h.xwz * fly);
            Tq[1] = Tq[1] + (h.xwz * flx - h.xwx * flz);
            Tq[2] = Tq[2] + (h.xwx * fly - h.xwy * flx);
        }
    });
}

// Block sum of F, T into the block's partial slot of this step; the sums
// over blocks are formed after the chunk of steps by k_reduce_hist (in block
// order: deterministic), so no fence or atomic is needed here.
__device__ __forceinline__ void reduce_ft(const DevLattice &L, const StepArgs &a, double v[6], double (*sh)[6],
                                          uint32_t tid, uint32_t bs, uint32_t nblocks, uint32_t bid) {
    block_sum6(v, sh, tid, bs);
    if (tid == 0) {
        double *o = &L.partials[((size_t)a.slot * nblocks + bid) * 8];
#pragma unroll
        for (int j = 0; j < 6; ++j) o[j] = v[j];
    }
}

#if SIMLAB_KSTORE < 0
// force history of a chunk: block s sums the partials of step s
__global__ void __launch_bounds__(256) k_reduce_hist(const double *partials, uint32_t nblocks, double *hist) {
    __shared__ double sh[16][6];
    const uint32_t s = blockIdx.x, tid = threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t b = tid; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int j = 0; j < 6; ++j) v[j] += partials[((size_t)s * nblocks + b) * 8 + j];
    block_sum6(v, sh, tid, blockDim.x);
    if (tid == 0)
#pragma unroll
        for (int j = 0; j < 6; ++j) hist[(size_t)s * HIST_STRIDE + j] = v[j];
}
#endif

// ---------------------------------------------------------------- special cells
// k_special processes, for a step of parity P, the wall-record cells, one
// thread per record (blocks [0, wall_blocks)), and the momentum exchange of
// the halfway wall cells (the blocks after those): Bouzidi / moving-wall links with momentum exchange
// (shader.rs:401-436, 808-814), the collision, the record save (:848-860);
// then the FP64 force/torque reduction of all records (wave shuffles -> LDS
// -> per-block partial -> the last block to finish sums the partials in
// block order: deterministic).
// A cell that is the inner neighbour of face cells (here or in k_bulk) also
// rebuilds their outgoing populations for the next step from its own
// post-collision state, which is what the next step's face_gather would
// read: the face-cell slots it writes are touched by no other thread of
// this step.
// block `bid` of a lattice's special grid (SPECIAL_BLOCK threads)
template <class C, int P, bool FORCE, bool SPONGE>
__device__ __forceinline__ void special_block(const DevLattice &L, const StepArgs &a, uint32_t bid, double (*sh)[6]) {
    using T = typename C::T;
    const uint32_t tid = threadIdx.x, bs = SPECIAL_BLOCK;
    const uint32_t seed_next = a.seed + 2654435761u;   // lattice.rs:676 seed of the next step
    const uint32_t nblocks = L.wall_blocks + L.hw_blocks;
    if (bid >= L.wall_blocks) {
        // momentum exchange of the halfway wall cells (their bounce-back is
        // done by k_bulk): reads only last step's outputs, which no thread
        // of this step writes
        const uint32_t j = (bid - L.wall_blocks) * bs + tid;
        T F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
        if (j < L.nhw) {
            const uint32_t n = L.hw_cells[j];
            const uint32_t fg = L.flags[n];
            uint32_t x, y, z;
            coords(L, n, x, y, z);
            Nbr nb;
            nb.make(L, x, y, z, n);
            halfway_force<C, P>(L, a, nb, (fg >> 1) & 0x7FFFEu, F, Tq);
        }
        double v[6] = {(double)F[0], (double)F[1], (double)F[2], (double)Tq[0], (double)Tq[1], (double)Tq[2]};
        reduce_ft(L, a, v, sh, tid, bs, nblocks, bid);
        return;
    }
    const uint32_t wb = bid;
    const uint32_t r = wb * bs + tid;
    T F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
    T pp[6];
    bool src = false;
    FaceSrc<T> fsrc;
    uint32_t sx = 0, sy = 0, sz = 0;
    T *rfpost = (T *)L.rfpost;
    T *rmac = (T *)L.rmac;
    T *rft = (T *)L.rft;
    const size_t NR = L.nrec;
    if (r < L.nrec) {
        // a spinning part's records are its swept box, x fastest: the cell
        // follows from r, and the flag load does not wait for a cell load
        uint32_t n;
        if (L.moving) {
            const uint32_t BX = L.box_n[0], BY = L.box_n[1];
            n = ((r / (BX * BY) + (uint32_t)L.box_lo[2]) * L.ny + (r / BX) % BY + (uint32_t)L.box_lo[1]) * L.nx + r % BX +
                (uint32_t)L.box_lo[0];
        } else {
            n = L.rcell[r];
        }
        const uint32_t fg = L.flags[n];
        if ((fg & 3u) == 0 && rec_of(fg) == r + 1) {
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
                // record loads go per link: few registers, so the one-kernel
                // step keeps the bulk tiles' occupancy; the latency is hidden
                // behind the bulk blocks that run beside these
                // the q < 1/2 rule reads the link's own incoming f_i only when
                // the opposite link is not walled, and then no link writes f_i:
                // f[i] still holds it, no copy of the inputs is needed
                sfor<18>([&](auto I) {
                    constexpr int i = I + 1;
                    constexpr int o = OPP[i];
                    if (mask & (1u << i)) {
                        const T wi = wgt<T>(i);
                        const uint32_t qwi = L.rq[(size_t)((i - 1) >> 1) * NR + r];
                        const uint16_t qh = (uint16_t)(((i - 1) & 1) ? (qwi >> 16) : (qwi & 0xffffu));
                        const T q = fmax((T)(float)__builtin_bit_cast(_Float16, qh), (T)0.001);
                        const T fpi = rfpost[(size_t)(i - 1) * NR + r] + wi;
                        const T fpo = rfpost[(size_t)(o - 1) * NR + r] + wi;
                        const T cx = (T)CX[i], cy = (T)CY[i], cz = (T)CZ[i];
                        const T xwx = pcx + q * cx, xwy = pcy + q * cy, xwz = pcz + q * cz;
                        const T uwx = omy * xwz - omz * xwy, uwy = omz * xwx - omx * xwz, uwz = omx * xwy - omy * xwx;
                        const T cu = cx * uwx + cy * uwy + cz * uwz;
                        T fb;
                        if (q >= (T)0.5) fb = div_(fpi + ((T)2 * q - (T)1) * fpo - (T)6 * wi * cu, (T)2 * q);
                        else if (!(mask & (1u << o))) fb = (T)2 * q * fpi + ((T)1 - (T)2 * q) * (f[i] + wi) - (T)6 * wi * cu;
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
            collide<T, FORCE, SPONGE>(L, x, y, z, f, rho, u, hook, pp);
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
            src = is_face_source(L, x, y, z);
            fsrc = post_state<T, SPONGE>(L, f, rho, u, pp, src);
            sx = x;
            sy = y;
            sz = z;
        } else if (fg != 0u) {
            // a solid cell of a spinning part (the momentum it took when
            // covered), or a record another cell holds. A plain fluid cell
            // (flags 0: a box cell far from the part) has F = T = 0: its
            // record was reset when it left the part's neighbourhood and is
            // not written while it is far
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                F[k] = rft[k * NR + r];
                Tq[k] = rft[(3 + k) * NR + r];
            }
        }
    }
    if (__ballot(src)) face_writes_wave<C, P>(L, src, sx, sy, sz, fsrc, seed_next);
    double v[6] = {(double)F[0], (double)F[1], (double)F[2], (double)Tq[0], (double)Tq[1], (double)Tq[2]};
    reduce_ft(L, a, v, sh, tid, bs, nblocks, bid);
}

template <class C, int P, bool FORCE, bool SPONGE>
__global__ void __launch_bounds__(SPECIAL_BLOCK) k_special(DevLattice L, StepArgs a) {
    __shared__ double sh[16][6];
    special_block<C, P, FORCE, SPONGE>(L, a, blockIdx.x, sh);
}

// ---------------------------------------------------------------- bulk
// Block b of nb taken in XCD order: the GPU deals consecutive workgroups to
// its 8 XCDs in turn (b % 8); this gives XCD k the k-th contiguous run of
// the nb tiles, so neighbouring tiles (which share the cache lines that
// straddle their x-edges in a parity-1 step) meet in one L2. A bijection on
// [0, nb).
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t nb) {
    const uint32_t xcd = b & 7u, j = b >> 3, q = nb >> 3, r = nb & 7u;
    return (xcd < r ? xcd * (q + 1u) : r * (q + 1u) + (xcd - r) * q) + j;
}

// the row-aligned parity-1 variant where the lattice allows it (16- and
// 32-bit storage)
template <class C, int P> constexpr bool rows_variant() {
    return P == 1 && (std::is_same_v<C, CF16> || std::is_same_v<C, CBF16> || std::is_same_v<C, CF32>);
}

// Every fluid cell with no wall record (face sources and halfway wall cells
// included). Grid:
// (xblocks * yblocks, nz); block (bx, by) = `by` whole x-rows (split in
// xblocks pieces when nx > bx) at one z, so the +-1 x-offsets of parity 1
// stay inside the block and waves touch whole cache lines.
// tile bxy (of xblocks * yblocks) of z-plane z; thread tid of the tile's
// bx * by threads (x fastest); threads past the tile idle
template <class C, int P, bool FORCE, bool SPONGE, bool ROWS = false>
__device__ __forceinline__ void bulk_block(const DevLattice &L, const StepArgs &a, uint32_t bxy, uint32_t z, uint32_t tid) {
    using T = typename C::T;
    static_assert(!ROWS || P == 1, "row-aligned addressing is for the parity-1 step");
    // A wave is 64 consecutive x of one row (bx is a multiple of 64, tiles
    // are whole waves): the tile, the row and the plane are the same for all
    // its lanes, so they (and the row and plane neighbour offsets) live in
    // scalar registers; only x differs per lane.
    bxy = __builtin_amdgcn_readfirstlane(bxy);
    z = __builtin_amdgcn_readfirstlane(z);
    const uint32_t xb = bxy % L.xblocks, yb = bxy / L.xblocks;
    const uint32_t tx = tid % L.bx, ty = __builtin_amdgcn_readfirstlane(tid / L.bx);
    const uint32_t x = xb * L.bx + tx;
    const uint32_t y = yb * L.by + ty;
    const bool inside = x < L.nx && y < L.ny && ty < L.by;
    const uint32_t n = (z * L.ny + y) * L.nx + x;
    // a wave is 64 consecutive cells of one row (bx is a multiple of 64):
    // where they all lie in plain-fluid groups the flags are known to be 0
    uint32_t fg;
    if (L.plain && __all(inside)) {
        const uint32_t n0 = __builtin_amdgcn_readfirstlane(n);
        const uint32_t g0 = n0 / POP_GROUP, g1 = (n0 + 63u) / POP_GROUP;
        const bool pl = ((L.plain[g0 >> 5] >> (g0 & 31u)) & 1u) && ((L.plain[g1 >> 5] >> (g1 & 31u)) & 1u);
        fg = pl ? 0u : L.flags[n];
    } else {
        fg = inside ? L.flags[n] : 1u;
    }
    // The populations are loaded without waiting for the flag, so the two
    // round trips overlap: every cell that may collide (all but the face
    // cells, known from the coordinates) loads, every index being inside the
    // buffer (Nbr wraps at the domain ends); solid and wall-record cells
    // among them drop what they loaded.
    const bool may = inside && !is_face_cell(L, x, y, z);
    Nbr nb;
    nb.make(L, x, y, z, n);
    RowAddr ra;
    if constexpr (ROWS) ra = row_addr<C>(L, x, y, z);
    // FP16 without the absorbing layer steps in the stored scale (x 2^15):
    // no multiply per load or store, the same bits
    constexpr bool RAW = std::is_same_v<C, CF16> && !SPONGE;
    constexpr int SL = RAW ? 15 : 0;
    T f[19];
    if (may) {
        if constexpr (ROWS) gather_pre_rows<C, RAW>(ra, f);
        else gather_pre<C, P, RAW>(L, nb, f);
    }
    // ours: fluid cells without a record (plain or halfway wall cells);
    // solid, face and wall-record cells are not
    const bool mine = may && (fg & 3u) == 0 && rec_of(fg) == 0;
    const uint32_t hwm = (fg & FLAG_HALFWAY) ? ((fg >> 1) & 0x7FFFEu) : 0u;
    T rho, u[3], pp[6];
    if (mine) {
        if (hwm) halfway_links<C, P, SL>(L, a, nb, hwm, f);
        collide<T, FORCE, SPONGE, NoHook, SL>(L, x, y, z, f, rho, u, NoHook{}, pp);
        if constexpr (ROWS) scatter_post_rows<C, RAW>(ra, n, f, a.seed);
        else scatter_post<C, P, false, false, RAW>(L, nb, f, a.seed);
    }
    const bool src = mine && is_face_source(L, x, y, z);
    if (__ballot(src)) {
        // the source's post-collision state from the collision's own
        // rho, u and stress (the moments of f where the layer acts)
        FaceSrc<T> s;
        if (src) s = post_state<T, SPONGE>(L, f, rho, u, pp, src);
        // a source at x = 1 or nx - 2 that feeds no y or z face cell leaves
        // its state for k_xfaces instead of holding its wave for one lane
        bool xs = false;
        if constexpr (std::is_same_v<T, float>) {
            xs = src && L.xdefer && (x == 1 || x == L.nx - 2) && (L.per[1] || (y != 1 && y != L.ny - 2)) &&
                 (L.per[2] || (z != 1 && z != L.nz - 2));
            if (xs) {
                float4 *o = (float4 *)(L.xsrc + ((size_t)((x == 1 ? 0u : 1u) * L.nz + z) * L.ny + y) * 12);
                o[0] = make_float4(s.rho, s.ux, s.uy, s.uz);
                o[1] = make_float4(s.p[0], s.p[1], s.p[2], s.p[3]);
                // the step's tag, as bits
                *(uint4 *)(o + 2) = make_uint4(__float_as_uint(s.p[4]), __float_as_uint(s.p[5]), a.mark_lo, a.mark_hi);
            }
        }
        const bool coop = src && !xs;
        if (__ballot(coop)) face_writes_wave<C, P>(L, coop, x, y, z, s, a.seed + 2654435761u);
    }
}

// The x-face cells of the sources that left their state in xsrc: one thread
// per face cell, for its five directions into the domain (c_x = +1 on the -x
// face: 1, 7, 9, 13, 15; c_x = -1 on the +x face: 2, 8, 10, 14, 16), the
// writes face_writes_wave would have done (same rule, values, slots and
// dither).
template <class C, int Q>
__global__ void __launch_bounds__(256) k_xfaces(DevLattice L, uint32_t seed, uint32_t mark_lo, uint32_t mark_hi) {
    using T = typename C::T;
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ne = 2u * L.nz * L.ny;
    if (e >= ne) return;
    const uint32_t side = e / (L.nz * L.ny), r = e - side * L.nz * L.ny;
    const uint32_t z = r / L.ny, y = r - z * L.ny;
    // the source must be a colliding x-only source of the bulk kernel
    if (!(L.per[1] || (y >= 1 && y <= L.ny - 2 && y != 1 && y != L.ny - 2))) return;
    if (!(L.per[2] || (z >= 1 && z <= L.nz - 2 && z != 1 && z != L.nz - 2))) return;
    // a colliding x-only source of this step left its state with this
    // step's tag (one memory round trip, no flag load)
    const float4 *o = (const float4 *)(L.xsrc + (size_t)e * 12);
    const float4 a0 = o[0], a1 = o[1];
    const uint4 a2 = *(const uint4 *)(o + 2);
    if (a2.z != mark_lo || a2.w != mark_hi) return;
    FaceSrc<T> fb;
    fb.rho = a0.x; fb.ux = a0.y; fb.uy = a0.z; fb.uz = a0.w;
    fb.p[0] = a1.x; fb.p[1] = a1.y; fb.p[2] = a1.z; fb.p[3] = a1.w; fb.p[4] = __uint_as_float(a2.x); fb.p[5] = __uint_as_float(a2.y);
    face_rule(L, L.vel_face[side] ? 2u : 3u, fb);
    const int bx = side ? (int)L.nx - 1 : 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int i = (k == 0 ? 1 : 5 + 2 * k + (k > 2 ? 2 : 0)) + (int)side;
        if (face_live(L, bx, (int)y, (int)z, i)) face_store<C, Q>(L, bx, (int)y, (int)z, i, face_value(i, fb), seed);
    }
}

// the sponge variant of parity 1 in 16-bit storage at 5 waves per SIMD like
// the plain one (it needs one VGPR more than 96 left alone)
template <class C, int P, bool FORCE, bool SPONGE> constexpr int bulk_waves() {
    return ((std::is_same_v<C, CF16> || std::is_same_v<C, CBF16>) && P == 1 && !FORCE && SPONGE) ? 5 : 1;
}
// z-planes of a bulk launch: z = zbase + blockIdx.y, and z += zskip_len from
// zskip_at on (all planes: {0, nz, 0})
struct ZMap {
    uint32_t zbase, zskip_at, zskip_len;
};
template <class C, int P, bool FORCE, bool SPONGE, bool ROWS = false>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(bulk_waves<C, P, FORCE, SPONGE>())))
k_bulk(DevLattice L, StepArgs a, ZMap zm) {
    uint32_t bxy = blockIdx.x, zl = blockIdx.y;
    if (L.xcd_swizzle) {
        const uint32_t lin = xcd_order(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
        bxy = lin % gridDim.x;
        zl = lin / gridDim.x;
    }
    uint32_t z = zm.zbase + zl;
    if (z >= zm.zskip_at) z += zm.zskip_len;
    bulk_block<C, P, FORCE, SPONGE, ROWS>(L, a, bxy, z, threadIdx.y * blockDim.x + threadIdx.x);
}

// One launch per step, one stream: blocks [0, wall_blocks + hw_blocks) do
// the special cells (dispatched first, they run beside the first bulk
// blocks), the others SPECIAL_BLOCK / (bx * by) bulk tiles each (tile
// index = z * xblocks * yblocks + tile in the plane). No second stream and
// no events per step.
// parity 1 with FP16 or FP32 storage and neither body force nor sponge: at
// least 5 waves per SIMD (96 VGPRs), as the bulk kernel's parity 1 has (the
// compiler fits them without spilling); the other variants would spill and
// keep their natural count
template <class C, int P, bool FORCE, bool SPONGE, bool ROWS = false> constexpr int step_waves() {
    if (ROWS && P == 1 && !FORCE) return 6;
    return ((std::is_same_v<C, CF16> || std::is_same_v<C, CF32>) && P == 1 && !FORCE && !SPONGE) ? 5 : 1;
}
template <class C, int P, bool FORCE, bool SPONGE, bool ROWS = false>
__global__ void __launch_bounds__(SPECIAL_BLOCK) __attribute__((amdgpu_waves_per_eu(step_waves<C, P, FORCE, SPONGE, ROWS>())))
k_step(DevLattice L, StepArgs a) {
    __shared__ double sh[16][6];
    const uint32_t nsp = L.wall_blocks + L.hw_blocks;
    if (blockIdx.x < nsp) {
        special_block<C, P, FORCE, SPONGE>(L, a, blockIdx.x, sh);
        return;
    }
    const uint32_t tt = L.bx * L.by;   // divides SPECIAL_BLOCK (checked by the host)
    uint32_t b = blockIdx.x - nsp;
    if (L.xcd_swizzle) b = xcd_order(b, gridDim.x - nsp);
    const uint32_t g = b * (SPECIAL_BLOCK / tt) + threadIdx.x / tt;
    const uint32_t tiles = L.xblocks * L.yblocks;
    if (g >= tiles * L.nz) return;   // whole tiles (whole waves) only
    bulk_block<C, P, FORCE, SPONGE, ROWS>(L, a, g % tiles, g / tiles, threadIdx.x % tt);
}

// ---------------------------------------------------------------- batches
// Several independent lattices in one launch per kernel and step: block b
// of the launch belongs to lattice i with boff[i] <= b < boff[i + 1]; the
// lattices' device views are read from `lats` (the index is uniform per
// block, so they come through the scalar cache). All lattices of a batch
// share storage, forcing/sponge flags and the step index (same seed).
__device__ __forceinline__ uint32_t batch_member(const uint32_t *boff, uint32_t nl, uint32_t b) {
    uint32_t i = 0;
    while (i + 1 < nl && b >= boff[i + 1]) ++i;
    return __builtin_amdgcn_readfirstlane(i);   // uniform: keeps the lattice's fields in scalar registers
}

// SPECIAL_BLOCK threads per block; a lattice's tile (bx * by threads, set
// by the host to divide SPECIAL_BLOCK) fills the block SPECIAL_BLOCK / tile
// times, so no lattice's blocks idle on another lattice's larger tile
// ROWS: every member on row-aligned addressing (parity 1, set by the host)
template <class C, int P, bool FORCE, bool SPONGE, bool ROWS = false>
__global__ void __launch_bounds__(SPECIAL_BLOCK) __attribute__((amdgpu_waves_per_eu(ROWS && !FORCE ? 6 : 1)))
k_bulk_batch(const DevLattice *__restrict__ lats,
                                                              const uint32_t *__restrict__ boff, uint32_t nl, StepArgs a) {
    const uint32_t i = batch_member(boff, nl, blockIdx.x);
    const DevLattice &L = lats[i];
    const uint32_t local = __builtin_amdgcn_readfirstlane(blockIdx.x - boff[i]);
    const uint32_t tt = L.bx * L.by;
    const uint32_t g = local * (SPECIAL_BLOCK / tt) + threadIdx.x / tt;
    const uint32_t tiles = L.xblocks * L.yblocks;
    if (g >= tiles * L.nz) return;   // whole tiles (whole waves) only
    bulk_block<C, P, FORCE, SPONGE, ROWS>(L, a, g % tiles, g / tiles, threadIdx.x % tt);
}

template <class C, int P, bool FORCE, bool SPONGE>
__global__ void __launch_bounds__(SPECIAL_BLOCK) k_special_batch(const DevLattice *__restrict__ lats,
                                                                 const uint32_t *__restrict__ boff, uint32_t nl, StepArgs a) {
    __shared__ double sh[16][6];
    const uint32_t i = batch_member(boff, nl, blockIdx.x);
    special_block<C, P, FORCE, SPONGE>(lats[i], a, blockIdx.x - boff[i], sh);
}

// One launch per step for a whole batch: blocks [0, boff[nl]) are the
// members' special blocks (member i from boff[i]), the rest their bulk
// tiles (member i from bulk[i], bulk = boff + nl + 1, counted after the
// special blocks), each member at its own tile (bx * by divides
// SPECIAL_BLOCK). No second stream and no events per step; the host orders
// the bulk ranges largest member first, so the small members' tiles fill the
// large member's tail.
template <class C, int P, bool FORCE, bool SPONGE, bool ROWS = false>
__global__ void __launch_bounds__(SPECIAL_BLOCK) __attribute__((amdgpu_waves_per_eu(step_waves<C, P, FORCE, SPONGE, ROWS>())))
k_step_batch(const DevLattice *__restrict__ lats, const uint32_t *__restrict__ boff, uint32_t nl, StepArgs a) {
    __shared__ double sh[16][6];
    const uint32_t nsp = boff[nl];
    if (blockIdx.x < nsp) {
        const uint32_t i = batch_member(boff, nl, blockIdx.x);
        special_block<C, P, FORCE, SPONGE>(lats[i], a, blockIdx.x - boff[i], sh);
        return;
    }
    const uint32_t *bulk = boff + nl + 1;
    const uint32_t b = blockIdx.x - nsp;
    const uint32_t i = batch_member(bulk, nl, b);
    const DevLattice &L = lats[i];
    uint32_t local = __builtin_amdgcn_readfirstlane(b - bulk[i]);
    // XCD order within the member's blocks (its first block's XCD only
    // relabels the runs)
    if (L.xcd_swizzle) local = xcd_order(local, bulk[i + 1] - bulk[i]);
    const uint32_t tt = L.bx * L.by;
    const uint32_t g = local * (SPECIAL_BLOCK / tt) + threadIdx.x / tt;
    const uint32_t tiles = L.xblocks * L.yblocks;
    if (g >= tiles * L.nz) return;   // whole tiles (whole waves) only
    bulk_block<C, P, FORCE, SPONGE, ROWS>(L, a, g % tiles, g / tiles, threadIdx.x % tt);
}

// ---------------------------------------------------------------- moving geometry
// Re-voxelization of a spinning part's swept box every step (shader.rs:865-1104):
// tri_scatter finds the lattice links each rotated triangle cuts, voxelize
// classifies every box cell (solid now / before from the body-frame
// occupancy bits), builds its wall links and q, refills cells the part
// uncovered and books the momentum of covered and uncovered cells, qclear
// resets the link scratch. Body-frame arithmetic is in f32 as in the WGSL,
// with no fused multiply-adds, so every lattice finds the same links and q
// bits for the same rotation whatever the compiler schedules.
struct V3f {
    float x, y, z;
};
__device__ __forceinline__ V3f v3(float x, float y, float z) { return V3f{x, y, z}; }
__device__ __forceinline__ V3f operator+(V3f a, V3f b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3f operator-(V3f a, V3f b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot3(V3f a, V3f b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ V3f cross3(V3f a, V3f b) {
#pragma clang fp contract(off)
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// rotation columns r[k] = body axis k: to_body(v) = (r0.v, r1.v, r2.v)
__device__ __forceinline__ V3f to_body(const float (*m)[4], V3f v) {
#pragma clang fp contract(off)
    return v3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}
__device__ __forceinline__ V3f from_body(const float (*m)[4], V3f b) {
#pragma clang fp contract(off)
    return v3(m[0][0] * b.x + m[1][0] * b.y + m[2][0] * b.z, m[0][1] * b.x + m[1][1] * b.y + m[2][1] * b.z,
              m[0][2] * b.x + m[1][2] * b.y + m[2][2] * b.z);
}
__device__ __forceinline__ bool occ_at(const DevLattice &L, V3f b) {
#pragma clang fp contract(off)
    const int gx = (int)__builtin_rintf((b.x - L.occ_lo[0]) * L.occ_invh);
    const int gy = (int)__builtin_rintf((b.y - L.occ_lo[1]) * L.occ_invh);
    const int gz = (int)__builtin_rintf((b.z - L.occ_lo[2]) * L.occ_invh);
    if (gx < 0 || gy < 0 || gz < 0 || gx >= L.occ_n[0] || gy >= L.occ_n[1] || gz >= L.occ_n[2]) return false;
    const uint32_t l = (uint32_t)((gz * L.occ_n[1] + gy) * L.occ_n[0] + gx);
    return ((L.occ[l >> 5] >> (l & 31u)) & 1u) != 0u;
}
__device__ __forceinline__ V3f centerf(const DevLattice &L) { return v3((float)L.center[0], (float)L.center[1], (float)L.center[2]); }
// every cell asks with the same arithmetic for itself and its neighbours
__device__ __forceinline__ bool solid_at(const DevLattice &L, const StepArgs &a, int px, int py, int pz, bool before) {
    const V3f c = centerf(L);
    const V3f pc = v3((float)px - c.x, (float)py - c.y, (float)pz - c.z);
    return occ_at(L, to_body(before ? a.q : a.r, pc));
}
// the cell's body-frame position lies in the marked cells of the near-part
// grid's plane 0 (within NEAR_CELLS) or plane 1 (within NEAR_CELLS + 1)
__device__ __forceinline__ bool near_part(const DevLattice &L, V3f b, int plane = 0) {
    const int gx = (int)__builtin_floorf(b.x - L.near_lo[0]);
    const int gy = (int)__builtin_floorf(b.y - L.near_lo[1]);
    const int gz = (int)__builtin_floorf(b.z - L.near_lo[2]);
    if (gx < 0 || gy < 0 || gz < 0 || gx >= L.near_n[0] || gy >= L.near_n[1] || gz >= L.near_n[2]) return false;
    const uint32_t l = (uint32_t)((gz * L.near_n[1] + gy) * L.near_n[0] + gx);
    return ((L.near[plane * L.near_words + (l >> 5)] >> (l & 31u)) & 1u) != 0u;
}
// box cell k -> lattice coordinates and position relative to the centre
__device__ __forceinline__ void box_cell(const DevLattice &L, uint32_t k, int &px, int &py, int &pz, V3f &pc) {
    const uint32_t BX = L.box_n[0], BY = L.box_n[1];
    px = (int)(k % BX) + L.box_lo[0];
    py = (int)((k / BX) % BY) + L.box_lo[1];
    pz = (int)(k / (BX * BY)) + L.box_lo[2];
    const V3f c = centerf(L);
    pc = v3((float)px - c.x, (float)py - c.y, (float)pz - c.z);
}

// a cleared hit entry: no nearest hit (above every hit's bits, which are
// those of a t in [0, 1]), farthest 0
__device__ __forceinline__ void hits_clear(uint32_t *o) {
#pragma unroll
    for (int w = 0; w < 9; ++w) *(uint2 *)(o + 2 * w) = make_uint2(0x7f7fffffu, 0u);
}

#if SIMLAB_KSTORE < 0
// TRI_LANES lanes per triangle, eight triangles per wave; a triangle's
// lanes share the lattice cells of its bounding box (grown by 1.415, as
// shader.rs:915-927). After subdivision a triangle spans about a cell, so
// that box holds 34 cells on average and at most about 80: a whole wave per
// triangle left half its lanes idle and repeated the triangle's setup on
// all 64 (on the 384^3 prop, 16 lanes per triangle stepped 0.7 % slower
// than 8, 32 lanes 1.7 %). Moller-Trumbore for the link segment
// p + s c_i, s in [0, 1] (shader.rs ray_tri), split into what depends on
// the triangle and the direction only (pv = c_i x e2, e1 . pv and its reciprocal, n . c_i: once
// per lane) and what depends on the cell (tv = p - v0, tv x e1 and its dot
// with e2: once per cell), each the same float operation on the same
// operands in the same order as the test written out per link, so the hits
// and q bits are those of the reference's arithmetic.
constexpr uint32_t TRI_LANES = 8;
__global__ void __launch_bounds__(256) k_tri_scatter(DevLattice L, StepArgs a) {
#pragma clang fp contract(off)
    const uint32_t t = blockIdx.x * (256 / TRI_LANES) + threadIdx.x / TRI_LANES;
    const uint32_t lane = threadIdx.x % TRI_LANES;
    if (t >= L.ntri) return;
    const float4 t0 = L.tris[3 * t], t1 = L.tris[3 * t + 1], t2 = L.tris[3 * t + 2];
    const V3f A = centerf(L) + from_body(a.r, v3(t0.x, t0.y, t0.z));
    const V3f e1 = from_body(a.r, v3(t1.x, t1.y, t1.z));
    const V3f e2 = from_body(a.r, v3(t2.x, t2.y, t2.z));
    const V3f nrm = cross3(e1, e2);
    const float nl = __builtin_sqrtf(dot3(nrm, nrm));
    if (nl < 1e-12f) return;
    const V3f nn = v3(nrm.x / nl, nrm.y / nl, nrm.z / nl);
    const V3f B = A + e1, Cc = A + e2;
    int lo[3], hi[3];
    {
        const float av[3] = {A.x, A.y, A.z}, bv[3] = {B.x, B.y, B.z}, cv[3] = {Cc.x, Cc.y, Cc.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float mn = fminf(fminf(av[k], bv[k]), cv[k]), mx = fmaxf(fmaxf(av[k], bv[k]), cv[k]);
            lo[k] = max((int)__builtin_ceilf(mn - 1.415f), L.box_lo[k]);
            hi[k] = min((int)__builtin_floorf(mx + 1.415f), L.box_lo[k] + (int)L.box_n[k] - 1);
        }
    }
    if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) return;
    // per odd direction i = 2j + 1: n . c_i, pv = c_i x e2, 1 / (e1 . pv)
    float dn[9], inv[9];
    V3f pv[9];
    uint32_t okd = 0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int i = 2 * j + 1;
        const V3f ci = v3((float)rcx(i), (float)rcy(i), (float)rcz(i));
        dn[j] = dot3(nn, ci);
        pv[j] = cross3(ci, e2);
        const float det = dot3(e1, pv[j]);
        inv[j] = 1.0f;
        if (!(__builtin_fabsf(det) < 1e-12f)) {
            inv[j] = 1.0f / det;
            okd |= 1u << j;
        }
    }
    const uint32_t ex = (uint32_t)(hi[0] - lo[0] + 1), ey = (uint32_t)(hi[1] - lo[1] + 1), ez = (uint32_t)(hi[2] - lo[2] + 1);
    for (uint32_t c = lane; c < ex * ey * ez; c += TRI_LANES) {
        const int x = lo[0] + (int)(c % ex), y = lo[1] + (int)((c / ex) % ey), z = lo[2] + (int)(c / (ex * ey));
        const V3f p = v3((float)x, (float)y, (float)z);
        const V3f tv = p - A;
        const float d0 = dot3(nn, tv);
        if (__builtin_fabsf(d0) > 1.415f) continue;
        const uint32_t k = ((uint32_t)(z - L.box_lo[2]) * L.box_n[1] + (uint32_t)(y - L.box_lo[1])) * L.box_n[0] +
                           (uint32_t)(x - L.box_lo[0]);
        const V3f qv = cross3(tv, e1);
        const float tq = dot3(e2, qv);
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int i = 2 * j + 1;
            const float d1 = d0 + dn[j];
            if (d0 * d1 > 0.0f && __builtin_fabsf(d0) > 1e-6f) continue;
            if (!(okd & (1u << j))) continue;
            // ray_tri for the segment p + s c_i, s in [0, 1]
            const float u = dot3(tv, pv[j]) * inv[j];
            if (u < 0.0f || u > 1.0f) continue;
            const V3f ci = v3((float)rcx(i), (float)rcy(i), (float)rcz(i));
            const float v = dot3(ci, qv) * inv[j];
            if (v < 0.0f || u + v > 1.0f) continue;
            float tt = tq * inv[j];
            if (tt < -1e-4f || tt > 1.0f) continue;
            // max(t, 0), with +0 for t = -0: the hit bits are ordered as
            // unsigned integers, which sort non-negative floats only
            tt = tt > 0.0f ? tt : 0.0f;
            uint32_t *o = L.qscr + (((size_t)a.hbuf * L.nrec + k) * 9 + j) * 2;
            const uint32_t tb = __float_as_uint(tt);
            atomicMin(o, tb);
            atomicMax(o + 1, tb);
        }
    }
}

__global__ void __launch_bounds__(256) k_qclear(DevLattice L) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2ull * L.nrec * 9u) return;
    *(uint2 *)(L.qscr + 2 * k) = make_uint2(0x7f7fffffu, 0u);
}
// the entries of a step's half that its triangles can have hit: only cells
// near the part at its rotation (every other entry is still clear)
__global__ void __launch_bounds__(256) k_qclear_near(DevLattice L, StepArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L.nrec) return;
    int px, py, pz;
    V3f pc;
    box_cell(L, k, px, py, pz, pc);
    if (!near_part(L, to_body(a.r, pc))) return;
    hits_clear(L.qscr + ((size_t)a.hbuf * L.nrec + k) * 18);
}
#endif

// f16 round to nearest even of a q in [0, 1] (pack2x16float)
__device__ __forceinline__ uint32_t half_rne_bits(float q) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)q);
}

// the refill's store: this step's inputs of the uncovered cell, written as
// the cell's previous outputs (store_prev_out, guarded, per-store dither)
template <class C, int Q>
__device__ __forceinline__ void refill_store(const DevLattice &L, const Nbr &nb, const typename C::T fe[19], uint32_t seed) {
    sfor<19>([&](auto I) {
        constexpr int i = I;
        uint32_t cell, dir;
        if constexpr (Q == 0 || i == 0) {
            cell = nb.n;
            dir = OPP[i];
        } else {
            if (!keeps_inside<i>(L, nb)) return;
            cell = nb.template at<CX[i], CY[i], CZ[i]>();
            dir = i;
        }
        const uint32_t d16 = C::SR ? (cell_hash(nb.n * 19u + dir, seed) >> 16) : 0u;
        pop<C>(L, dir, cell) = enc<C>(fe[i], d16);
    });
}

template <class C, int P> __global__ void __launch_bounds__(256) k_voxelize(DevLattice L, StepArgs a) {
    using T = typename C::T;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L.nrec) return;
    const uint32_t BX = L.box_n[0], BY = L.box_n[1], BZ = L.box_n[2];
    const uint32_t lx = k % BX, ly = (k / BX) % BY, lz = k / (BX * BY);
    const int px = (int)lx + L.box_lo[0], py = (int)ly + L.box_lo[1], pz = (int)lz + L.box_lo[2];
    const uint32_t n = ((uint32_t)pz * L.ny + (uint32_t)py) * L.nx + (uint32_t)px;
    const V3f c = centerf(L);
    const V3f pc = v3((float)px - c.x, (float)py - c.y, (float)pz - c.z);
    const size_t NR = L.nrec;
    T *rfpost = (T *)L.rfpost;
    T *rmac = (T *)L.rmac;
    T *rft = (T *)L.rft;
    const V3f bnow = to_body(a.r, pc);
    // the cell's body-frame position at the previous voxelization, and
    // whether it was near the part then
    const V3f bprev = to_body(a.vq, pc);
    const bool near_prev = a.vq_valid && near_part(L, bprev);
    // the step before's hits (the other half) are all read: clear its
    // entries near the part then, for the next step's triangles (hclear
    // implies vq_valid)
    if (a.hclear && near_prev) hits_clear(L.qscr + ((size_t)(a.hbuf ^ 1u) * NR + k) * 18);
    if (!near_part(L, bnow)) {
        // far from the part: fluid now and before, no solid neighbour, no
        // hit; the outcome of the full rule below is known. Farther still
        // (outside plane 1) the cell steps as a plain fluid cell in the
        // bulk kernel and its record is left as it is.
        const bool in1 = near_part(L, bnow, 1);
        if (a.vq_valid && !near_prev) {
            // far at the previous voxelization too: its links, q and F, T
            // were reset then (or at an earlier one) and nothing has written
            // them since; only the flag may change (plane 1)
            if (in1 != near_part(L, bprev, 1)) L.flags[n] = in1 ? (k + 1u) << 2 : 0u;
            return;
        }
#pragma unroll
        for (int w = 0; w < 9; ++w) L.rq[(size_t)w * NR + k] = 0x3C003C00u;
        L.rmask[k] = 0u;
#pragma unroll
        for (int j = 0; j < 6; ++j) rft[j * NR + k] = (T)0;
        L.flags[n] = in1 ? (k + 1u) << 2 : 0u;
        return;
    }
    const bool solid_now = solid_at(L, a, px, py, pz, false);
    const bool solid_before = solid_at(L, a, px, py, pz, true);
    const T pcx = (T)pc.x, pcy = (T)pc.y, pcz = (T)pc.z;
    if (solid_now) {
        L.flags[n] = 1u | ((k + 1u) << 2);
        L.rmask[k] = 0u;
        T F[3] = {0, 0, 0};
        if (!solid_before) {
            // a cell the part covered: the part takes its fluid's momentum
            const T rho = rmac[k];
            F[0] = rho * rmac[NR + k];
            F[1] = rho * rmac[2 * NR + k];
            F[2] = rho * rmac[3 * NR + k];
        }
        rft[k] = F[0]; rft[NR + k] = F[1]; rft[2 * NR + k] = F[2];
        rft[3 * NR + k] = pcy * F[2] - pcz * F[1];
        rft[4 * NR + k] = pcz * F[0] - pcx * F[2];
        rft[5 * NR + k] = pcx * F[1] - pcy * F[0];
        return;
    }
    // link hits of this step: a link is kept at the cell it leaves in the
    // positive direction; the cell at the other end reads it mirrored. The
    // reference also counts the hits (shader.rs:998-1016) and asks for two
    // of them below; with one hit the nearest and farthest are the same
    // value, so that their distance alone decides, and the count is not kept
    const uint32_t *hits = L.qscr + (size_t)a.hbuf * NR * 18;
    float qv[19], qx[19];
    bool hit[19];
    for (int i = 1; i < 19; ++i) {
        qv[i] = 9.0f;
        qx[i] = -1.0f;
        hit[i] = false;
        if (i & 1) {
            const uint2 o = *(const uint2 *)(hits + ((size_t)k * 9 + (i - 1) / 2) * 2);
            if (o.x != 0x7f7fffffu) {
                qv[i] = __uint_as_float(o.x);
                qx[i] = __uint_as_float(o.y);
                hit[i] = true;
            }
        } else {
            const int qx_ = (int)lx + rcx(i), qy_ = (int)ly + rcy(i), qz_ = (int)lz + rcz(i);
            if (qx_ >= 0 && qy_ >= 0 && qz_ >= 0 && qx_ < (int)BX && qy_ < (int)BY && qz_ < (int)BZ) {
                const uint32_t k2 = ((uint32_t)qz_ * BY + (uint32_t)qy_) * BX + (uint32_t)qx_;
                const uint2 o = *(const uint2 *)(hits + ((size_t)k2 * 9 + (i - 2) / 2) * 2);
                if (o.x != 0x7f7fffffu) {
                    qv[i] = 1.0f - __uint_as_float(o.y);
                    qx[i] = 1.0f - __uint_as_float(o.x);
                    hit[i] = true;
                }
            }
        }
    }
    uint32_t mask = 0;
    for (int i = 1; i < 19; ++i) {
        // a wall where the neighbour is solid, or where the link crosses a
        // sheet thinner than a cell (two hits); q from the nearest hit
        if (solid_at(L, a, px + rcx(i), py + rcy(i), pz + rcz(i), false)) {
            if (!hit[i]) qv[i] = 1.0f;
            mask |= 1u << i;
        } else if (qx[i] - qv[i] > 1e-5f) {
            mask |= 1u << i;
        } else {
            qv[i] = 1.0f;
        }
    }
    for (int w = 0; w < 9; ++w)
        L.rq[(size_t)w * NR + k] = half_rne_bits(fminf(qv[2 * w + 1], 1.0f)) | (half_rne_bits(fminf(qv[2 * w + 2], 1.0f)) << 16);
    L.rmask[k] = mask;
    T F[3] = {0, 0, 0};
    if (solid_before) {
        // a cell the part uncovered: refill what it would have sent out last
        // step with the equilibrium at the wall velocity and the density of
        // its fluid neighbours, plus their mean non-equilibrium stress
        T rs = 0, cnt = 0;
        T pis[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 1; i < 19; ++i) {
            const bool fluid_now = !solid_at(L, a, px + rcx(i), py + rcy(i), pz + rcz(i), false);
            const bool fluid_before = !solid_at(L, a, px + rcx(i), py + rcy(i), pz + rcz(i), true);
            if (!(fluid_now && fluid_before)) continue;
            const int qx_ = (int)lx + rcx(i), qy_ = (int)ly + rcy(i), qz_ = (int)lz + rcz(i);
            if (qx_ < 0 || qy_ < 0 || qz_ < 0 || qx_ >= (int)BX || qy_ >= (int)BY || qz_ >= (int)BZ) continue;
            const uint32_t k2 = ((uint32_t)qz_ * BY + (uint32_t)qy_) * BX + (uint32_t)qx_;
            const T r2 = rmac[k2];
            rs = rs + r2;
            cnt = cnt + (T)1;
            T fp[19];
            fp[0] = (T)0;
            for (int j = 1; j < 19; ++j) fp[j] = rfpost[(size_t)(j - 1) * NR + k2];
            const Mom<T> mo = moments(fp);
            const T ux = rmac[NR + k2], uy = rmac[2 * NR + k2], uz = rmac[3 * NR + k2];
            pis[3] = pis[3] + (mo.sxy - r2 * ux * uy);
            pis[4] = pis[4] + (mo.sxz - r2 * ux * uz);
            pis[5] = pis[5] + (mo.syz - r2 * uy * uz);
            const T tr = (mo.sxx + mo.syy + mo.szz) / (T)3;
            const T uu = ux * ux + uy * uy + uz * uz;
            pis[0] = pis[0] + (mo.sxx - tr - r2 * (ux * ux - uu / (T)3));
            pis[1] = pis[1] + (mo.syy - tr - r2 * (uy * uy - uu / (T)3));
            pis[2] = pis[2] + (mo.szz - tr - r2 * (uz * uz - uu / (T)3));
        }
        const T rho = cnt > (T)0 ? rs / cnt : (T)1;
        const V3f uwf = cross3(v3(a.omega[0], a.omega[1], a.omega[2]), pc);
        const T uwx = (T)uwf.x, uwy = (T)uwf.y, uwz = (T)uwf.z;
        T fe[19];
        feq_arr<T>(rho - (T)1, uwx, uwy, uwz, fe);
        if (cnt > (T)0) {
#pragma unroll
            for (int j = 0; j < 6; ++j) pis[j] = pis[j] / cnt;
            T fq[19];
            fneq_arr<T>(pis, fq);
#pragma unroll
            for (int j = 0; j < 19; ++j) fe[j] = fe[j] + fq[j];
        }
        Nbr nb;
        nb.make(L, (uint32_t)px, (uint32_t)py, (uint32_t)pz, n);
        refill_store<C, 1 - P>(L, nb, fe, a.seed);
        F[0] = -rho * uwx;
        F[1] = -rho * uwy;
        F[2] = -rho * uwz;
        for (int i = 1; i < 19; ++i) rfpost[(size_t)(i - 1) * NR + k] = fe[i];
        rmac[k] = rho;
        rmac[NR + k] = uwx;
        rmac[2 * NR + k] = uwy;
        rmac[3 * NR + k] = uwz;
    }
    rft[k] = F[0]; rft[NR + k] = F[1]; rft[2 * NR + k] = F[2];
    rft[3 * NR + k] = pcy * F[2] - pcz * F[1];
    rft[4 * NR + k] = pcz * F[0] - pcx * F[2];
    rft[5 * NR + k] = pcx * F[1] - pcy * F[0];
    L.flags[n] = (k + 1u) << 2;
}

// ---------------------------------------------------------------- partial bounce-back and its adjoint
// Partial bounce-back (new; the reference has no porous cells): a cell with
// solid fraction phi blends its post-collision populations with the bounced
// inputs, f*_i = (1 - phi) f^c_i + phi f^in_opp(i) (w_i = w_opp(i), so the
// same holds for the deviations f~), and the porous solid takes the force
// F = phi sum_i c_i (f^c_i - f^in_opp(i)) (minus the fluid's momentum change).
// The plain regularized collision (Cs = 0, no body force, no sponge) is the
// forward map the discrete adjoint differentiates.

// forward intermediates of the regularized collision, for its vector-Jacobian product
template <class T> struct RegState {
    T rho1, rho, ux, uy, uz, k1;
};

template <class T> __device__ __forceinline__ T step_k1(const DevLattice &L) {
    // 1 - 1/tau with tau as collide() computes it for Cs = 0
    const T tau0 = (T)L.tau0;
    const T tau = (T)0.5 * (tau0 + sqrt_(tau0 * tau0));
    return (T)1 - rcp_(tau);
}

// adjoint of the moments -> (rho1, j, S) -> f: fbar_k = r1 + c_k.jb + c_k c_k : Sb
template <class T>
__device__ __forceinline__ void moments_vjp(T r1, T jbx, T jby, T jbz, const T sb[6], T fb[19]) {
    sfor<19>([&](auto I) {
        constexpr int i = I;
        T v = r1 + cdot<i>(jbx, jby, jbz);
        if constexpr (CX[i] != 0) v = v + sb[0];
        if constexpr (CY[i] != 0) v = v + sb[1];
        if constexpr (CZ[i] != 0) v = v + sb[2];
        if constexpr (CX[i] * CY[i] != 0) v = v + (T)(CX[i] * CY[i]) * sb[3];
        if constexpr (CX[i] * CZ[i] != 0) v = v + (T)(CX[i] * CZ[i]) * sb[4];
        if constexpr (CY[i] * CZ[i] != 0) v = v + (T)(CY[i] * CZ[i]) * sb[5];
        fb[i] = v;
    });
}

// Given the adjoint of (rho1 through rho, u, Pi) of a state computed from f
// (rho = 1 + rho1, u = j / rho, Pi = S - rho u u - rho1/3 I), add into fb.
template <class T>
__device__ __forceinline__ void macro_vjp(T rho, T ux, T uy, T uz, T rbar, T ubx, T uby, T ubz, const T pb[6], T fb[19]) {
    // Pi = S - rho u u - rho1/3 I
    T sb[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) sb[k] = pb[k];
    rbar = rbar - (ux * ux * pb[0] + uy * uy * pb[1] + uz * uz * pb[2] + ux * uy * pb[3] + ux * uz * pb[4] + uy * uz * pb[5]);
    rbar = rbar - (pb[0] + pb[1] + pb[2]) * ((T)1 / (T)3);
    ubx = ubx - rho * ((T)2 * ux * pb[0] + uy * pb[3] + uz * pb[4]);
    uby = uby - rho * ((T)2 * uy * pb[1] + ux * pb[3] + uz * pb[5]);
    ubz = ubz - rho * ((T)2 * uz * pb[2] + ux * pb[4] + uy * pb[5]);
    // u = j / rho
    const T ir = rcp_(rho);
    rbar = rbar - (ubx * ux + uby * uy + ubz * uz) * ir;
    moments_vjp<T>(rbar, ubx * ir, uby * ir, ubz * ir, sb, fb);
}

// vector-Jacobian product of the plain regularized collision at the state
// st: fb = (d out / d f)^T ab
template <class T> __device__ __forceinline__ void collide_vjp(const RegState<T> &st, const T ab[19], T fb[19]) {
    const T usq = st.ux * st.ux + st.uy * st.uy + st.uz * st.uz;
    T rbar = 0, r1bar = 0, ubx = 0, uby = 0, ubz = 0;
    T pb[6] = {0, 0, 0, 0, 0, 0};
    sfor<19>([&](auto I) {
        constexpr int i = I;
        const T w = wgt<T>(i);
        const T a = ab[i] * w;
        const T cu = cdot<i>(st.ux, st.uy, st.uz);
        r1bar = r1bar + a;
        rbar = rbar + a * ((T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * usq);
        const T g = st.rho * ((T)3 + (T)9 * cu);
        if constexpr (CX[i] != 0) ubx = ubx + a * (T)CX[i] * g;
        if constexpr (CY[i] != 0) uby = uby + a * (T)CY[i] * g;
        if constexpr (CZ[i] != 0) ubz = ubz + a * (T)CZ[i] * g;
        ubx = ubx - a * (T)3 * st.rho * st.ux;
        uby = uby - a * (T)3 * st.rho * st.uy;
        ubz = ubz - a * (T)3 * st.rho * st.uz;
        // k1 * 4.5 w (sum_ab c_a c_b P_ab - tr P / 3)
        const T q = st.k1 * (T)4.5 * a;
        pb[0] = pb[0] + q * ((T)(CX[i] * CX[i]) - (T)1 / (T)3);
        pb[1] = pb[1] + q * ((T)(CY[i] * CY[i]) - (T)1 / (T)3);
        pb[2] = pb[2] + q * ((T)(CZ[i] * CZ[i]) - (T)1 / (T)3);
        if constexpr (CX[i] * CY[i] != 0) pb[3] = pb[3] + q * (T)(2 * CX[i] * CY[i]);
        if constexpr (CX[i] * CZ[i] != 0) pb[4] = pb[4] + q * (T)(2 * CX[i] * CZ[i]);
        if constexpr (CY[i] * CZ[i] != 0) pb[5] = pb[5] + q * (T)(2 * CY[i] * CZ[i]);
    });
    macro_vjp<T>(st.rho, st.ux, st.uy, st.uz, rbar + r1bar, ubx, uby, ubz, pb, fb);
}

template <class T> __device__ __forceinline__ RegState<T> reg_state(const DevLattice &L, const T f[19]) {
    const Mom<T> m = moments(f);
    RegState<T> st;
    st.rho1 = m.rho1;
    st.rho = (T)1 + m.rho1;
    const T ir = rcp_(st.rho);
    st.ux = m.jx * ir;
    st.uy = m.jy * ir;
    st.uz = m.jz * ir;
    st.k1 = step_k1<T>(L);
    return st;
}

// one forward step of a lattice with partial bounce-back: the plain
// collision, the blend, the porous force (per-block FP64 partials of the
// step's slot), the fused face update. P: parity.
template <class C, int P>
__global__ void __launch_bounds__(1024) k_pbb_step(DevLattice L, StepArgs a, const float *phi) {
    using T = typename C::T;
    __shared__ double sh[16][6];
    const uint32_t xb = blockIdx.x % L.xblocks, yb = blockIdx.x / L.xblocks;
    const uint32_t x = xb * L.bx + threadIdx.x;
    const uint32_t y = yb * L.by + threadIdx.y;
    const uint32_t z = blockIdx.y;
    const bool inside = x < L.nx && y < L.ny;
    const uint32_t n = (z * L.ny + y) * L.nx + x;
    const uint32_t fg = inside ? L.flags[n] : 1u;
    const bool mine = (fg & 3u) == 0;
    T F[3] = {0, 0, 0};
    T f[19];
    FaceSrc<T> s;
    bool src = false;
    if (mine) {
        Nbr nb;
        nb.make(L, x, y, z, n);
        gather_pre<C, P>(L, nb, f);
        T fin[19];
#pragma unroll
        for (int i = 0; i < 19; ++i) fin[i] = f[i];
        T rho, u[3], pp[6];
        collide<T, false, false>(L, x, y, z, f, rho, u, NoHook{}, pp);
        const T ph = (T)phi[n];
        if (ph != (T)0) {
            sfor<19>([&](auto I) {
                constexpr int i = I;
                const T d = f[i] - fin[OPP[i]];
                if constexpr (CX[i] != 0) F[0] = F[0] + (T)CX[i] * ph * d;
                if constexpr (CY[i] != 0) F[1] = F[1] + (T)CY[i] * ph * d;
                if constexpr (CZ[i] != 0) F[2] = F[2] + (T)CZ[i] * ph * d;
            });
#pragma unroll
            for (int i = 0; i < 19; ++i) f[i] = ((T)1 - ph) * f[i] + ph * fin[OPP[i]];
        }
        src = is_face_source(L, x, y, z);
        if (src) s = face_source(L, f);
        scatter_post<C, P, false>(L, nb, f, a.seed);
    }
    if (__ballot(src)) face_writes_wave<C, P>(L, src, x, y, z, s, a.seed + 2654435761u);
    double v[6] = {(double)F[0], (double)F[1], (double)F[2], 0.0, 0.0, 0.0};
    const uint32_t tid = threadIdx.y * blockDim.x + threadIdx.x;
    block_sum6(v, sh, tid, blockDim.x * blockDim.y);
    if (tid == 0) {
        const uint32_t nblk = gridDim.x * gridDim.y, bid = blockIdx.y * gridDim.x + blockIdx.x;
        double *o = &L.partials[((size_t)a.slot * nblk + bid) * 8];
#pragma unroll
        for (int j = 0; j < 6; ++j) o[j] = v[j];
    }
}

// Adjoint of the face update that followed a step of parity P (the face
// cells' inputs of the next step, written from their source cells' outputs).
// One thread per source cell: it recomputes its own step-P output f* from
// the step's input state `fwd`, reads the adjoints of the face slots it
// wrote, adds their vector-Jacobian product to the adjoints of its own
// output slots, and clears the face slots' adjoints.
template <class C, int P>
__global__ void __launch_bounds__(256) k_adj_faces(DevLattice L, DevLattice A, const float *phi) {
    using T = typename C::T;
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L.ncell) return;
    uint32_t x, y, z;
    coords(L, m, x, y, z);
    if ((L.flags[m] & 3u) != 0 || !is_face_source(L, x, y, z)) return;
    Nbr nb;
    nb.make(L, x, y, z, m);
    T f[19], fin[19];
    gather_pre<C, P>(L, nb, f);
#pragma unroll
    for (int i = 0; i < 19; ++i) fin[i] = f[i];
    T rho, u[3], pp[6];
    collide<T, false, false>(L, x, y, z, f, rho, u, NoHook{}, pp);
    const T ph = (T)phi[m];
    if (ph != (T)0)
#pragma unroll
        for (int i = 0; i < 19; ++i) f[i] = ((T)1 - ph) * f[i] + ph * fin[OPP[i]];
    // the source state the face rule saw: rho, u, Pi of f*
    const FaceSrc<T> s = face_source(L, f);
    T rbar = 0, ubx = 0, uby = 0, ubz = 0;
    T pb[6] = {0, 0, 0, 0, 0, 0};
    T *adj = (T *)A.pops;
    const bool lo[3] = {!L.per[0] && x == 1, !L.per[1] && y == 1, !L.per[2] && z == 1};
    const bool hi[3] = {!L.per[0] && x == L.nx - 2, !L.per[1] && y == L.ny - 2, !L.per[2] && z == L.nz - 2};
    for (int dz = -1; dz <= 1; ++dz) {
        if ((dz < 0 && !lo[2]) || (dz > 0 && !hi[2])) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            if ((dy < 0 && !lo[1]) || (dy > 0 && !hi[1])) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                if ((dx < 0 && !lo[0]) || (dx > 0 && !hi[0])) continue;
                if (dx == 0 && dy == 0 && dz == 0) continue;
                const int bx = (int)x + dx, by = (int)y + dy, bz = (int)z + dz;
                const uint32_t ty = L.flags[((uint32_t)bz * L.ny + (uint32_t)by) * L.nx + (uint32_t)bx] & 3u;
                // the face cell's state: velocity face takes rho from inside
                // and u = u_inf; open face takes rho = 1 and u from inside
                const T rb = ty == 2u ? s.rho : (T)1;
                const T vx = ty == 2u ? (T)L.u_inf[0] : s.ux, vy = ty == 2u ? (T)L.u_inf[1] : s.uy,
                        vz = ty == 2u ? (T)L.u_inf[2] : s.uz;
                const T usq = vx * vx + vy * vy + vz * vz;
                T rbb = 0, vbx = 0, vby = 0, vbz = 0;
                for (int i = 0; i < 19; ++i) {
                    if (!face_live(L, bx, by, bz, i)) continue;
                    // the slot face_store<C, P> wrote direction i of b to
                    const int cx = rcx(i), cy = rcy(i), cz = rcz(i);
                    uint32_t cell, dir;
                    if (P == 0 || i == 0) {
                        cell = ((uint32_t)bz * L.ny + (uint32_t)by) * L.nx + (uint32_t)bx;
                        dir = (uint32_t)ropp(i);
                    } else {
                        int tx = bx + cx, tyy = by + cy, tz = bz + cz;
                        tx = tx < 0 ? tx + (int)L.nx : (tx >= (int)L.nx ? tx - (int)L.nx : tx);
                        tyy = tyy < 0 ? tyy + (int)L.ny : (tyy >= (int)L.ny ? tyy - (int)L.ny : tyy);
                        tz = tz < 0 ? tz + (int)L.nz : (tz >= (int)L.nz ? tz - (int)L.nz : tz);
                        cell = ((uint32_t)tz * L.ny + (uint32_t)tyy) * L.nx + (uint32_t)tx;
                        dir = (uint32_t)i;
                    }
                    T *slot = adj + pop_index(dir, cell);
                    const T ab = *slot;
                    *slot = (T)0;
                    if (ab == (T)0) continue;
                    const T w = rwgt<T>(i);
                    const T cu = (T)cx * vx + (T)cy * vy + (T)cz * vz;
                    // f_b,i = w (rb - 1 + rb (3 cu + 4.5 cu^2 - 1.5 u^2)) + 4.5 w (c c : Pi - tr Pi / 3)
                    rbb = rbb + ab * w * ((T)1 + (T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * usq);
                    const T g = rb * w * ab * ((T)3 + (T)9 * cu);
                    vbx = vbx + g * (T)cx - (T)3 * rb * w * ab * vx;
                    vby = vby + g * (T)cy - (T)3 * rb * w * ab * vy;
                    vbz = vbz + g * (T)cz - (T)3 * rb * w * ab * vz;
                    const T q = (T)4.5 * w * ab;
                    pb[0] = pb[0] + q * ((T)(cx * cx) - (T)1 / (T)3);
                    pb[1] = pb[1] + q * ((T)(cy * cy) - (T)1 / (T)3);
                    pb[2] = pb[2] + q * ((T)(cz * cz) - (T)1 / (T)3);
                    pb[3] = pb[3] + q * (T)(2 * cx * cy);
                    pb[4] = pb[4] + q * (T)(2 * cx * cz);
                    pb[5] = pb[5] + q * (T)(2 * cy * cz);
                }
                if (ty == 2u) {
                    rbar = rbar + rbb;
                } else {
                    ubx = ubx + vbx;
                    uby = uby + vby;
                    ubz = ubz + vbz;
                }
            }
        }
    }
    // back through rho, u, Pi of f* to f* itself (u = j / rho, no force)
    T fb[19];
    macro_vjp<T>(s.rho, s.ux, s.uy, s.uz, rbar, ubx, uby, ubz, pb, fb);
    // add to the adjoints of this cell's step-P output slots
    sfor<19>([&](auto I) {
        constexpr int i = I;
        T *slot;
        if constexpr (P == 0 || i == 0) slot = adj + pop_index(OPP[i], nb.n);
        else slot = adj + pop_index(i, nb.template at<CX[i], CY[i], CZ[i]>());
        *slot = *slot + fb[i];
    });
}

// Adjoint of one step of parity P: each fluid cell reads the adjoints of
// the slots it wrote, goes back through the blend and the collision at its
// recomputed forward state (inputs from `fwd` = L.pops), adds the drag
// functional's derivative (weight wj on this step's x-force), accumulates
// d/d phi and writes the adjoints of its inputs into the slots it read (the
// same set of slots: in-place like the forward).
template <class C, int P>
__global__ void __launch_bounds__(1024) k_adj_step(DevLattice L, DevLattice A, const float *phi, double *grad, double wj, int axis) {
    using T = typename C::T;
    const uint32_t xb = blockIdx.x % L.xblocks, yb = blockIdx.x / L.xblocks;
    const uint32_t x = xb * L.bx + threadIdx.x;
    const uint32_t y = yb * L.by + threadIdx.y;
    const uint32_t z = blockIdx.y;
    if (x >= L.nx || y >= L.ny) return;
    const uint32_t n = (z * L.ny + y) * L.nx + x;
    if ((L.flags[n] & 3u) != 0) return;
    Nbr nb;
    nb.make(L, x, y, z, n);
    T fin[19], fc[19];
    gather_pre<C, P>(L, nb, fin);
#pragma unroll
    for (int i = 0; i < 19; ++i) fc[i] = fin[i];
    T rho, u[3], pp[6];
    collide<T, false, false>(L, x, y, z, fc, rho, u, NoHook{}, pp);
    const RegState<T> st = reg_state<T>(L, fin);
    const T ph = (T)phi[n];
    const T w = (T)wj;
    // adjoints of this cell's outputs (the slots scatter_post<P> wrote)
    T ao[19];
    gather_post<C, P>(A, nb, ao);
    T ac[19], ain[19];
    T gphi = 0;
#pragma unroll
    for (int i = 0; i < 19; ++i) ain[i] = 0;
    sfor<19>([&](auto I) {
        constexpr int i = I;
        const T cax = axis == 0 ? (T)CX[i] : (axis == 1 ? (T)CY[i] : (T)CZ[i]);
        // blend: f*_i = (1 - phi) fc_i + phi fin_opp(i); force term phi c (fc_i - fin_opp(i))
        const T d = fc[i] - fin[OPP[i]];
        gphi = gphi + ao[i] * (fin[OPP[i]] - fc[i]) + w * cax * d;
        ac[i] = ((T)1 - ph) * ao[i] + w * ph * cax;
        ain[OPP[i]] = ain[OPP[i]] + ph * ao[i] - w * ph * cax;
    });
    T fb[19];
    collide_vjp<T>(st, ac, fb);
    grad[n] = grad[n] + (double)gphi;
    // write the adjoints of the inputs into the slots gather_pre<P> read
    T *adj = (T *)A.pops;
    sfor<19>([&](auto I) {
        constexpr int i = I;
        const T v = fb[i] + ain[i];
        if constexpr (P == 0 || i == 0) adj[pop_index(i, nb.n)] = v;
        else adj[pop_index(OPP[i], nb.template at<-CX[i], -CY[i], -CZ[i]>())] = v;
    });
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
    const uint32_t rec = rec_of(fg);
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
#if SIMLAB_KSTORE < 0
__global__ void __launch_bounds__(256) k_sum_final(const double *partials, uint32_t n, double *out) {
    __shared__ double sh[4];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = sh[0] + sh[1] + sh[2] + sh[3];
}
#endif

// ---------------------------------------------------------------- probes
// shader.rs:736-784 (run with the module of the other parity after a step)
// Eight lanes per vertex, one per corner of the interpolation cell (the
// eight of a vertex sit in one wave), summed with lane shuffles.
template <class C, int P> __global__ void __launch_bounds__(256) k_sample_verts(DevLattice L, StepArgs a) {
    using T = typename C::T;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t v = gid >> 3;
    const int k = (int)(gid & 7u);
    float acc = 0.0f, wsum = 0.0f;
    if (v < L.nvert) {
        const float4 pb = L.verts[2 * v], nn = L.verts[2 * v + 1];
        const float bx = pb.x + L.vertex_offset * nn.x, by = pb.y + L.vertex_offset * nn.y, bz = pb.z + L.vertex_offset * nn.z;
        // CENTER + from_body(b)
        const float sx = (float)L.center[0] + (a.r[0][0] * bx + a.r[1][0] * by + a.r[2][0] * bz);
        const float sy = (float)L.center[1] + (a.r[0][1] * bx + a.r[1][1] * by + a.r[2][1] * bz);
        const float sz = (float)L.center[2] + (a.r[0][2] * bx + a.r[1][2] * by + a.r[2][2] * bz);
        const float b0x = floorf(sx), b0y = floorf(sy), b0z = floorf(sz);
        const float frx = sx - b0x, fry = sy - b0y, frz = sz - b0z;
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const int cx = (int)b0x + ox, cy = (int)b0y + oy, cz = (int)b0z + oz;
        if (cx >= 1 && cy >= 1 && cz >= 1 && cx < (int)L.nx - 1 && cy < (int)L.ny - 1 && cz < (int)L.nz - 1) {
            const float w = (ox ? frx : 1.0f - frx) * (oy ? fry : 1.0f - fry) * (oz ? frz : 1.0f - frz);
            T rho, u[3];
            const bool ok = cell_macro<C, P>(L, ((uint32_t)cz * L.ny + (uint32_t)cy) * L.nx + (uint32_t)cx, rho, u);
            if (ok && w > 0.0f) {
                acc = w * (float)rho;
                wsum = w;
            }
        }
    }
#pragma unroll
    for (int m = 4; m >= 1; m >>= 1) {
        acc += __shfl_xor(acc, m, 8);
        wsum += __shfl_xor(wsum, m, 8);
    }
    if (v < L.nvert && k == 0 && wsum > 1e-6f) {
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

// a kernel launch, with its dispatch recording its own start and end in
// ev0 / ev1 when they are given (hipExtLaunchKernel: the timestamps of the
// dispatch itself, no packets between the kernels)
template <class K, class... Args>
hipError_t launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, Args... args) {
    if (ev0) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev0, ev1, 0, args...);
    else kernel<<<grid, block, 0, s>>>(args...);
    return hipGetLastError();
}


template <class C, int P, bool FORCE, bool SPONGE>
hipError_t step_launch4(const DevLattice &L, const StepArgs &a, hipStream_t s_special, hipStream_t s_bulk, hipEvent_t ev0,
                        hipEvent_t ev1) {
    k_special<C, P, FORCE, SPONGE><<<L.wall_blocks + L.hw_blocks, SPECIAL_BLOCK, 0, s_special>>>(L, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if constexpr (rows_variant<C, P>())
        if (L.rows_aligned)
            return launch_timed(k_bulk<C, P, FORCE, SPONGE, true>, dim3(L.xblocks * L.yblocks, L.nz), dim3(L.bx, L.by),
                                s_bulk, ev0, ev1, L, a, ZMap{0u, L.nz, 0u});
    return launch_timed(k_bulk<C, P, FORCE, SPONGE>, dim3(L.xblocks * L.yblocks, L.nz), dim3(L.bx, L.by), s_bulk, ev0, ev1,
                        L, a, ZMap{0u, L.nz, 0u});
}

template <class C, int P, bool FORCE, bool SPONGE>
hipError_t special_launch4(const DevLattice &L, const StepArgs &a, hipStream_t s) {
    k_special<C, P, FORCE, SPONGE><<<L.wall_blocks + L.hw_blocks, SPECIAL_BLOCK, 0, s>>>(L, a);
    return hipGetLastError();
}

template <class C, int P, bool FORCE, bool SPONGE>
hipError_t bulk_z_launch4(const DevLattice &L, const StepArgs &a, hipStream_t s, uint32_t nplanes, const ZMap &zm,
                          hipEvent_t ev0, hipEvent_t ev1) {
    if (nplanes == 0) return hipSuccess;
    if constexpr (rows_variant<C, P>())
        if (L.rows_aligned)
            return launch_timed(k_bulk<C, P, FORCE, SPONGE, true>, dim3(L.xblocks * L.yblocks, nplanes), dim3(L.bx, L.by), s,
                                ev0, ev1, L, a, zm);
    return launch_timed(k_bulk<C, P, FORCE, SPONGE>, dim3(L.xblocks * L.yblocks, nplanes), dim3(L.bx, L.by), s, ev0, ev1,
                        L, a, zm);
}

// the (FORCE, SPONGE) variant of a launch template for this lattice
#define SIMLAB_VARIANT(FN, C, P, ...)                                              \
    do {                                                                           \
        const bool force_ = L.forcing != 0, sponge_ = L.sponge != 0;               \
        if (!force_ && !sponge_) return FN<C, P, false, false>(__VA_ARGS__);       \
        if (force_ && !sponge_) return FN<C, P, true, false>(__VA_ARGS__);        \
        if (!force_ && sponge_) return FN<C, P, false, true>(__VA_ARGS__);        \
        return FN<C, P, true, true>(__VA_ARGS__);                                  \
    } while (0)

template <class C, int P, bool FORCE, bool SPONGE>
hipError_t fused_launch4(const DevLattice &L, const StepArgs &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    const uint64_t tiles = (uint64_t)L.xblocks * L.yblocks * L.nz;
    const uint32_t per = SPECIAL_BLOCK / (L.bx * L.by);
    const uint32_t nb = L.wall_blocks + L.hw_blocks + cdiv(tiles, per);
    if constexpr (rows_variant<C, P>())
        if (L.rows_aligned) return launch_timed(k_step<C, P, FORCE, SPONGE, true>, dim3(nb), dim3(SPECIAL_BLOCK), s, ev0, ev1, L, a);
    return launch_timed(k_step<C, P, FORCE, SPONGE>, dim3(nb), dim3(SPECIAL_BLOCK), s, ev0, ev1, L, a);
}

template <class C, int P>
hipError_t fused_launch(const DevLattice &L, const StepArgs &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    const bool force = L.forcing != 0, sponge = L.sponge != 0;
    if (!force && !sponge) return fused_launch4<C, P, false, false>(L, a, s, ev0, ev1);
    if (force && !sponge) return fused_launch4<C, P, true, false>(L, a, s, ev0, ev1);
    if (!force && sponge) return fused_launch4<C, P, false, true>(L, a, s, ev0, ev1);
    return fused_launch4<C, P, true, true>(L, a, s, ev0, ev1);
}

template <class C>
hipError_t fused_launch_p(const DevLattice &L, int p, const StepArgs &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    return p == 0 ? fused_launch<C, 0>(L, a, s, ev0, ev1) : fused_launch<C, 1>(L, a, s, ev0, ev1);
}

template <class C, int P>
hipError_t step_launch(const DevLattice &L, const StepArgs &a, hipStream_t s1, hipStream_t s2, hipEvent_t ev0,
                       hipEvent_t ev1) {
    const bool force = L.forcing != 0, sponge = L.sponge != 0;
    if (!force && !sponge) return step_launch4<C, P, false, false>(L, a, s1, s2, ev0, ev1);
    if (force && !sponge) return step_launch4<C, P, true, false>(L, a, s1, s2, ev0, ev1);
    if (!force && sponge) return step_launch4<C, P, false, true>(L, a, s1, s2, ev0, ev1);
    return step_launch4<C, P, true, true>(L, a, s1, s2, ev0, ev1);
}

template <class C>
hipError_t step_launch_p(const DevLattice &L, int p, const StepArgs &a, hipStream_t s1, hipStream_t s2, hipEvent_t ev0,
                         hipEvent_t ev1) {
    return p == 0 ? step_launch<C, 0>(L, a, s1, s2, ev0, ev1) : step_launch<C, 1>(L, a, s1, s2, ev0, ev1);
}

}  // namespace

// ---------------------------------------------------------------- launchers
// Every storage's kernels are compiled in a translation unit of their own
// (SIMLAB_KSTORE = its simlab_storage value), which fills that storage's
// table of launchers; the storage-independent kernels and the dispatch on
// L.storage are compiled with SIMLAB_KSTORE = -1.
#if SIMLAB_KSTORE >= 0

namespace {
#if SIMLAB_KSTORE == 0
using KC = CF32;
#elif SIMLAB_KSTORE == 1
using KC = CF16;
#elif SIMLAB_KSTORE == 2
using KC = CBF16;
#else
using KC = CF64;
#endif
constexpr bool KC_WIDE = std::is_same_v<KC, CF32> || std::is_same_v<KC, CF64>;

hipError_t t_init(const DevLattice &L, hipStream_t s) {
    k_init<KC><<<cdiv(L.ncell, 256), 256, 0, s>>>(L);
    return hipGetLastError();
}

hipError_t t_faces(const DevLattice &L, int p, uint32_t seed, hipStream_t s) {
    if (L.nface == 0) return hipSuccess;
    const uint32_t nb = cdiv(L.nface, 256);
    if (p == 0) k_faces<KC, 0><<<nb, 256, 0, s>>>(L, seed);
    else k_faces<KC, 1><<<nb, 256, 0, s>>>(L, seed);
    return hipGetLastError();
}

hipError_t t_xfaces(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    const uint32_t seed = a.seed + 2654435761u;   // the face writes' dither seed
    if (!L.xdefer) return hipSuccess;
    if constexpr (std::is_same_v<KC, CF64>) {
        return hipErrorInvalidValue;
    } else {
        const uint32_t nb = cdiv(2ull * L.nz * L.ny, 256);
        if (p == 0) k_xfaces<KC, 0><<<nb, 256, 0, s>>>(L, seed, a.mark_lo, a.mark_hi);
        else k_xfaces<KC, 1><<<nb, 256, 0, s>>>(L, seed, a.mark_lo, a.mark_hi);
        return hipGetLastError();
    }
}

hipError_t t_stream_collide(const DevLattice &L, int p, const StepArgs &a, hipStream_t s1, hipStream_t s2, hipEvent_t ev0,
                            hipEvent_t ev1) {
    return step_launch_p<KC>(L, p, a, s1, s2, ev0, ev1);
}

hipError_t t_special(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    if (p == 0) SIMLAB_VARIANT(special_launch4, KC, 0, L, a, s);
    SIMLAB_VARIANT(special_launch4, KC, 1, L, a, s);
}

hipError_t t_bulk_z(const DevLattice &L, int p, const StepArgs &a, hipStream_t s, uint32_t zbase, uint32_t nplanes,
                    uint32_t zskip_at, uint32_t zskip_len, hipEvent_t ev0, hipEvent_t ev1) {
    const ZMap zm{zbase, zskip_at, zskip_len};
    if (p == 0) SIMLAB_VARIANT(bulk_z_launch4, KC, 0, L, a, s, nplanes, zm, ev0, ev1);
    SIMLAB_VARIANT(bulk_z_launch4, KC, 1, L, a, s, nplanes, zm, ev0, ev1);
}

hipError_t t_step_fused(const DevLattice &L, int p, const StepArgs &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (L.bx * L.by > SPECIAL_BLOCK || SPECIAL_BLOCK % (L.bx * L.by) != 0) return hipErrorInvalidConfiguration;
    return fused_launch_p<KC>(L, p, a, s, ev0, ev1);
}

hipError_t t_macros(const DevLattice &L, int p, float4 *out, hipStream_t s) {
    const uint32_t nb = cdiv(L.ncell, 256);
    if (p == 0) k_macros<KC, 0><<<nb, 256, 0, s>>>(L, out);
    else k_macros<KC, 1><<<nb, 256, 0, s>>>(L, out);
    return hipGetLastError();
}

hipError_t t_mass_partial(const DevLattice &L, int p, double *partials, uint32_t nblocks, hipStream_t s) {
    if (p == 0) k_mass_partial<KC, 0><<<nblocks, 256, 0, s>>>(L, partials);
    else k_mass_partial<KC, 1><<<nblocks, 256, 0, s>>>(L, partials);
    return hipGetLastError();
}

hipError_t t_samplers(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    // samplers use the module of the other parity (lattice.rs:733-737)
    const int q = 1 - p;
    if (L.nvert) {
        const uint32_t nbv = cdiv((uint64_t)L.nvert * 8, 256);
        if (q == 0) k_sample_verts<KC, 0><<<nbv, 256, 0, s>>>(L, a);
        else k_sample_verts<KC, 1><<<nbv, 256, 0, s>>>(L, a);
    }
    if (L.npix) {
        if (q == 0) k_sample_slices<KC, 0><<<cdiv(L.npix, 256), 256, 0, s>>>(L);
        else k_sample_slices<KC, 1><<<cdiv(L.npix, 256), 256, 0, s>>>(L);
    }
    return hipGetLastError();
}

hipError_t t_pbb_step(const DevLattice &L, int p, const StepArgs &a, const float *phi, hipStream_t s) {
    if constexpr (!KC_WIDE) {
        return hipErrorInvalidValue;
    } else {
        const dim3 grid(L.xblocks * L.yblocks, L.nz), block(L.bx, L.by);
        if (p == 0) k_pbb_step<KC, 0><<<grid, block, 0, s>>>(L, a, phi);
        else k_pbb_step<KC, 1><<<grid, block, 0, s>>>(L, a, phi);
        return hipGetLastError();
    }
}

hipError_t t_adjoint_step(const DevLattice &L, const DevLattice &A, int p, const float *phi, double *grad, double wj,
                          int axis, hipStream_t s) {
    if constexpr (!KC_WIDE) {
        return hipErrorInvalidValue;
    } else {
        const dim3 grid(L.xblocks * L.yblocks, L.nz), block(L.bx, L.by);
        const uint32_t nb = cdiv(L.ncell, 256);
        if (p == 0) {
            k_adj_faces<KC, 0><<<nb, 256, 0, s>>>(L, A, phi);
            k_adj_step<KC, 0><<<grid, block, 0, s>>>(L, A, phi, grad, wj, axis);
        } else {
            k_adj_faces<KC, 1><<<nb, 256, 0, s>>>(L, A, phi);
            k_adj_step<KC, 1><<<grid, block, 0, s>>>(L, A, phi, grad, wj, axis);
        }
        return hipGetLastError();
    }
}

template <int P, bool FORCE, bool SPONGE>
hipError_t batch_launch4(const BatchLaunch &b, const StepArgs &a, hipStream_t s1, hipStream_t s2) {
    // row-aligned addressing when every member has it
    constexpr bool RV = rows_variant<KC, P>();
    const bool rows = RV && b.rows;
    if (b.fused) {
        if (rows) k_step_batch<KC, P, FORCE, SPONGE, RV><<<b.nsp + b.nbulk, SPECIAL_BLOCK, 0, s1>>>(b.lats, b.boff_sp, b.nl, a);
        else k_step_batch<KC, P, FORCE, SPONGE><<<b.nsp + b.nbulk, SPECIAL_BLOCK, 0, s1>>>(b.lats, b.boff_sp, b.nl, a);
        return hipGetLastError();
    }
    k_special_batch<KC, P, FORCE, SPONGE><<<b.nsp, SPECIAL_BLOCK, 0, s1>>>(b.lats, b.boff_sp, b.nl, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (rows) k_bulk_batch<KC, P, FORCE, SPONGE, RV><<<b.nbulk, SPECIAL_BLOCK, 0, s2>>>(b.lats, b.boff_bulk, b.nl, a);
    else k_bulk_batch<KC, P, FORCE, SPONGE><<<b.nbulk, SPECIAL_BLOCK, 0, s2>>>(b.lats, b.boff_bulk, b.nl, a);
    return hipGetLastError();
}
template <int P> hipError_t batch_launch2(const BatchLaunch &b, const StepArgs &a, hipStream_t s1, hipStream_t s2) {
    if (!b.force && !b.sponge) return batch_launch4<P, false, false>(b, a, s1, s2);
    if (b.force && !b.sponge) return batch_launch4<P, true, false>(b, a, s1, s2);
    if (!b.force && b.sponge) return batch_launch4<P, false, true>(b, a, s1, s2);
    return batch_launch4<P, true, true>(b, a, s1, s2);
}
hipError_t t_batch_step(const BatchLaunch &b, int p, const StepArgs &a, hipStream_t s1, hipStream_t s2) {
    return p == 0 ? batch_launch2<0>(b, a, s1, s2) : batch_launch2<1>(b, a, s1, s2);
}

hipError_t t_voxelize(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    const uint32_t nb = cdiv(L.nrec, 256);
    if (p == 0) k_voxelize<KC, 0><<<nb, 256, 0, s>>>(L, a);
    else k_voxelize<KC, 1><<<nb, 256, 0, s>>>(L, a);
    return hipGetLastError();
}
}  // namespace

#if SIMLAB_KSTORE == 0
#define KTABLE launchers_fp32
#elif SIMLAB_KSTORE == 1
#define KTABLE launchers_fp16
#elif SIMLAB_KSTORE == 2
#define KTABLE launchers_bf16
#else
#define KTABLE launchers_fp64
#endif
#ifndef __HIP_DEVICE_COMPILE__   // a host table of host functions
extern const StorageLaunchers KTABLE;
const StorageLaunchers KTABLE = {t_init,       t_faces,          t_xfaces,   t_stream_collide, t_step_fused,
                                 t_macros,     t_mass_partial,   t_samplers, t_pbb_step,       t_adjoint_step,
                                 t_batch_step, t_voxelize,       t_special,  t_bulk_z};
#endif
#undef KTABLE

#else   // SIMLAB_KSTORE < 0: storage-independent kernels and the dispatch

namespace {
const StorageLaunchers &kl(int32_t storage) {
    switch (storage) {
        case SIMLAB_STORAGE_FP32: return launchers_fp32;
        case SIMLAB_STORAGE_FP16: return launchers_fp16;
        case SIMLAB_STORAGE_BF16: return launchers_bf16;
        default: return launchers_fp64;
    }
}
}  // namespace

hipError_t launch_init(const DevLattice &L, hipStream_t s) { return kl(L.storage).init(L, s); }
hipError_t launch_faces(const DevLattice &L, int p, uint32_t seed, hipStream_t s) { return kl(L.storage).faces(L, p, seed, s); }
hipError_t launch_xfaces(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    return kl(L.storage).xfaces(L, p, a, s);
}
hipError_t launch_stream_collide(const DevLattice &L, int p, const StepArgs &a, hipStream_t s_special, hipStream_t s_bulk,
                                 hipEvent_t ev0, hipEvent_t ev1) {
    return kl(L.storage).stream_collide(L, p, a, s_special, s_bulk, ev0, ev1);
}
hipError_t launch_step_fused(const DevLattice &L, int p, const StepArgs &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    return kl(L.storage).step_fused(L, p, a, s, ev0, ev1);
}
hipError_t launch_special(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    return kl(L.storage).special(L, p, a, s);
}
hipError_t launch_bulk_planes(const DevLattice &L, int p, const StepArgs &a, hipStream_t s, uint32_t zbase, uint32_t nplanes,
                              uint32_t zskip_at, uint32_t zskip_len, hipEvent_t ev0, hipEvent_t ev1) {
    return kl(L.storage).bulk_z(L, p, a, s, zbase, nplanes, zskip_at, zskip_len, ev0, ev1);
}
hipError_t launch_macros(const DevLattice &L, int p, float4 *out, hipStream_t s) { return kl(L.storage).macros(L, p, out, s); }
hipError_t launch_mass(const DevLattice &L, int p, double *partials, uint32_t nblocks, double *out, hipStream_t s) {
    hipError_t e = kl(L.storage).mass_partial(L, p, partials, nblocks, s);
    if (e != hipSuccess) return e;
    k_sum_final<<<1, 256, 0, s>>>(partials, nblocks, out);
    return hipGetLastError();
}
hipError_t launch_reduce_hist(const DevLattice &L, uint32_t nsteps, hipStream_t s) {
    k_reduce_hist<<<nsteps, 256, 0, s>>>(L.partials, L.wall_blocks + L.hw_blocks, L.hist);
    return hipGetLastError();
}
hipError_t launch_reduce_partials(const double *partials, uint32_t nparts, uint32_t nsteps, double *hist, hipStream_t s) {
    k_reduce_hist<<<nsteps, 256, 0, s>>>(partials, nparts, hist);
    return hipGetLastError();
}
hipError_t launch_samplers(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    return kl(L.storage).samplers(L, p, a, s);
}
hipError_t launch_pbb_step(const DevLattice &L, int p, const StepArgs &a, const float *phi, hipStream_t s) {
    return kl(L.storage).pbb_step(L, p, a, phi, s);
}
hipError_t launch_adjoint_step(const DevLattice &L, const DevLattice &A, int p, const float *phi, double *grad, double wj,
                               int axis, hipStream_t s) {
    return kl(L.storage).adjoint_step(L, A, p, phi, grad, wj, axis, s);
}
hipError_t launch_batch_step(const BatchLaunch &b, int p, const StepArgs &a, hipStream_t s1, hipStream_t s2) {
    return kl(b.storage).batch_step(b, p, a, s1, s2);
}
hipError_t launch_scatter(const DevLattice &L, const StepArgs &a, hipStream_t s) {
    if (L.ntri) k_tri_scatter<<<cdiv(L.ntri, 256 / TRI_LANES), 256, 0, s>>>(L, a);
    return hipGetLastError();
}
hipError_t launch_qclear_near(const DevLattice &L, const StepArgs &a, hipStream_t s) {
    k_qclear_near<<<cdiv(L.nrec, 256), 256, 0, s>>>(L, a);
    return hipGetLastError();
}
hipError_t launch_voxelize(const DevLattice &L, int p, const StepArgs &a, hipStream_t s) {
    return kl(L.storage).voxelize(L, p, a, s);
}
hipError_t launch_qclear(const DevLattice &L, hipStream_t s) {
    k_qclear<<<cdiv(2ull * L.nrec * 9, 256), 256, 0, s>>>(L);
    return hipGetLastError();
}

#endif  // SIMLAB_KSTORE

}  // namespace simlab
