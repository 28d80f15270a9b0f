// Memory ceilings of the streaming access patterns (measurement only): the
// loads and stores of one D3Q19 step on the grouped 16-bit population layout
// (slot (i, n) = (n / 64) * 19 * 64 + i * 64 + n % 64) with no arithmetic,
// on a periodic nx x ny x nz grid, so that the stream-collide kernels'
// achieved bandwidth can be set against what their access pattern moves at
// all on this GPU.

#include "../../include/simlab_cfd.h"
#include "hip_util.h"
#include "simlab_internal.h"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace simlab {
namespace {

constexpr int Q = 19;
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

__device__ __forceinline__ size_t pslot(int i, uint32_t n) {
    return (size_t)(n / 64) * (Q * 64) + (size_t)i * 64 + (n % 64);
}

struct Grid {
    uint32_t nx, ny, nz, n;
    __device__ __forceinline__ uint32_t nb(uint32_t x, uint32_t y, uint32_t z, int dx, int dy, int dz) const {
        x = (x + nx + (uint32_t)dx) % nx;
        y = (y + ny + (uint32_t)dy) % ny;
        z = (z + nz + (uint32_t)dz) % nz;
        return (z * ny + y) * nx + x;
    }
};

// pattern 0: AA even step, in place (n, i) -> (n, opp i); 1: AA odd step,
// in place (n - c_i, opp i) -> (n + c_i, i); 2: two-lattice pull
// a(n - c_i, i) -> b(n, i); 3: out of place, aligned a(n, i) -> b(n, opp i)
template <int P> __global__ __launch_bounds__(256) void k_pattern(Grid g, const uint16_t *__restrict__ a,
                                                                   uint16_t *__restrict__ b, uint16_t *c) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= g.n) return;
    const uint32_t x = n % g.nx, y = (n / g.nx) % g.ny, z = n / (g.nx * g.ny);
    uint16_t v[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        if constexpr (P == 0) v[i] = c[pslot(i, n)];
        if constexpr (P == 1) v[i] = c[pslot(OPP[i], i == 0 ? n : g.nb(x, y, z, -CX[i], -CY[i], -CZ[i]))];
        if constexpr (P == 2) v[i] = a[pslot(i, i == 0 ? n : g.nb(x, y, z, -CX[i], -CY[i], -CZ[i]))];
        if constexpr (P == 3) v[i] = a[pslot(i, n)];
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const uint16_t w = (uint16_t)(v[i] + 1u);
        if constexpr (P == 0) c[pslot(OPP[i], n)] = w;
        if constexpr (P == 1) c[pslot(i, i == 0 ? n : g.nb(x, y, z, CX[i], CY[i], CZ[i]))] = w;
        if constexpr (P == 2) b[pslot(i, n)] = w;
        if constexpr (P == 3) b[pslot(OPP[i], n)] = w;
    }
}

template <int P> void launch(const Grid &g, uint16_t *a, uint16_t *b, int r, hipStream_t st) {
    const dim3 grid((g.n + 255) / 256);
    // the two-lattice patterns swap source and destination every pass
    uint16_t *src = (P >= 2 && (r & 1)) ? b : a, *dst = (P >= 2 && (r & 1)) ? a : b;
    hipLaunchKernelGGL(k_pattern<P>, grid, dim3(256), 0, st, g, src, dst, a);
}

}  // namespace
}  // namespace simlab

using namespace simlab;

extern "C" int simlab_cfd_probe_stream_patterns(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t reps, double gbs[4]) {
    return guarded([&]() -> int {
        try {
            if (!gbs || nx == 0 || ny == 0 || nz == 0 || reps == 0) return fail(SIMLAB_ERR_INVALID, "bad argument");
            const uint64_t n = (uint64_t)nx * ny * nz;
            if (n % 64 != 0 || n >= (1ull << 31)) return fail(SIMLAB_ERR_INVALID, "cells must be a multiple of 64 below 2^31");
            Grid g{nx, ny, nz, (uint32_t)n};
            DevBuf a, b;
            size_t total = 0;
            a.alloc(n * Q * 2, total);
            b.alloc(n * Q * 2, total);
            hip_check(hipMemset(a.p, 0, a.bytes), "clear");
            hip_check(hipMemset(b.p, 0, b.bytes), "clear");
            hipStream_t st = nullptr;
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            for (int p = 0; p < 4; ++p) {
                auto one = [&](int r) {
                    switch (p) {
                        case 0: launch<0>(g, a.as<uint16_t>(), b.as<uint16_t>(), r, st); break;
                        case 1: launch<1>(g, a.as<uint16_t>(), b.as<uint16_t>(), r, st); break;
                        case 2: launch<2>(g, a.as<uint16_t>(), b.as<uint16_t>(), r, st); break;
                        default: launch<3>(g, a.as<uint16_t>(), b.as<uint16_t>(), r, st); break;
                    }
                };
                for (int r = 0; r < 4; ++r) one(r);
                hip_check(hipGetLastError(), "probe kernel");
                hip_check(hipEventRecord(e0, st), "event");
                for (uint32_t r = 0; r < reps; ++r) one((int)r);
                hip_check(hipEventRecord(e1, st), "event");
                hip_check(hipEventSynchronize(e1), "probe kernel");
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
                gbs[p] = 2.0 * Q * 2.0 * (double)n / ((double)ms / reps * 1e-3) / 1e9;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return SIMLAB_OK;
        } catch (const HipError &e) {
            return fail(e.code, e.what());
        }
    });
}
