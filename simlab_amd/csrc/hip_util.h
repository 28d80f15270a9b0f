// HIP error handling and owned device buffers shared by the host sources
// of libsimlab_cfd.so (lattice.cpp, mlattice.cpp).
#pragma once

#include "../../include/simlab_cfd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>

namespace simlab {

// A failed HIP call, with the ABI status code it maps to.
struct HipError : std::runtime_error {
    int code;
    HipError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();   // clear the sticky error of failed launches
        throw HipError(e == hipErrorOutOfMemory ? SIMLAB_ERR_OOM : SIMLAB_ERR_GPU,
                       std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Device buffers owned by a lattice; freed in the destructor.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t n, size_t &total) {
        bytes = std::max<size_t>(n, 16);
        hip_check(hipMalloc(&p, bytes), "hipMalloc");
        total += bytes;
    }
    template <class T> T *as() const { return (T *)p; }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

}  // namespace simlab
