// Host helpers shared by the engine and the ranker: error type, HIP call
// check, grow-only device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace ia {

struct DevError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define IA_CHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            throw ::ia::DevError(std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

// Grow-only device buffer.
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    template <class T>
    T *get(size_t n) {
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (bytes > cap) {
            if (p) IA_CHECK(hipFree(p));
            p = nullptr;
            const size_t want = std::max(bytes, cap + cap / 2);
            IA_CHECK(hipMalloc(&p, want));
            cap = want;
        }
        return (T *)p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace ia
