// hr_host.h -- host-side helpers shared by the library's HIP translation
// units: status propagation and grow-only device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "hr_error.h"

namespace hr {

inline hr_status fail(hr_status s, const std::string& msg) { return set_error(s, msg); }

#define HR_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::hr::fail(e_ == hipErrorOutOfMemory ? HR_ERR_OOM : HR_ERR_DEVICE,           \
                              std::string(#call) + ": " + hipGetErrorString(e_));               \
    } while (0)

#define HR_TRY(call)                                                                            \
    do {                                                                                        \
        hr_status s_ = (call);                                                                  \
        if (s_ != HR_OK) return s_;                                                             \
    } while (0)

// grow-only device buffer (MemoryManager::allocate*, src/RT/Structure/MemoryManager.h:73-139)
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    hr_status reserve(size_t n)
    {
        if (n <= cap && p) return HR_OK;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(HR_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                        hipGetErrorString(e));
        }
        cap = std::max<size_t>(n, 1);
        return HR_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    size_t bytes() const { return cap * sizeof(T); }
};

}  // namespace hr
