// vx_schedule.h -- host-side check of the schedules the fill and emit
// launches index through (a batch's longest-first fill order, the emission
// order the scan writes, the server's per-flight order).  Plain C++: no HIP
// types, so the check is also compiled and tested on its own on the CPU.
#pragma once

#include <stdint.h>

#include <vector>

namespace vx {

// True iff a[0..n) holds every index of [0, n) exactly once: each launch
// workgroup then indexes a distinct, in-range request.
inline bool is_schedule(const int32_t *a, int32_t n) {
    if (n < 0 || (n > 0 && !a)) return false;
    std::vector<unsigned char> seen((size_t)n, 0);
    for (int32_t i = 0; i < n; i++) {
        const int32_t v = a[i];
        if (v < 0 || v >= n || seen[(size_t)v]) return false;
        seen[(size_t)v] = 1;
    }
    return true;
}

}  // namespace vx
