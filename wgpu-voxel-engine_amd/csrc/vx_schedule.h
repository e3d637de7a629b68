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

// True iff a[0..m) holds every emit item exactly once, an item being the
// code c = chunk << 2 | slab with chunk in [0, n) and slab < per[chunk] (4
// for a 64^3 chunk, else 1), and m is the number of items.
inline bool is_item_schedule(const int32_t *a, int32_t m, const int32_t *per, int32_t n) {
    if (n < 0 || m < 0 || (m > 0 && !a) || (n > 0 && !per)) return false;
    int64_t items = 0;
    for (int32_t i = 0; i < n; i++) {
        if (per[i] < 1 || per[i] > 4) return false;
        items += per[i];
    }
    if (items != m) return false;
    std::vector<unsigned char> seen((size_t)n * 4, 0);
    for (int32_t k = 0; k < m; k++) {
        const int32_t c = a[k];
        if (c < 0) return false;
        const int32_t chunk = c >> 2, slab = c & 3;
        if (chunk >= n || slab >= per[chunk] || seen[(size_t)c]) return false;
        seen[(size_t)c] = 1;
    }
    return true;
}

}  // namespace vx
