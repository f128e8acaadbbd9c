// Overlap test of two row-strided blocks of doubles, host only and free of
// HIP headers so that host-side tests can include it alone.
#pragma once

#include <cstdint>

namespace crs {

// True when the doubles a[r][i] (r < ra) and b[r][i] (r < rb), i < n, both
// with row stride ld, share a byte.  Entry (r1, i1) of a meets (r2, i2) of
// b when |d - 8 k ld - 8 j| < 8 with d = b - a in bytes, k = r1 - r2 and
// j = i1 - i2 (|j| < n), i.e. when |d - 8 k ld| < 8 n for some k in
// [-(rb - 1), ra - 1]; since 8 n <= 8 ld only the two multiples of 8 ld
// either side of d can qualify.
inline bool rows_overlap(const void *a, int64_t ra, const void *b, int64_t rb,
                         int64_t n, int64_t ld)
{
    if (n <= 0 || ra <= 0 || rb <= 0 || ld < n) return false;
    const __int128 d = (__int128)(uintptr_t)b - (__int128)(uintptr_t)a;
    const __int128 s = (__int128)8 * ld, w = (__int128)8 * n;
    __int128 k = d / s, e = d % s;
    if (e < 0) { e += s; --k; }          // d = k s + e, 0 <= e < s
    const bool lo = e < w && k >= -(rb - 1) && k <= ra - 1;
    const bool hi = s - e < w && k + 1 >= -(rb - 1) && k + 1 <= ra - 1;
    return lo || hi;
}

}  // namespace crs
