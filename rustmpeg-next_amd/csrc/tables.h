// tables.h - normative integer tables shared by the host layer and the kernels.
//
// Transform matrices are generated on the host from one cosine table (DCT-II,
// all sizes) and one sine row per size (DST-VII / DCT-VIII), the construction
// H.266's integer transMatrix follows; rmh_transform_matrix exposes the result
// so tests can compare it entry by entry. Interpolation and deblocking tables
// are H.266 Tables 33, 34 and 43 (crates/rustmpeg-video/src/vvc/inter.rs:57-80,
// vvc/predict.rs:86-119, vvc/filter.rs:184-198).
#pragma once
#include <stdint.h>

namespace rmh {

// Angle-indexed cosine magnitudes 64*sqrt(2)*cos(pi*m/128), m = 0..64, as the
// normative integers (m = 0 is the DC row value 64).
static const int16_t kCos64[65] = {
    64, 91, 90, 90, 90, 90, 90, 90, 89, 88, 88, 87, 87, 86, 85, 84, 83, 83, 82, 81, 80, 79,
    78, 77, 75, 73, 73, 71, 70, 69, 67, 65, 64, 62, 61, 59, 57, 56, 54, 52, 50, 48, 46, 44,
    43, 41, 38, 37, 36, 33, 31, 28, 25, 24, 22, 20, 18, 15, 13, 11, 9,  7,  4,  2,  0};

// sin(pi*i/(2N+1)) magnitudes, i = 1..N, for N = 4, 8, 16, 32.
static const int16_t kSin4[4] = {29, 55, 74, 84};
static const int16_t kSin8[8] = {17, 32, 46, 60, 71, 78, 85, 86};
static const int16_t kSin16[16] = {8, 17, 25, 33, 40, 48, 55, 62, 68, 73, 77, 81, 85, 87, 88, 88};
static const int16_t kSin32[32] = {4,  9,  13, 17, 21, 26, 30, 34, 38, 42, 46, 50, 53, 56, 60, 63,
                                   66, 68, 72, 74, 77, 78, 80, 82, 84, 85, 86, 87, 88, 89, 90, 90};

// Entry (k, j) of the N-point matrix of `type` (0 DCT-II, 1 DST-VII, 2 DCT-VIII);
// returns 0 and sets *ok = false for unsupported (type, N).
inline int transform_entry(int type, int n, int k, int j, bool* ok) {
    *ok = true;
    if (type == 0) {
        if (n != 2 && n != 4 && n != 8 && n != 16 && n != 32 && n != 64) {
            *ok = false;
            return 0;
        }
        if (k == 0) return 64;
        int m = ((64 / n) * k * (2 * j + 1)) % 256;
        int s = 1;
        if (m > 128) m = 256 - m;
        if (m > 64) {
            m = 128 - m;
            s = -1;
        }
        return s * kCos64[m];
    }
    const int16_t* t = n == 4 ? kSin4 : n == 8 ? kSin8 : n == 16 ? kSin16 : n == 32 ? kSin32 : nullptr;
    if (!t || (type != 1 && type != 2)) {
        *ok = false;
        return 0;
    }
    int p = 2 * n + 1;
    int r = type == 1 ? (2 * k + 1) * (j + 1) : (p - (2 * k + 1) * (2 * j + 1)) / 2;
    int m = r % (2 * p);
    if (m < 0) m += 2 * p;
    int s = 1;
    if (m >= p) {
        m -= p;
        s = -1;
    }
    if (m > n) m = p - m;
    return m == 0 ? 0 : s * t[m - 1];
}

// LFNST matrix set for an intra-mode index 0..94 (g_lfnstLut, vvc/tr_tables.rs:478).
inline int lfnst_set(int mode) {
    if (mode <= 1) return 0;
    if (mode <= 12) return 1;
    if (mode <= 23) return 2;
    if (mode <= 44) return 3;
    if (mode <= 55) return 2;
    return 1;
}

}  // namespace rmh
