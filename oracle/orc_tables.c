/*
 * orc_tables.c - transform matrices, LFNST set table and coefficient scans for
 * the CPU oracle (test infrastructure only; see vvc_oracle.h).
 *
 * The integer DCT-II matrices of every size (vvc/dct.rs:58-208) are samples of
 * one cosine: entry (k, j) of the N-point matrix is the 64-point cosine
 * magnitude at angle index (64/N)*k*(2j+1) (units of pi/128), folded into the
 * first quadrant with its sign. DST-VII and DCT-VIII (vvc/tr_tables.rs:14-164)
 * likewise sample one sine per size: sin(pi*(2k+1)(j+1)/(2N+1)) and
 * cos(pi*(2k+1)(2j+1)/(2(2N+1))) = sin(pi*r/(2N+1)), r = (2N+1-(2k+1)(2j+1))/2.
 * The magnitudes below are the normative integer values (H.266 transMatrix);
 * a CPU test compares every generated matrix with the reference tables.
 */
#include <string.h>

#include "vvc_oracle.h"
#include "lfnst_data.h"

/* Cosine magnitudes at angle index m = 0..64 (m = 0 is the 64 of the DC row). */
static const int COS64[65] = {
    64, 91, 90, 90, 90, 90, 90, 90, 89, 88, 88, 87, 87, 86, 85, 84, 83, 83, 82, 81, 80, 79,
    78, 77, 75, 73, 73, 71, 70, 69, 67, 65, 64, 62, 61, 59, 57, 56, 54, 52, 50, 48, 46, 44,
    43, 41, 38, 37, 36, 33, 31, 28, 25, 24, 22, 20, 18, 15, 13, 11, 9, 7, 4, 2, 0};

/* Sine magnitudes sin(pi*i/(2N+1)), i = 1..N, for N = 4, 8, 16, 32. */
static const int SIN4[4] = {29, 55, 74, 84};
static const int SIN8[8] = {17, 32, 46, 60, 71, 78, 85, 86};
static const int SIN16[16] = {8, 17, 25, 33, 40, 48, 55, 62, 68, 73, 77, 81, 85, 87, 88, 88};
static const int SIN32[32] = {4,  9,  13, 17, 21, 26, 30, 34, 38, 42, 46, 50, 53, 56, 60, 63,
                              66, 68, 72, 74, 77, 78, 80, 82, 84, 85, 86, 87, 88, 89, 90, 90};

static int dct2_entry(int n, int k, int j) {
    if (k == 0) return 64;
    int m = ((64 / n) * k * (2 * j + 1)) % 256;
    int sign = 1;
    if (m > 128) m = 256 - m;
    if (m > 64) {
        m = 128 - m;
        sign = -1;
    }
    return sign * COS64[m];
}

static const int* sin_table(int n) {
    switch (n) {
        case 4: return SIN4;
        case 8: return SIN8;
        case 16: return SIN16;
        case 32: return SIN32;
        default: return 0;
    }
}

/* sin(pi * r / (2n+1)) scaled, for any integer r. */
static int sin_entry(int n, int r) {
    const int* t = sin_table(n);
    int p = 2 * n + 1;
    int m = r % (2 * p);
    if (m < 0) m += 2 * p;
    int sign = 1;
    if (m >= p) {
        m -= p;
        sign = -1;
    }
    if (m > n) m = p - m;
    return m == 0 ? 0 : sign * t[m - 1];
}

int orc_transform_matrix(int type, int n, int32_t* out) {
    if (type == RMH_TR_DCT2) {
        if (n != 2 && n != 4 && n != 8 && n != 16 && n != 32 && n != 64) return -1;
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++) out[k * n + j] = dct2_entry(n, k, j);
        return 0;
    }
    if (type != RMH_TR_DST7 && type != RMH_TR_DCT8) return -1;
    if (!sin_table(n)) return -1;
    int p = 2 * n + 1;
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
            if (type == RMH_TR_DST7)
                out[k * n + j] = sin_entry(n, (2 * k + 1) * (j + 1));
            else
                out[k * n + j] = sin_entry(n, (p - (2 * k + 1) * (2 * j + 1)) / 2);
        }
    return 0;
}

/* g_lfnstLut (vvc/tr_tables.rs:478-482): PLANAR/DC use set 0; the angular
 * modes fall into sets 1, 2, 3, 2, 1 by direction (and wide angles 67..94 into 1). */
int orc_lfnst_set(int mode) {
    if (mode < 0 || mode > 94) return -1;
    if (mode <= 1) return 0;
    if (mode <= 12) return 1;
    if (mode <= 23) return 2;
    if (mode <= 44) return 3;
    if (mode <= 55) return 2;
    return 1;
}

/* TrQuant::getTransposeFlag (vvc/lfnst.rs:97-101). */
int orc_lfnst_transposes(int mode) { return (mode >= 35 && mode < 67) || mode >= 81; }

int orc_lfnst_coeff(int set, int idx, int big, int i, int j) {
    return big ? ORC_LFNST8[set][idx][i][j] : ORC_LFNST4[set][idx][i][j];
}

/* coefficient_group_log2 (vvc/transform.rs:100-178). */
void orc_cg_log2(int w, int h, int* lw, int* lh) {
    int lx = 0, ly = 0;
    while ((1 << lx) < w) lx++;
    while ((1 << ly) < h) ly++;
    if (lx > 7) lx = 7;
    if (ly > 7) ly = 7;
    int a, b;
    if (lx == 0) {
        a = 0;
        b = ly < 4 ? ly : 4;
    } else if (ly == 0) {
        a = lx < 4 ? lx : 4;
        b = 0;
    } else if (lx == 1) {
        a = 1;
        b = ly <= 2 ? 1 : 3;
    } else if (ly == 1) {
        a = lx == 2 ? 1 : 3;
        b = 1;
    } else {
        a = 2;
        b = 2;
    }
    *lw = a;
    *lh = b;
}

/* diagonal_scan (vvc/transform.rs:189-212): (x, y) pairs. */
static int diag_scan(int w, int h, int* xs, int* ys) {
    int line = 0, col = 0;
    for (int i = 0; i < w * h; i++) {
        xs[i] = col;
        ys[i] = line;
        if (col == w - 1 || line == 0) {
            line += col + 1;
            col = 0;
            if (line >= h) {
                col += line - (h - 1);
                line = h - 1;
            }
        } else {
            col++;
            line--;
        }
    }
    return w * h;
}

/* CoeffScan::new(w, h).raster (vvc/transform.rs:37-79); returns the length. */
int orc_coeff_scan(int w, int h, int32_t* raster) {
    int lw, lh;
    orc_cg_log2(w, h, &lw, &lh);
    int cw = 1 << lw, chh = 1 << lh;
    int capw = w < 32 ? w : 32, caph = h < 32 ? h : 32;
    int gw = capw >> lw, gh = caph >> lh;
    int gx[1024], gy[1024], ix[64], iy[64];
    int ng = diag_scan(gw, gh, gx, gy);
    int ni = diag_scan(cw, chh, ix, iy);
    int n = 0;
    for (int g = 0; g < ng; g++)
        for (int i = 0; i < ni; i++) {
            int x = gx[g] * cw + ix[i];
            int y = gy[g] * chh + iy[i];
            raster[n++] = y * w + x;
        }
    return n;
}
