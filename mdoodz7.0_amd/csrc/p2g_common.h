// Device helpers shared by the marker and P2G kernels.
#pragma once
#include <hip/hip_runtime.h>

// Nearest node along one axis, ties to the lower index, clamped to [0, n-1]:
// the reference expression ceil(distance/dx + 0.5) - 1 of
// ParticleRoutines.c:2300-2310.  Division and ceil are correctly rounded on
// both host and device and no contraction applies, so the index is
// bit-identical to the reference.
__device__ __forceinline__ int near_index(double pos, double origin, double h, int n) {
    const double distance = pos - origin;
    int i = (int)ceil(distance / h + 0.5) - 1;
    i = i < 0 ? 0 : i;
    return i > n - 1 ? n - 1 : i;
}

// The same index without the division on almost every call: t = d*(1/h) +
// 0.5 differs from the reference's RN(RN(d/h) + 0.5) by a few ulps at most,
// so ceil(t) is the reference's result whenever t is farther than a margin
// (1e-12 relative, far above those ulps) from an integer; only inside that
// margin (markers within ~1e-12 cells of a mid-cell tie) is the exact
// division taken.
__device__ __forceinline__ int near_index_fast(double pos, double origin, double h, double inv_h, int n) {
    const double distance = pos - origin;
    const double t = __fma_rn(distance, inv_h, 0.5);
    const double f = ceil(t);
    const double margin = 1e-12 * (fabs(t) + 1.0);
    int i = (f - t > margin && t - (f - 1.0) > margin) ? (int)f - 1 : (int)ceil(distance / h + 0.5) - 1;
    i = i < 0 ? 0 : i;
    return i > n - 1 ? n - 1 : i;
}

// Natural logarithm for the per-marker term of geometric averaging, the
// largest remaining instruction cost of the P2G kernels once the weights are
// cached.  Accurate to ~2 ulp (far below the 1e-12 P2G tolerance) at about a
// third of the library log's instructions:
//   v = 2^k m, m in [sqrt(1/2), sqrt(2));  s = (m - 1)/(m + 1);
//   log m = 2 atanh(s) = 2s + 2s z (1/3 + z/5 + ... + z^9/21),  z = s^2,
// the series cut where z^10/23 < 2^-54 (|s| <= 0.1716);
//   log v = k ln2_hi + (2s + (k ln2_lo + 2s z P(z))),
// ln2_hi = ln 2 with its low 21 mantissa bits cleared (k ln2_hi is exact for
// any exponent) and ln2_lo = ln 2 - ln2_hi.  Non-positive, infinite, NaN and
// subnormal arguments go to the library log.
__device__ __noinline__ double p2g_log_lib(double v) { return log(v); }
__device__ __forceinline__ double p2g_log(double v) {
    if (!(v >= 2.2250738585072014e-308) || !(v <= 1.7976931348623157e308)) return p2g_log_lib(v);
    int e;
    double m = frexp(v, &e);                      // v = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752440) { m = 2.0 * m; e -= 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 21.0;
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double s2 = s + s;
    const double k = (double)e;
    const double r = fma(s2 * z, p, k * 1.9082149292705877e-10);
    return k * 0.6931471803691238 + (s2 + r);
}

// How a node set lines up with the cells along one axis:
//   CELL_AXIS: node index k covers cell k + off (both halves);
//   VERT_AXIS: node index k covers the east half of cell k-1 and the west
//              half of cell k.
enum { CELL_AXIS = 0, VERT_AXIS = 1 };

struct p2g_axes {
    int xkind, xoff, zkind, zoff;
};

// centroid 1 centres, 0 vertices, -1 Vx nodes, -2 Vz nodes.
__host__ __device__ inline p2g_axes p2g_axes_of(int centroid) {
    switch (centroid) {
    case 1:  return {CELL_AXIS, 0, CELL_AXIS, 0};
    case 0:  return {VERT_AXIS, 0, VERT_AXIS, 0};
    case -1: return {VERT_AXIS, 0, CELL_AXIS, -1};
    default: return {CELL_AXIS, -1, VERT_AXIS, 0};
    }
}

// Visit the marker ranges [b, e) (in sorted order) whose expected nearest
// node is (hk, hl).  Ranges are quarter cells; halves of one cell that are
// adjacent in the key order are merged.
template <typename F>
__device__ __forceinline__ void for_node_ranges(const unsigned *__restrict__ qoff, int Ncx, int Ncz,
                                                p2g_axes ax, int hk, int hl, F &&f) {
    int cxs[2], hxs[2], nxr;   // (cell column, half: 0 west, 1 east, 2 both)
    if (ax.xkind == CELL_AXIS) { cxs[0] = hk + ax.xoff; hxs[0] = 2; nxr = 1; }
    else { cxs[0] = hk - 1; hxs[0] = 1; cxs[1] = hk; hxs[1] = 0; nxr = 2; }
    int czs[2], hzs[2], nzr;
    if (ax.zkind == CELL_AXIS) { czs[0] = hl + ax.zoff; hzs[0] = 2; nzr = 1; }
    else { czs[0] = hl - 1; hzs[0] = 1; czs[1] = hl; hzs[1] = 0; nzr = 2; }
    for (int a = 0; a < nzr; a++) {
        const int cz = czs[a];
        if (cz < 0 || cz >= Ncz) continue;
        for (int b = 0; b < nxr; b++) {
            const int cx = cxs[b];
            if (cx < 0 || cx >= Ncx) continue;
            const unsigned base = 4u * ((unsigned)cx + (unsigned)cz * (unsigned)Ncx);
            if (hzs[a] == 2 && hxs[b] == 2) {
                f(qoff[base], qoff[base + 4]);
            } else if (hzs[a] == 2) {          // both z halves, one x half
                f(qoff[base + hxs[b]], qoff[base + hxs[b] + 1]);
                f(qoff[base + 2 + hxs[b]], qoff[base + 3 + hxs[b]]);
            } else if (hxs[b] == 2) {          // both x halves of one z half
                f(qoff[base + 2 * hzs[a]], qoff[base + 2 * hzs[a] + 2]);
            } else {
                const unsigned q = base + hxs[b] + 2 * hzs[a];
                f(qoff[q], qoff[q + 1]);
            }
        }
    }
}
