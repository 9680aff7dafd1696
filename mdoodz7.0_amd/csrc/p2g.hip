// Marker -> grid interpolation (P2Mastah, ParticleRoutines.c:2170-2665 and
// P2Mastah_Fused, ParticleRoutines.c:1895-2163) on cell-sorted markers.
//
// One thread owns one node and walks the quarter-cell ranges of the markers
// whose nearest node it is (plus the 3x3 neighbour homes for the 9-cell
// stencil), summing weights and weighted values in a fixed order: no
// atomics, bit-reproducible from run to run.  Each marker's nearest node is
// recomputed with the reference expression; a marker whose exact node differs
// from the one its sort key implies (only possible within round-off of a
// cell edge for the Vx/Vz node sets) raises a flag and the call is redone
// with a one-node wider search that tests every marker exactly.
#include "mdg_internal.h"
#include "p2g_common.h"

#define P2G_MAXF 4
#define P2G_MAXPH 32
#define P2G_MAXTAB 1024
#define P2G_TILE 256     // nodes per block of the row-tiled kernels

struct p2g_fld {
    const double      *src;   // per-phase table (flag 0) or sorted per-marker (flag 1)
    double            *dst;
    const signed char *ntype;
    int                flag, avg, n_src;
};
struct p2g_set {
    p2g_fld f[P2G_MAXF];
};

struct p2g_geo {
    int nx, nz, Ncx, Ncz;
    p2g_axes ax;
    const double *X, *Z;
    double X0, Z0, dx, dz, dx_itp, dz_itp;
    double idx, idz, ix_itp, iz_itp;   // reciprocals of dx, dz, dx_itp, dz_itp
    int R, S;   // stencil radius (0: 1-cell, 1: 9-cell) and search radius
    int set;    // node set: 0 centres, 1 vertices, 2 Vx, 3 Vz
};

template <typename F>
__device__ __forceinline__ void p2g_visit(const p2g_geo &g, const unsigned *__restrict__ qoff,
                                          const double *__restrict__ mx,
                                          const double *__restrict__ mz, int k, int l,
                                          int *__restrict__ stray, F &&acc) {
    for (int b = -g.S; b <= g.S; b++) {
        const int hl = l + b;
        if (hl < 0 || hl >= g.nz) continue;
        for (int a = -g.S; a <= g.S; a++) {
            const int hk = k + a;
            if (hk < 0 || hk >= g.nx) continue;
            for_node_ranges(qoff, g.Ncx, g.Ncz, g.ax, hk, hl, [&](unsigned s, unsigned e) {
                for (unsigned i = s; i < e; i++) {
                    const double x = mx[i], z = mz[i];
                    const int ip = near_index_fast(x, g.X0, g.dx, g.idx, g.nx);
                    const int jp = near_index_fast(z, g.Z0, g.dz, g.idz, g.nz);
                    if (g.S == g.R && (ip != hk || jp != hl)) { *stray = 1; continue; }
                    const int ox = k - ip, oz = l - jp;
                    if (ox < -g.R || ox > g.R || oz < -g.R || oz > g.R) continue;
                    // distances exactly as the reference forms them:
                    // |X[ip] - x| for the home node, |X[ip] +- dx - x| otherwise
                    const double dxm = (ox == 0) ? fabs(g.X[ip] - x) : fabs(g.X[ip] + ox * g.dx - x);
                    const double dzm = (oz == 0) ? fabs(g.Z[jp] - z) : fabs(g.Z[jp] + oz * g.dz - z);
                    acc(i, (1.0 - dxm * g.ix_itp) * (1.0 - dzm * g.iz_itp));
                }
            });
        }
    }
}

__device__ __forceinline__ double mark_val(const p2g_fld &f, int p, unsigned i) {
    // a phase outside the table reads as NaN instead of out of bounds
    double v = (f.flag == 0) ? ((p < f.n_src) ? f.src[p] : __builtin_nan("")) : f.src[i];
    if (f.avg == 1) v = 1.0 / v;
    if (f.avg == 2) v = p2g_log(v);
    return v;
}

template <int NF>
__global__ void __launch_bounds__(256) k_p2g(p2g_geo g, p2g_set fs, const unsigned *__restrict__ qoff,
                                             const double *__restrict__ mx,
                                             const double *__restrict__ mz,
                                             const int *__restrict__ mph, int *__restrict__ stray) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= g.nx * g.nz) return;
    const int k = node % g.nx, l = node / g.nx;
    double W = 0.0, BW[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) BW[f] = 0.0;
    p2g_visit(g, qoff, mx, mz, k, l, stray, [&](unsigned i, double w) {
        const int p = mph[i];
        W += w;
#pragma unroll
        for (int f = 0; f < NF; f++) BW[f] += mark_val(fs.f[f], p, i) * w;
    });
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const p2g_fld &F = fs.f[f];
        if (fabs(W) < 1e-30) continue;
        if (F.ntype && (F.ntype[node] == 30 || F.ntype[node] == 31)) continue;
        double v = BW[f] / W;
        if (F.avg == 1) v = 1.0 / v;
        if (F.avg == 2) v = exp(v);
        F.dst[node] = v;
    }
}

// The 9-cell stencil when every marker's nearest node is the one its sort
// key implies (the node set's cached state says so).  A block owns
// P2G_TILE consecutive nodes of one node row; the markers that reach them
// lie in the 3 (cell-centred rows) or 4 (rows on cell edges) cell rows
// around it, over the block's column span plus one or two cells.  The block
// streams those rows in chunks with coalesced loads into LDS as (x, z, home
// node, transformed value), each marker's home taken from its sort key and
// its value transformed once per chunk instead of once per visit; every
// node thread then walks, chunk by chunk, the sorted span of its own
// neighbour cells and sums the markers whose home is one of its 3 x 3
// neighbours, with the reference's weight expression (|X[home] + o dx - x|
// for a neighbour at offset o).  Fixed order: bit-reproducible.
template <int NF> struct tile9_cap { static constexpr int v = (NF <= 2) ? 1024 : 640; };

template <int NF>
__global__ void __launch_bounds__(P2G_TILE)
k_p2g9_tile(p2g_geo g, p2g_set fs, const unsigned *__restrict__ qoff, const unsigned *__restrict__ keys,
            const double *__restrict__ mx, const double *__restrict__ mz, const int *__restrict__ mph,
            int nbx, int nblocks) {
    constexpr int CAP = tile9_cap<NF>::v;
    __shared__ double lx[CAP], lz[CAP], lv[NF][CAP];
    __shared__ int lh[CAP];
    const int b = blockIdx.x, xcd = b & 7, per = nblocks >> 3, rem = nblocks & 7;
    const int tile = xcd * per + min(xcd, rem) + (b >> 3);
    const int l = tile / nbx, k0 = (tile - l * nbx) * P2G_TILE;
    const int kend = min(k0 + P2G_TILE, g.nx);
    const int t = threadIdx.x, k = k0 + t;
    const bool own = k < kend;
    const bool xc = (g.ax.xkind == CELL_AXIS), zc = (g.ax.zkind == CELL_AXIS);
    // cell columns a node (and the block) draws from; cell rows of node row l
    const int bx0 = max(xc ? k0 - 1 + g.ax.xoff : k0 - 2, 0);
    const int bx1 = min(xc ? kend + g.ax.xoff : kend, g.Ncx - 1);
    const int tx0 = max(xc ? k - 1 + g.ax.xoff : k - 2, 0);
    const int tx1 = min(xc ? k + 1 + g.ax.xoff : k + 1, g.Ncx - 1);
    const int cz0 = max(zc ? l - 1 + g.ax.zoff : l - 2, 0);
    const int cz1 = min(zc ? l + 1 + g.ax.zoff : l + 1, g.Ncz - 1);
    const int node = own ? l * g.nx + k : -1;
    double W = 0.0, BW[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) BW[f] = 0.0;
    for (int cz = cz0; cz <= cz1 && bx0 <= bx1; cz++) {
        const unsigned S = qoff[4u * (unsigned)(bx0 + cz * g.Ncx)];
        const unsigned E = qoff[4u * (unsigned)(bx1 + cz * g.Ncx) + 4u];
        unsigned ts = 0u, te = 0u;   // this node's span in the row
        if (own && tx0 <= tx1) {
            ts = qoff[4u * (unsigned)(tx0 + cz * g.Ncx)];
            te = qoff[4u * (unsigned)(tx1 + cz * g.Ncx) + 4u];
        }
        for (unsigned cs = S; cs < E; cs += CAP) {
            const unsigned ce = min(E, cs + (unsigned)CAP);
            __syncthreads();   // previous chunk consumed
            for (unsigned j = cs + t; j < ce; j += P2G_TILE) {
                const unsigned q = j - cs;
                lx[q] = mx[j];
                lz[q] = mz[j];
                const unsigned key = keys[j], cell = key >> 2, qd = key & 3u;
                const int cx = (int)(cell % (unsigned)g.Ncx), czm = (int)(cell / (unsigned)g.Ncx);
                const int hk = xc ? cx - g.ax.xoff : cx + (int)(qd & 1u);
                const int hl = zc ? czm - g.ax.zoff : czm + (int)(qd >> 1);
                lh[q] = hk + hl * g.nx;
                const int p = mph[j];
#pragma unroll
                for (int f = 0; f < NF; f++) lv[f][q] = mark_val(fs.f[f], p, j);
            }
            __syncthreads();
            const unsigned a = max(ts, cs), e = min(te, ce);
            for (unsigned i = a; i < e; i++) {
                const unsigned q = i - cs;
                const int h = lh[q], hl = h / g.nx, hk = h - hl * g.nx;
                const int ox = k - hk, oz = l - hl;
                if (ox < -1 || ox > 1 || oz < -1 || oz > 1) continue;
                const double x = lx[q], z = lz[q];
                const double dxm = (ox == 0) ? fabs(g.X[hk] - x) : fabs(g.X[hk] + ox * g.dx - x);
                const double dzm = (oz == 0) ? fabs(g.Z[hl] - z) : fabs(g.Z[hl] + oz * g.dz - z);
                const double w = (1.0 - dxm * g.ix_itp) * (1.0 - dzm * g.iz_itp);
                W += w;
#pragma unroll
                for (int f = 0; f < NF; f++) BW[f] += lv[f][q] * w;
            }
        }
    }
    if (!own) return;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const p2g_fld &F = fs.f[f];
        if (fabs(W) < 1e-30) continue;
        if (F.ntype && (F.ntype[node] == 30 || F.ntype[node] == 31)) continue;
        double v = BW[f] / W;
        if (F.avg == 1) v = 1.0 / v;
        if (F.avg == 2) v = exp(v);
        F.dst[node] = v;
    }
}

// Per-marker weight of the 1-cell stencil for one node set, in sorted
// order: the same nearest node and weight the kernels compute inline, kept
// while the markers do not move (MDOODZ interpolates 15-25 fields per step
// from the same positions).  A marker whose nearest node is not the one its
// (cell, quarter) key implies raises the flag; the node set then keeps the
// inline path, which handles it.
__global__ void __launch_bounds__(256) k_p2g_weights(p2g_geo g, int n, const unsigned *__restrict__ keys,
                                                     const double *__restrict__ mx, const double *__restrict__ mz,
                                                     double *__restrict__ w, int *__restrict__ stray) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = mx[i], z = mz[i];
    const int ip = near_index_fast(x, g.X0, g.dx, g.idx, g.nx);
    const int jp = near_index_fast(z, g.Z0, g.dz, g.idz, g.nz);
    const unsigned key = keys[i], cell = key >> 2, qd = key & 3u;
    if (cell >= (unsigned)g.Ncx * (unsigned)g.Ncz) { w[i] = 0.0; return; }   // phase -1: in no range
    const int cx = (int)(cell % (unsigned)g.Ncx), cz = (int)(cell / (unsigned)g.Ncx);
    const int hx = (int)(qd & 1u), hz = (int)(qd >> 1);
    const int ek = (g.ax.xkind == CELL_AXIS) ? cx - g.ax.xoff : cx + hx;
    const int el = (g.ax.zkind == CELL_AXIS) ? cz - g.ax.zoff : cz + hz;
    if (ip != ek || jp != el) *stray = 1;
    const double dxm = fabs(g.X[ip] - x), dzm = fabs(g.Z[jp] - z);
    w[i] = (1.0 - dxm * g.ix_itp) * (1.0 - dzm * g.iz_itp);
}

// Row-tiled variant for the 1-cell stencil.  A block owns P2G_TILE
// consecutive nodes of ROWS node rows; the markers of the cell rows those
// nodes draw from are contiguous in sorted order, so the block streams them
// with coalesced loads, turns each into (home node, w, w*value) in LDS once,
// and each node thread then sums its own quarter-cell ranges out of LDS in
// the same marker order as k_p2g.  Node sets whose rows sit on cell edges
// (vertices, Vz nodes) take half of each of two cell rows; with ROWS = 4
// rows per block a cell row is computed (K+1)/K = 1.25 times instead of
// twice.  Consecutive tiles are handed to the same XCD (blocks are
// dispatched round-robin over the 8 XCDs), so a cell row shared by two
// blocks is read from HBM once and from that XCD's L2 the second time.
template <int NF> struct tile_cap { static constexpr int v = (NF <= 2) ? 1280 : 768; };

template <int NF, int ROWS, bool PRE>
__global__ void __launch_bounds__(P2G_TILE)
k_p2g_tile(p2g_geo g, p2g_set fs, const unsigned *__restrict__ qoff, const double *__restrict__ mx,
           const double *__restrict__ mz, const double *__restrict__ mw, const int *__restrict__ mph,
           int need_phase, int nbx, int nblocks, int *__restrict__ stray) {
    constexpr int CAP = tile_cap<NF>::v;
    __shared__ double lw[CAP];
    __shared__ double lwv[NF][CAP];
    __shared__ int lhome[PRE ? 1 : CAP];   // homes are checked when the weights are cached
    // XCD-aware order: hardware block b runs on XCD b % 8; XCD x gets the
    // contiguous logical tiles [start(x), start(x) + count(x)).
    const int b = blockIdx.x, xcd = b & 7, per = nblocks >> 3, rem = nblocks & 7;
    const int tile = xcd * per + min(xcd, rem) + (b >> 3);
    const int lt = tile / nbx, k0 = (tile - lt * nbx) * P2G_TILE;
    const int l0 = lt * ROWS, l1 = min(l0 + ROWS, g.nz);     // node rows [l0, l1)
    const int kend = min(k0 + P2G_TILE, g.nx);
    const int t = threadIdx.x, k = k0 + t;
    const bool own = k < kend;
    // each node's marker ranges (at most 4 for the 1-cell stencil)
    unsigned rs[ROWS][4], re[ROWS][4];
    int nr[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        nr[r] = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) { rs[r][q] = 0u; re[r][q] = 0u; }
        if (own && l0 + r < l1)
            for_node_ranges(qoff, g.Ncx, g.Ncz, g.ax, k, l0 + r, [&](unsigned s0, unsigned e0) {
                if (nr[r] < 4) {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (q == nr[r]) { rs[r][q] = s0; re[r][q] = e0; }
                    nr[r]++;
                }
            });
    }
    // the block's cell rows and column span
    int cx0, cx1, cz0, cz1;
    if (g.ax.xkind == CELL_AXIS) { cx0 = k0 + g.ax.xoff; cx1 = kend - 1 + g.ax.xoff; }
    else { cx0 = k0 - 1; cx1 = kend - 1; }
    if (g.ax.zkind == CELL_AXIS) { cz0 = l0 + g.ax.zoff; cz1 = l1 - 1 + g.ax.zoff; }
    else { cz0 = l0 - 1; cz1 = l1 - 1; }
    cx0 = max(cx0, 0); cx1 = min(cx1, g.Ncx - 1);
    cz0 = max(cz0, 0); cz1 = min(cz1, g.Ncz - 1);
    double W[ROWS], BW[ROWS][NF];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        W[r] = 0.0;
#pragma unroll
        for (int f = 0; f < NF; f++) BW[r][f] = 0.0;
    }
    bool bad = false;
    for (int cz = cz0; cz <= cz1 && cx0 <= cx1; cz++) {
        const unsigned S = qoff[4u * (unsigned)(cx0 + cz * g.Ncx)];
        const unsigned E = qoff[4u * (unsigned)(cx1 + cz * g.Ncx) + 4u];
        for (unsigned cs = S; cs < E; cs += CAP) {
            const unsigned ce = min(E, cs + (unsigned)CAP);
            __syncthreads();   // previous chunk fully consumed
            // Every thread first issues all loads of its (up to PER) markers so
            // that the whole chunk is in flight at once, then computes.
            constexpr int PER = CAP / P2G_TILE;
            double xs[PER], zs[PER], vs[NF][PER];
            int ps[PER];
#pragma unroll
            for (int m = 0; m < PER; m++) {
                const unsigned j = cs + t + (unsigned)m * P2G_TILE;
                xs[m] = 0.0; zs[m] = 0.0; ps[m] = 0;
#pragma unroll
                for (int f = 0; f < NF; f++) vs[f][m] = 0.0;
                if (j < ce) {
                    if (PRE) {
                        xs[m] = mw[j];      // the cached weight
                    } else {
                        xs[m] = mx[j];
                        zs[m] = mz[j];
                    }
                    if (need_phase) ps[m] = mph[j];
#pragma unroll
                    for (int f = 0; f < NF; f++)
                        if (fs.f[f].flag == 1) vs[f][m] = fs.f[f].src[j];
                }
            }
#pragma unroll
            for (int m = 0; m < PER; m++) {
                const unsigned j = cs + t + (unsigned)m * P2G_TILE;
                if (j >= ce) continue;
                const unsigned q = j - cs;
                double w;
                if (PRE) {
                    w = xs[m];
                } else {
                    const double x = xs[m], z = zs[m];
                    const int ip = near_index_fast(x, g.X0, g.dx, g.idx, g.nx);
                    const int jp = near_index_fast(z, g.Z0, g.dz, g.idz, g.nz);
                    const double dxm = fabs(g.X[ip] - x), dzm = fabs(g.Z[jp] - z);
                    w = (1.0 - dxm * g.ix_itp) * (1.0 - dzm * g.iz_itp);
                    lhome[q] = ip + jp * g.nx;
                }
                lw[q] = w;
#pragma unroll
                for (int f = 0; f < NF; f++) {
                    const p2g_fld &F = fs.f[f];
                    const int p = ps[m];
                    double v = (F.flag == 0) ? ((p < F.n_src) ? F.src[p] : __builtin_nan("")) : vs[f][m];
                    if (F.avg == 1) v = 1.0 / v;
                    if (F.avg == 2) v = p2g_log(v);
                    lwv[f][q] = v * w;
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < ROWS; r++) {
                const int node = (l0 + r) * g.nx + k;
#pragma unroll
                for (int qr = 0; qr < 4; qr++) {
                    if (qr >= nr[r]) continue;
                    const unsigned a = max(rs[r][qr], cs), e = min(re[r][qr], ce);
                    for (unsigned i = a; i < e; i++) {
                        const unsigned q = i - cs;
                        if (!PRE && lhome[q] != node) { bad = true; continue; }
                        W[r] += lw[q];
#pragma unroll
                        for (int f = 0; f < NF; f++) BW[r][f] += lwv[f][q];
                    }
                }
            }
        }
    }
    if (bad) *stray = 1;
    if (!own) return;
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        if (l0 + r >= l1) break;
        const int node = (l0 + r) * g.nx + k;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const p2g_fld &F = fs.f[f];
            if (fabs(W[r]) < 1e-30) continue;
            if (F.ntype && (F.ntype[node] == 30 || F.ntype[node] == 31)) continue;
            double v = BW[r][f] / W[r];
            if (F.avg == 1) v = 1.0 / v;
            if (F.avg == 2) v = exp(v);
            F.dst[node] = v;
        }
    }
}

// Phase fractions (P2Mastah with prop = 1): per-thread phase accumulators in LDS.
__global__ void __launch_bounds__(128) k_p2g_phase(p2g_geo g, int nph, const signed char *ntype,
                                                   const unsigned *__restrict__ qoff,
                                                   const double *__restrict__ mx,
                                                   const double *__restrict__ mz,
                                                   const int *__restrict__ mph,
                                                   double *__restrict__ perc,
                                                   int *__restrict__ stray) {
    __shared__ double acc[P2G_MAXPH * 128];
    const int t = threadIdx.x;
    const int node = blockIdx.x * blockDim.x + t;
    for (int p = 0; p < nph; p++) acc[p * 128 + t] = 0.0;
    if (node >= g.nx * g.nz) return;
    const int k = node % g.nx, l = node / g.nx;
    double W = 0.0;
    p2g_visit(g, qoff, mx, mz, k, l, stray, [&](unsigned i, double w) {
        W += w;
        const int p = mph[i];
        if (p >= 0 && p < nph) acc[p * 128 + t] += w;
    });
    const size_t nn = (size_t)g.nx * g.nz;
    const bool skip = fabs(W) < 1e-30 || (ntype && (ntype[node] == 30 || ntype[node] == 31));
    for (int p = 0; p < nph; p++) {
        const double a = acc[p * 128 + t];
        perc[p * nn + node] = skip ? a : a / W;
    }
}

extern int mdg_gather_sorted(mdg_ctx *c, const double *src, double *dst);

static int setup_geo(mdg_ctx *c, int centroid, int stencil, p2g_geo *g) {
    node_set ns;
    int rc = mdg_node_set(c, centroid, &ns);
    if (rc) return rc;
    if (stencil != 1 && stencil != 9) return mdg_fail(c, MDG_ERR_ARG, "interp_stencil must be 1 or 9");
    if (!c->qoff.p) return mdg_fail(c, MDG_ERR_STATE, "mdg_markers_set has not been called");
    g->nx = ns.nx; g->nz = ns.nz; g->Ncx = c->Ncx; g->Ncz = c->Ncz;
    g->ax = p2g_axes_of(centroid);
    g->X = ns.X; g->Z = ns.Z; g->X0 = ns.X0; g->Z0 = ns.Z0; g->dx = c->dx; g->dz = c->dz;
    g->dx_itp = (stencil == 9) ? 3.0 * c->dx / 2.0 : c->dx / 2.0;   // ParticleRoutines.c:2218-2221
    g->dz_itp = (stencil == 9) ? 3.0 * c->dz / 2.0 : c->dz / 2.0;
    // weights multiply by the reciprocal (the reference divides; the two
    // differ by at most an ulp of each factor)
    g->idx = 1.0 / c->dx; g->idz = 1.0 / c->dz;
    g->ix_itp = 1.0 / g->dx_itp; g->iz_itp = 1.0 / g->dz_itp;
    g->R = (stencil == 9) ? 1 : 0;
    g->S = g->R;
    g->set = (centroid == 1) ? 0 : (centroid == 0) ? 1 : (centroid == -1) ? 2 : 3;
    return MDG_OK;
}

// Computes the cached 1-cell-stencil weights of g's node set if they are
// not current (once per node set after each mdg_markers_set).
static int ensure_weights(mdg_ctx *c, const p2g_geo &g) {
    if (c->mw_state[g.set] != 0) return MDG_OK;
    if (g.R != 0) {
        // the 9-cell stencil needs the node set's stray state only; the
        // cached weights are those of the 1-cell stencil
        p2g_geo g1;
        const int centroid = g.set == 0 ? 1 : g.set == 1 ? 0 : g.set == 2 ? -1 : -2;
        int rc = setup_geo(c, centroid, 1, &g1);
        return rc ? rc : ensure_weights(c, g1);
    }
    const int n = c->n_markers;
    HIPCHK(c, c->mw[g.set].reserve((size_t)n + 1));
    HIPCHK(c, hipMemsetAsync(c->flag_dev.p, 0, sizeof(int), c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (n > 0)
        hipLaunchKernelGGL(k_p2g_weights, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, g, n, (const unsigned *)c->keys2.p,
                           (const double *)c->mx.p, (const double *)c->mz.p, c->mw[g.set].p, c->flag_dev.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    int stray = 0;
    HIPCHK(c, hipMemcpyAsync(&stray, c->flag_dev.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->weights_ms = ms;
    c->mw_state[g.set] = stray ? 2 : 1;
    return MDG_OK;
}

static int launch_fields(mdg_ctx *c, p2g_geo g, const p2g_set &s, int nf) {
    const int nn = g.nx * g.nz;
    if (g.R == 0 && g.S == 0) {
        // 1-cell stencil, exact homes: row-tiled kernel
        int need_phase = 0;
        for (int f = 0; f < nf; f++) need_phase |= (s.f[f].flag == 0);
        // node rows on cell edges in z draw from two cell rows: 4 rows a block
        const int rows = (g.ax.zkind == VERT_AXIS) ? 4 : 1;
        const int nbx = (g.nx + P2G_TILE - 1) / P2G_TILE, nb = nbx * ((g.nz + rows - 1) / rows);
        dim3 grid(nb), block(P2G_TILE);
        const bool pre = (c->mw_state[g.set] == 1);
        const double *mw = c->mw[g.set].p;
#define TILE_LAUNCH(NF, R, P) hipLaunchKernelGGL((k_p2g_tile<NF, R, P>), grid, block, 0, c->stream, g, s, c->qoff.p, c->mx.p, c->mz.p, mw, c->mphase.p, need_phase, nbx, nb, c->flag_dev.p)
#define TILE_ROWS(NF) do { if (pre) { if (rows == 4) TILE_LAUNCH(NF, 4, true); else TILE_LAUNCH(NF, 1, true); } \
                           else { if (rows == 4) TILE_LAUNCH(NF, 4, false); else TILE_LAUNCH(NF, 1, false); } } while (0)
        switch (nf) {
        case 1: TILE_ROWS(1); break;
        case 2: TILE_ROWS(2); break;
        case 3: TILE_ROWS(3); break;
        default: TILE_ROWS(4); break;
        }
#undef TILE_ROWS
#undef TILE_LAUNCH
        HIPCHK(c, hipGetLastError());
        return MDG_OK;
    }
    dim3 grid(mdg_grid1d(nn, 256)), block(256);
    if (g.R == 1 && g.S == 1 && c->mw_state[g.set] == 1) {
        // 9-cell stencil on exact homes: row tiles
        const int nbx = (g.nx + P2G_TILE - 1) / P2G_TILE, nb = nbx * g.nz;
#define T9(NF) hipLaunchKernelGGL(k_p2g9_tile<NF>, dim3(nb), dim3(P2G_TILE), 0, c->stream, g, s, c->qoff.p, \
                                  (const unsigned *)c->keys2.p, c->mx.p, c->mz.p, c->mphase.p, nbx, nb)
        switch (nf) {
        case 1: T9(1); break;
        case 2: T9(2); break;
        case 3: T9(3); break;
        default: T9(4); break;
        }
#undef T9
        HIPCHK(c, hipGetLastError());
        return MDG_OK;
    }
    switch (nf) {
    case 1: hipLaunchKernelGGL(k_p2g<1>, grid, block, 0, c->stream, g, s, c->qoff.p, c->mx.p, c->mz.p, c->mphase.p, c->flag_dev.p); break;
    case 2: hipLaunchKernelGGL(k_p2g<2>, grid, block, 0, c->stream, g, s, c->qoff.p, c->mx.p, c->mz.p, c->mphase.p, c->flag_dev.p); break;
    case 3: hipLaunchKernelGGL(k_p2g<3>, grid, block, 0, c->stream, g, s, c->qoff.p, c->mx.p, c->mz.p, c->mphase.p, c->flag_dev.p); break;
    default: hipLaunchKernelGGL(k_p2g<4>, grid, block, 0, c->stream, g, s, c->qoff.p, c->mx.p, c->mz.p, c->mphase.p, c->flag_dev.p); break;
    }
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

// Runs `launch` once; if a stray marker was flagged, reruns it with the
// wide search.  The kernel time of the final launch is kept in last_ms.
template <typename L>
static int run_checked(mdg_ctx *c, p2g_geo &g, L &&launch) {
    for (int pass = 0; pass < 2; pass++) {
        HIPCHK(c, hipMemsetAsync(c->flag_dev.p, 0, sizeof(int), c->stream));
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        int rc = launch(g);
        if (rc) return rc;
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        int stray = 0;
        HIPCHK(c, hipMemcpyAsync(&stray, c->flag_dev.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        if (!stray) return MDG_OK;
        g.S = g.R + 1;
    }
    return MDG_OK;
}

struct staged_field {
    p2g_fld f;
    double *host_dst;     // non-null when results go back to a host array
    size_t  n;
};

// Resolve one field's source, destination and node tags onto the device.
static int stage_field(mdg_ctx *c, const double *src, int n_src, double *dst,
                       const signed char *ntype, int flag, int avg, size_t nn, size_t slot,
                       staged_field *sf, dbuf<double> &fld, dbuf<double> &tab,
                       dbuf<double> &nodes, dbuf<signed char> &tags) {
    const int base_flag = flag & ~MDG_SRC_SORTED;
    if (!src || !dst) return mdg_fail(c, MDG_ERR_ARG, "null field array");
    if (avg < 0 || avg > 2) return mdg_fail(c, MDG_ERR_ARG, "avg must be 0, 1 or 2");
    if (base_flag == 0) {
        if (n_src < 1 || n_src > P2G_MAXTAB)
            return mdg_fail(c, MDG_ERR_ARG, "phase table size must be 1..%d", P2G_MAXTAB);
        // tables are tiny: always copy (host or device source)
        HIPCHK(c, hipMemcpyAsync(tab.p + slot * P2G_MAXTAB, src, sizeof(double) * n_src,
                                 hipMemcpyDefault, c->stream));
        sf->f.src = tab.p + slot * P2G_MAXTAB;
    } else if (base_flag == 1) {
        if (n_src < c->n_markers) return mdg_fail(c, MDG_ERR_ARG, "per-marker field shorter than markers");
        if (flag & MDG_SRC_SORTED) {
            if (!mdg_is_device_ptr(src)) return mdg_fail(c, MDG_ERR_ARG, "sorted source must be device memory");
            sf->f.src = src;
        } else {
            double *d = fld.p + slot * (size_t)(c->n_markers + 1);
            int rc = mdg_gather_sorted(c, src, d);
            if (rc) return rc;
            sf->f.src = d;
        }
    } else {
        return mdg_fail(c, MDG_ERR_UNSUPPORTED, "flag %d (anisotropy flags are not implemented)", flag);
    }
    sf->f.flag = base_flag;
    sf->f.n_src = n_src;
    sf->f.avg = avg;
    sf->n = nn;
    if (mdg_is_device_ptr(dst)) {
        sf->f.dst = dst;
        sf->host_dst = nullptr;
    } else {
        sf->f.dst = nodes.p + slot * nn;
        sf->host_dst = dst;
        HIPCHK(c, hipMemcpyAsync(sf->f.dst, dst, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    sf->f.ntype = nullptr;
    if (ntype) {
        if (mdg_is_device_ptr(ntype)) sf->f.ntype = ntype;
        else {
            HIPCHK(c, hipMemcpyAsync(tags.p + slot * nn, ntype, nn, hipMemcpyHostToDevice, c->stream));
            sf->f.ntype = tags.p + slot * nn;
        }
    }
    return MDG_OK;
}

extern "C" int mdg_p2mastah_fused(mdg_ctx *c, int centroid, int stencil,
                                  const mdg_p2g_field *fields, int nfields) {
    if (!c) return MDG_ERR_ARG;
    if (nfields < 1 || nfields > 16 || !fields) return mdg_fail(c, MDG_ERR_ARG, "nfields must be 1..16");
    p2g_geo g;
    int rc = setup_geo(c, centroid, stencil, &g);
    if (rc || (rc = ensure_weights(c, g))) return rc;
    const size_t nn = (size_t)g.nx * g.nz, N = (size_t)c->n_markers + 1;
    HIPCHK(c, c->field_tmp.reserve(N * P2G_MAXF));
    HIPCHK(c, c->table_tmp.reserve(P2G_MAXTAB * P2G_MAXF));
    HIPCHK(c, c->node_tmp.reserve(nn * P2G_MAXF));
    HIPCHK(c, c->ntype_tmp.reserve(nn * P2G_MAXF));
    for (int f0 = 0; f0 < nfields; f0 += P2G_MAXF) {
        const int nf = (nfields - f0 < P2G_MAXF) ? nfields - f0 : P2G_MAXF;
        staged_field sf[P2G_MAXF];
        p2g_set s = {};
        for (int q = 0; q < nf; q++) {
            const mdg_p2g_field &F = fields[f0 + q];
            rc = stage_field(c, F.src, F.n_src, F.dst, F.node_type, F.flag, F.avg, nn, q, &sf[q],
                             c->field_tmp, c->table_tmp, c->node_tmp, c->ntype_tmp);
            if (rc) return rc;
            s.f[q] = sf[q].f;
        }
        rc = run_checked(c, g, [&](p2g_geo &gg) { return launch_fields(c, gg, s, nf); });
        if (rc) return rc;
        g.S = g.R;
        for (int q = 0; q < nf; q++)
            if (sf[q].host_dst)
                HIPCHK(c, hipMemcpyAsync(sf[q].host_dst, sf[q].f.dst, nn * sizeof(double),
                                         hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_p2mastah(mdg_ctx *c, const double *mat_prop, int n_mat_prop, double *node_field,
                            const signed char *node_type, int flag, int avg, int prop,
                            int centroid, int stencil, int nb_phases, double *const *phase_perc) {
    if (!c) return MDG_ERR_ARG;
    if (prop == 0) {
        mdg_p2g_field f = {mat_prop, n_mat_prop, node_field, node_type, flag, avg};
        return mdg_p2mastah_fused(c, centroid, stencil, &f, 1);
    }
    if (prop != 1) return mdg_fail(c, MDG_ERR_ARG, "prop must be 0 or 1");
    if (centroid != 0 && centroid != 1)
        return mdg_fail(c, MDG_ERR_ARG, "phase fractions exist on centres and vertices only");
    if (nb_phases < 1 || nb_phases > P2G_MAXPH || !phase_perc)
        return mdg_fail(c, MDG_ERR_ARG, "nb_phases must be 1..%d", P2G_MAXPH);
    p2g_geo g;
    int rc = setup_geo(c, centroid, stencil, &g);
    if (rc) return rc;
    const size_t nn = (size_t)g.nx * g.nz;
    HIPCHK(c, c->node_tmp.reserve(nn * nb_phases));
    HIPCHK(c, c->ntype_tmp.reserve(nn));
    const signed char *nt = nullptr;
    if (node_type) {
        if (mdg_is_device_ptr(node_type)) nt = node_type;
        else {
            HIPCHK(c, hipMemcpyAsync(c->ntype_tmp.p, node_type, nn, hipMemcpyHostToDevice, c->stream));
            nt = c->ntype_tmp.p;
        }
    }
    rc = run_checked(c, g, [&](p2g_geo &gg) {
        hipLaunchKernelGGL(k_p2g_phase, dim3(mdg_grid1d((long)nn, 128)), dim3(128), 0, c->stream, gg,
                           nb_phases, nt, c->qoff.p, c->mx.p, c->mz.p, c->mphase.p, c->node_tmp.p,
                           c->flag_dev.p);
        HIPCHK(c, hipGetLastError());
        return MDG_OK;
    });
    if (rc) return rc;
    for (int p = 0; p < nb_phases; p++) {
        if (!phase_perc[p]) return mdg_fail(c, MDG_ERR_ARG, "phase_perc[%d] is null", p);
        HIPCHK(c, hipMemcpyAsync(phase_perc[p], c->node_tmp.p + p * nn, nn * sizeof(double),
                                 hipMemcpyDefault, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}
