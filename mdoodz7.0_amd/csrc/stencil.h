// Matrix-free staggered-grid Stokes stencils (Picard operator and the
// viscosity-weighted augmented velocity block), shared by the FP64 outer
// operator and the FP32/FP64 multigrid levels.
//
// The coefficients are those of the Picard assembly of
// StokesAssemblyDecoupled.c:62-212 (x momentum), 562-725 (z momentum) and
// 1090-1145 (continuity) for incompressible rows without stress boundary
// conditions (SxxBC = SzzBC = 0) and without air cells: every row is
// multiplied by the cell volume dx*dz, Dirichlet-type neighbours (tags 0, 11,
// 13) carry no column (their correction is zero), and a tag-13 (free slip)
// or tag-11 (no slip) ghost neighbour changes the centre coefficient through
// the (2*Dir + Reg) factor of the reference.
//
// Arrays are in grid layout: Vx (Nx x Nz+1), Vz (Nx+1 x Nz), P and eta_n on
// cells (Nx-1 x Nz-1), eta_s on vertices (Nx x Nz); entries that are not
// free unknowns are zero.
#pragma once
#include <hip/hip_runtime.h>

struct sgrid {
    int Nx, Nz;            // vertices
    double dx, dz;
    // boundary kind of the tangential ghost rows / columns: 1 = free slip
    // (tag 13), 2 = no slip (tag 11); normal components are Dirichlet.
    int ghostS, ghostN, ghostW, ghostE;
    // penalty weight of the grad-div term: 0 = gamma * eta_n (viscosity
    // weighted), 1 = gamma alone (KillerSolver's constant penalty/dV)
    int pen_const;
};

__host__ __device__ inline int nvx(const sgrid &g) { return g.Nx * (g.Nz + 1); }
__host__ __device__ inline int nvz(const sgrid &g) { return (g.Nx + 1) * g.Nz; }
__host__ __device__ inline int ncell(const sgrid &g) { return (g.Nx - 1) * (g.Nz - 1); }

// Free velocity unknowns: Vx rows 1..Nz-1, columns 1..Nx-2; Vz rows
// 1..Nz-2, columns 1..Nx-1 (walls carry the Dirichlet normal component,
// ghost rows/columns the tangential boundary condition).
__host__ __device__ inline bool vx_free(const sgrid &g, int k, int l) {
    return k > 0 && k < g.Nx - 1 && l > 0 && l < g.Nz;
}
__host__ __device__ inline bool vz_free(const sgrid &g, int k, int l) {
    return k > 0 && k < g.Nx && l > 0 && l < g.Nz - 1;
}

// (2*Dir + Reg) of the reference for a neighbour that may be a ghost.
template <typename T>
__device__ inline T ghost_factor(int kind, bool is_ghost) {
    if (!is_ghost) return T(1);
    return kind == 2 ? T(2) : T(0);
}

// y = A_gamma * (u, v) on the velocity unknowns, with
//   A_gamma = A + gamma * C' diag(eta_n / dV) C      (dV = dx*dz),
// i.e. the Picard block plus gamma * (grad (eta div)) scaled like the rows.
// gamma = 0 gives the Picard velocity block A itself.
template <typename T>
__device__ inline T apply_vx_row(const sgrid &g, int k, int l, const T *__restrict__ u,
                                 const T *__restrict__ v, const T *__restrict__ en,
                                 const T *__restrict__ es, T gamma, T *diag_out) {
    const int Nx = g.Nx, Ncx = Nx - 1;
    const T dx = T(g.dx), dz = T(g.dz), dV = T(g.dx * g.dz);
    const int c = k + l * Nx;
    const T D11W = T(2) * en[(k - 1) + (l - 1) * Ncx], D11E = T(2) * en[k + (l - 1) * Ncx];
    const T D33S = es[k + (l - 1) * Nx], D33N = es[k + l * Nx];
    const bool gS = (l == 1), gN = (l == g.Nz - 1);
    const T fS = ghost_factor<T>(g.ghostS, gS), fN = ghost_factor<T>(g.ghostN, gN);
    const T regS = gS ? T(0) : T(1), regN = gN ? T(0) : T(1);
    const T idx2 = T(1) / (dx * dx), idz2 = T(1) / (dz * dz), idxz = T(1) / (dx * dz);
    const T uC = (D11E + D11W) * idx2 + (D33N * fN + D33S * fS) * idz2;
    T y = uC * u[c];
    y -= D11W * idx2 * u[c - 1] + D11E * idx2 * u[c + 1];
    y -= D33S * regS * idz2 * u[c - Nx] + D33N * regN * idz2 * u[c + Nx];
    const int NxV = Nx + 1;
    const int sw = k + (l - 1) * NxV, nw = k + l * NxV;
    y += (-D33S * v[sw] + D33S * v[sw + 1] + D33N * v[nw] - D33N * v[nw + 1]) * idxz;
    T d = uC;
    if (gamma != T(0)) {
        // gamma*dV*(eta_W div_W - eta_E div_E)/dx with div = C u / dV
        const int cW = (k - 1) + (l - 1) * Ncx, cE = k + (l - 1) * Ncx;
        const T divW = (u[c] - u[c - 1]) / dx + (v[k + l * NxV] - v[k + (l - 1) * NxV]) / dz;
        const T divE = (u[c + 1] - u[c]) / dx + (v[k + 1 + l * NxV] - v[k + 1 + (l - 1) * NxV]) / dz;
        const T wW = g.pen_const ? T(1) : en[cW], wE = g.pen_const ? T(1) : en[cE];
        y += gamma * (wW * divW - wE * divE) / dx;
        d += gamma * (wW + wE) * idx2;
    }
    if (diag_out) *diag_out = d * dV;
    return y * dV;
}

template <typename T>
__device__ inline T apply_vz_row(const sgrid &g, int k, int l, const T *__restrict__ u,
                                 const T *__restrict__ v, const T *__restrict__ en,
                                 const T *__restrict__ es, T gamma, T *diag_out) {
    const int Nx = g.Nx, Ncx = Nx - 1, NxV = Nx + 1;
    const T dx = T(g.dx), dz = T(g.dz), dV = T(g.dx * g.dz);
    const int c = k + l * NxV;
    const T D22S = T(2) * en[(k - 1) + (l - 1) * Ncx], D22N = T(2) * en[(k - 1) + l * Ncx];
    const T D33W = es[(k - 1) + l * Nx], D33E = es[k + l * Nx];
    const bool gW = (k == 1), gE = (k == Nx - 1);
    const T fW = ghost_factor<T>(g.ghostW, gW), fE = ghost_factor<T>(g.ghostE, gE);
    const T regW = gW ? T(0) : T(1), regE = gE ? T(0) : T(1);
    const T idx2 = T(1) / (dx * dx), idz2 = T(1) / (dz * dz), idxz = T(1) / (dx * dz);
    const T vC = (D33E * fE + D33W * fW) * idx2 + (D22N + D22S) * idz2;
    T y = vC * v[c];
    y -= D33W * regW * idx2 * v[c - 1] + D33E * regE * idx2 * v[c + 1];
    y -= D22S * idz2 * v[c - NxV] + D22N * idz2 * v[c + NxV];
    const int s = k + l * Nx;   // Vx(k, l) is SE, Vx(k-1, l) SW
    y += (-D33W * u[s - 1] + D33E * u[s] + D33W * u[s - 1 + Nx] - D33E * u[s + Nx]) * idxz;
    T d = vC;
    if (gamma != T(0)) {
        const int cS = (k - 1) + (l - 1) * Ncx, cN = (k - 1) + l * Ncx;
        const T divS = (u[k + l * Nx] - u[k - 1 + l * Nx]) / dx + (v[c] - v[c - NxV]) / dz;
        const T divN = (u[k + (l + 1) * Nx] - u[k - 1 + (l + 1) * Nx]) / dx + (v[c + NxV] - v[c]) / dz;
        const T wS = g.pen_const ? T(1) : en[cS], wN = g.pen_const ? T(1) : en[cN];
        y += gamma * (wS * divS - wN * divN) / dz;
        d += gamma * (wS + wN) * idz2;
    }
    if (diag_out) *diag_out = d * dV;
    return y * dV;
}

// The same two rows with the grid constants precomputed (reciprocals
// instead of divisions) and the diagonal returned: used by the multigrid
// smoother, whose result only has to be a good approximate inverse, so the
// last-bit difference from the divisions of apply_vx_row does not matter.
template <typename T>
struct stc {
    T dV, idx, idz, idx2, idz2, idxz;
};
template <typename T>
inline stc<T> make_stc(const sgrid &g) {
    return stc<T>{T(g.dx * g.dz), T(1.0 / g.dx), T(1.0 / g.dz), T(1.0 / (g.dx * g.dx)), T(1.0 / (g.dz * g.dz)),
                  T(1.0 / (g.dx * g.dz))};
}
template <typename T>
__device__ inline T apply_vx_rcp(const sgrid &g, const stc<T> &cf, int k, int l, const T *__restrict__ u,
                                 const T *__restrict__ v, const T *__restrict__ en, const T *__restrict__ es,
                                 T gamma, T *diag_out) {
    const int Nx = g.Nx, Ncx = Nx - 1, NxV = Nx + 1;
    const int c = k + l * Nx, cW = (k - 1) + (l - 1) * Ncx, cE = k + (l - 1) * Ncx;
    const T enW = en[cW], enE = en[cE];
    const T D33S = es[k + (l - 1) * Nx], D33N = es[k + l * Nx];
    const bool gS = (l == 1), gN = (l == g.Nz - 1);
    const T fS = ghost_factor<T>(g.ghostS, gS), fN = ghost_factor<T>(g.ghostN, gN);
    const T regS = gS ? T(0) : T(1), regN = gN ? T(0) : T(1);
    const T uc = u[c], uw = u[c - 1], ue = u[c + 1];
    const int sw = k + (l - 1) * NxV, nw = k + l * NxV;
    const T vsw = v[sw], vse = v[sw + 1], vnw = v[nw], vne = v[nw + 1];
    const T uC = T(2) * (enE + enW) * cf.idx2 + (D33N * fN + D33S * fS) * cf.idz2;
    T y = uC * uc - T(2) * cf.idx2 * (enW * uw + enE * ue);
    y -= cf.idz2 * (D33S * regS * u[c - Nx] + D33N * regN * u[c + Nx]);
    y += (D33S * (vse - vsw) + D33N * (vnw - vne)) * cf.idxz;
    // gamma * (eta_W div_W - eta_E div_E) / dx
    const T divW = (uc - uw) * cf.idx + (vnw - vsw) * cf.idz;
    const T divE = (ue - uc) * cf.idx + (vne - vse) * cf.idz;
    const T wW = g.pen_const ? T(1) : enW, wE = g.pen_const ? T(1) : enE;
    y += gamma * (wW * divW - wE * divE) * cf.idx;
    *diag_out = (uC + gamma * (wW + wE) * cf.idx2) * cf.dV;
    return y * cf.dV;
}
template <typename T>
__device__ inline T apply_vz_rcp(const sgrid &g, const stc<T> &cf, int k, int l, const T *__restrict__ u,
                                 const T *__restrict__ v, const T *__restrict__ en, const T *__restrict__ es,
                                 T gamma, T *diag_out) {
    const int Nx = g.Nx, Ncx = Nx - 1, NxV = Nx + 1;
    const int c = k + l * NxV, cS = (k - 1) + (l - 1) * Ncx, cN = (k - 1) + l * Ncx;
    const T enS = en[cS], enN = en[cN];
    const T D33W = es[(k - 1) + l * Nx], D33E = es[k + l * Nx];
    const bool gW = (k == 1), gE = (k == Nx - 1);
    const T fW = ghost_factor<T>(g.ghostW, gW), fE = ghost_factor<T>(g.ghostE, gE);
    const T regW = gW ? T(0) : T(1), regE = gE ? T(0) : T(1);
    const T vc = v[c], vs = v[c - NxV], vn = v[c + NxV];
    const int s = k + l * Nx;   // Vx(k, l) is SE, Vx(k-1, l) SW
    const T usw = u[s - 1], use = u[s], unw = u[s - 1 + Nx], une = u[s + Nx];
    const T vC = (D33E * fE + D33W * fW) * cf.idx2 + T(2) * (enN + enS) * cf.idz2;
    T y = vC * vc - cf.idx2 * (D33W * regW * v[c - 1] + D33E * regE * v[c + 1]);
    y -= T(2) * cf.idz2 * (enS * vs + enN * vn);
    y += (D33W * (unw - usw) + D33E * (use - une)) * cf.idxz;
    const T divS = (use - usw) * cf.idx + (vc - vs) * cf.idz;
    const T divN = (une - unw) * cf.idx + (vn - vc) * cf.idz;
    const T wS = g.pen_const ? T(1) : enS, wN = g.pen_const ? T(1) : enN;
    y += gamma * (wS * divS - wN * divN) * cf.idz;
    *diag_out = (vC + gamma * (wS + wN) * cf.idz2) * cf.dV;
    return y * cf.dV;
}

// Diagonal of A_gamma (Jacobi scaling of the smoothers)
template <typename T>
__device__ inline T diag_vx(const sgrid &g, int k, int l, const T *__restrict__ en,
                            const T *__restrict__ es, T gamma) {
    const int Nx = g.Nx, Ncx = Nx - 1;
    const T dx = T(g.dx), dz = T(g.dz);
    const T enW = en[(k - 1) + (l - 1) * Ncx], enE = en[k + (l - 1) * Ncx];
    const T fS = ghost_factor<T>(g.ghostS, l == 1), fN = ghost_factor<T>(g.ghostN, l == g.Nz - 1);
    const T wW = g.pen_const ? T(1) : enW, wE = g.pen_const ? T(1) : enE;
    const T d = T(2) * (enE + enW) / (dx * dx) + (es[k + l * Nx] * fN + es[k + (l - 1) * Nx] * fS) / (dz * dz) +
                gamma * (wW + wE) / (dx * dx);
    return d * T(g.dx * g.dz);
}
template <typename T>
__device__ inline T diag_vz(const sgrid &g, int k, int l, const T *__restrict__ en,
                            const T *__restrict__ es, T gamma) {
    const int Nx = g.Nx, Ncx = Nx - 1;
    const T dx = T(g.dx), dz = T(g.dz);
    const T enS = en[(k - 1) + (l - 1) * Ncx], enN = en[(k - 1) + l * Ncx];
    const T fW = ghost_factor<T>(g.ghostW, k == 1), fE = ghost_factor<T>(g.ghostE, k == Nx - 1);
    const T wS = g.pen_const ? T(1) : enS, wN = g.pen_const ? T(1) : enN;
    const T d = (es[k + l * Nx] * fE + es[(k - 1) + l * Nx] * fW) / (dx * dx) + T(2) * (enN + enS) / (dz * dz) +
                gamma * (wS + wN) / (dz * dz);
    return d * T(g.dx * g.dz);
}

// Pressure-gradient part of a momentum row (the B block): Vx row (k,l).
template <typename T>
__device__ inline T grad_vx_row(const sgrid &g, int k, int l, const T *__restrict__ p) {
    const int Ncx = g.Nx - 1;
    return (p[k + (l - 1) * Ncx] - p[(k - 1) + (l - 1) * Ncx]) * T(g.dz);
}
template <typename T>
__device__ inline T grad_vz_row(const sgrid &g, int k, int l, const T *__restrict__ p) {
    const int Ncx = g.Nx - 1;
    return (p[(k - 1) + l * Ncx] - p[(k - 1) + (l - 1) * Ncx]) * T(g.dx);
}
// Continuity row (the C block) of cell (k, l): dV * div.
template <typename T>
__device__ inline T div_cell(const sgrid &g, int k, int l, const T *__restrict__ u,
                             const T *__restrict__ v) {
    const int Nx = g.Nx, NxV = Nx + 1;
    return (u[k + 1 + (l + 1) * Nx] - u[k + (l + 1) * Nx]) * T(g.dz) +
           (v[k + 1 + (l + 1) * NxV] - v[k + 1 + l * NxV]) * T(g.dx);
}
