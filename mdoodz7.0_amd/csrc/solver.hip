// Stokes solve on the GPU, replacing KillerSolver (Solvers.c:1054-1454).
//
// KillerSolver solves J dx = F with Powell-Hestenes (augmented Lagrangian)
// iterations around GCR(6), preconditioned by a Cholesky factor of the
// penalised velocity block A - B*Dcm*C with Dcm = penalty/dV.  Here:
//   * the penalty is viscosity weighted, Dcm = gamma*eta_n/dV, so the
//     augmented block A_g = A - B*Dcm*C = A + gamma*C'diag(eta/dV)C stays well
//     scaled across a 1e6 viscosity contrast;
//   * the Powell-Hestenes pressure update and the velocity solve are fused
//     into one flexible GCR on the augmented system [A_g B; C 0] with the
//     block preconditioner  z_p = -(1+gamma) (eta_n/dV) r_p,
//                           z_u = A_g^-1 (r_u - B z_p),
//     the inverse Schur complement of the augmented system being
//     S^-1 - Dcm ~ -(1+gamma) eta/dV;
//   * A_g^-1 is approximated by conjugate gradients preconditioned with a
//     staggered geometric multigrid V-cycle (Chebyshev-Jacobi smoothing,
//     coarse grids with ceil(n/2) cells and viscosity averaged geometrically,
//     linear interpolation between non-nested grids, dense coarsest solve).
// The outer GCR, the operator and the residual are FP64; the inner CG and
// multigrid run in FP64 ("plain") or FP32 ("mixed", the FP64 outer
// iteration recovering the accuracy).
#include <cmath>
#include <memory>

#include "stokes_internal.h"

#define RB 256   // reduction block size

// ------------------------------------------------------------ small kernels
// Deterministic two-stage dot products: block partials, then one block sums
// them in a fixed order into dst (device scalar).
template <typename T>
__global__ void __launch_bounds__(RB) k_dot_part(int n, const T *__restrict__ a, const T *__restrict__ b,
                                                 double *__restrict__ part) {
    __shared__ double s[RB];
    double acc = 0.0;
    for (int i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) acc += (double)a[i] * (double)b[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void __launch_bounds__(RB) k_sum_part(int nb, const double *__restrict__ part, double *__restrict__ dst) {
    __shared__ double s[RB];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *dst = s[0];
}

template <typename T>
__global__ void k_fill(int n, T *__restrict__ a, T v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
template <typename A, typename B>
__global__ void k_convert(int n, const A *__restrict__ a, B *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (B)a[i];
}
// y = a*x + b*y with scalars taken from device memory when given
template <typename T>
__global__ void k_axpby(int n, T a, const T *__restrict__ x, T b, T *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// --------------------------------------------------- multigrid level kernels
// r = b - A_g x   (or y = A_g x when b == nullptr, negated); velocity only.
template <typename T>
__global__ void __launch_bounds__(256) k_vel_residual(sgrid g, T gamma, const T *__restrict__ x,
                                                      const T *__restrict__ b, const T *__restrict__ en,
                                                      const T *__restrict__ es, T *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g);
    const T *u = x, *v = x + nu;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        r[i] = vx_free(g, k, l) ? (b ? b[i] : T(0)) - apply_vx_row<T>(g, k, l, u, v, en, es, gamma, nullptr) : T(0);
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        r[i] = vz_free(g, k, l) ? (b ? b[i] : T(0)) - apply_vz_row<T>(g, k, l, u, v, en, es, gamma, nullptr) : T(0);
    }
}

// y = A_g p and partial sums of p.y (CG)
template <typename T>
__global__ void __launch_bounds__(RB) k_vel_apply_dot(sgrid g, T gamma, const T *__restrict__ p,
                                                      const T *__restrict__ en, const T *__restrict__ es,
                                                      T *__restrict__ y, double *__restrict__ part) {
    __shared__ double s[RB];
    const int i = blockIdx.x * RB + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g);
    const T *u = p, *v = p + nu;
    double acc = 0.0;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        const T q = vx_free(g, k, l) ? apply_vx_row<T>(g, k, l, u, v, en, es, gamma, nullptr) : T(0);
        y[i] = q;
        acc = (double)q * (double)p[i];
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        const T q = vz_free(g, k, l) ? apply_vz_row<T>(g, k, l, u, v, en, es, gamma, nullptr) : T(0);
        y[i] = q;
        acc = (double)q * (double)p[i];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

template <typename T>
__global__ void k_vel_diag_inv(sgrid g, T gamma, const T *__restrict__ en, const T *__restrict__ es,
                               T *__restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g);
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        dinv[i] = vx_free(g, k, l) ? T(1) / diag_vx<T>(g, k, l, en, es, gamma) : T(0);
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        dinv[i] = vz_free(g, k, l) ? T(1) / diag_vz<T>(g, k, l, en, es, gamma) : T(0);
    }
}

// One Chebyshev step:  r -= A_g d  (in place)
template <typename T>
__global__ void __launch_bounds__(256) k_cheb_res(sgrid g, T gamma, const T *__restrict__ d,
                                                  const T *__restrict__ en, const T *__restrict__ es,
                                                  T *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g);
    const T *u = d, *v = d + nu;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        if (vx_free(g, k, l)) r[i] -= apply_vx_row<T>(g, k, l, u, v, en, es, gamma, nullptr);
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        if (vz_free(g, k, l)) r[i] -= apply_vz_row<T>(g, k, l, u, v, en, es, gamma, nullptr);
    }
}
// x += d ; d = a*d + b*Dinv*r
template <typename T>
__global__ void k_cheb_upd(int n, T a, T b, const T *__restrict__ dinv, const T *__restrict__ r,
                           T *__restrict__ d, T *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T di = d[i];
    x[i] += di;
    d[i] = a * di + b * dinv[i] * r[i];
}
// d = s*Dinv*r (start of the smoother)
template <typename T>
__global__ void k_scale_dinv(int n, T s, const T *__restrict__ dinv, const T *__restrict__ r,
                             T *__restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s * dinv[i] * r[i];
}

// Linear interpolation weight of fine coordinate index f (position
// (f + off_f) * hf) onto coarse index K (position (K + off_c) * hc), with the
// coarse bracket clamped to [0, nc - 2].  Shared by P and R = P^T.
struct axis_map {
    double hf, hc, off_f, off_c;
    int nc;   // number of coarse coordinates
};
__host__ __device__ inline void bracket(const axis_map &m, int f, int *K0, double *w1) {
    double t = ((f + m.off_f) * m.hf) / m.hc - m.off_c;
    int k0 = (int)floor(t);
    if (k0 < 0) k0 = 0;
    if (k0 > m.nc - 2) k0 = m.nc - 2;
    double w = t - k0;
    if (w < 0) w = 0;
    if (w > 1) w = 1;
    *K0 = k0;
    *w1 = w;
}
__host__ __device__ inline double pweight(const axis_map &m, int f, int K) {
    int k0;
    double w1;
    bracket(m, f, &k0, &w1);
    if (K == k0) return 1.0 - w1;
    if (K == k0 + 1) return w1;
    return 0.0;
}

struct comp_maps {
    axis_map xm, zm;       // for one velocity component
    int nxf, nzf, nxc, nzc;
    int comp;              // 0: Vx (ghost rows in z), 1: Vz (ghost columns in x)
    int ghost_lo, ghost_hi;   // 1 free slip (+), 2 no slip (-)
};

// Coarse value at (K, L) with ghosts mapped to their interior neighbours and
// walls set to zero (the correction vanishes on Dirichlet nodes).
template <typename T>
__device__ inline T coarse_val(const comp_maps &m, const sgrid &gc, const T *__restrict__ xc, int K, int L) {
    if (m.comp == 0) {
        if (K <= 0 || K >= m.nxc - 1) return T(0);
        if (L == 0) return (m.ghost_lo == 1 ? T(1) : T(-1)) * xc[K + 1 * m.nxc];
        if (L == m.nzc - 1) return (m.ghost_hi == 1 ? T(1) : T(-1)) * xc[K + (m.nzc - 2) * m.nxc];
        return xc[K + L * m.nxc];
    } else {
        if (L <= 0 || L >= m.nzc - 1) return T(0);
        if (K == 0) return (m.ghost_lo == 1 ? T(1) : T(-1)) * xc[1 + L * m.nxc];
        if (K == m.nxc - 1) return (m.ghost_hi == 1 ? T(1) : T(-1)) * xc[m.nxc - 2 + L * m.nxc];
        return xc[K + L * m.nxc];
    }
}

// x_f += P x_c for both components
template <typename T>
__global__ void k_prolong(sgrid gf, sgrid gc, comp_maps mu, comp_maps mv, const T *__restrict__ xc,
                          T *__restrict__ xf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nuf = nvx(gf), nvf = nvz(gf), nuc = nvx(gc);
    const comp_maps *m;
    int k, l;
    const T *src;
    if (i < nuf) { k = i % gf.Nx; l = i / gf.Nx; if (!vx_free(gf, k, l)) return; m = &mu; src = xc; }
    else if (i < nuf + nvf) { const int j = i - nuf; k = j % (gf.Nx + 1); l = j / (gf.Nx + 1); if (!vz_free(gf, k, l)) return; m = &mv; src = xc + nuc; }
    else return;
    int K0, L0;
    double wx, wz;
    bracket(m->xm, k, &K0, &wx);
    bracket(m->zm, l, &L0, &wz);
    const T a = coarse_val<T>(*m, gc, src, K0, L0), b = coarse_val<T>(*m, gc, src, K0 + 1, L0);
    const T c = coarse_val<T>(*m, gc, src, K0, L0 + 1), d = coarse_val<T>(*m, gc, src, K0 + 1, L0 + 1);
    xf[i] += T((1 - wx) * (1 - wz)) * a + T(wx * (1 - wz)) * b + T((1 - wx) * wz) * c + T(wx * wz) * d;
}

// Sum of P weights times the fine residual onto coarse node (K, L) itself
// (no ghost folding).
template <typename T>
__device__ inline double restrict_node(const comp_maps &m, const sgrid &gf, const T *__restrict__ rf, int K, int L) {
    // fine index ranges whose bracket can include K / L
    const double tx0 = ((K - 1 + m.xm.off_c) * m.xm.hc) / m.xm.hf - m.xm.off_f;
    const double tx1 = ((K + 1 + m.xm.off_c) * m.xm.hc) / m.xm.hf - m.xm.off_f;
    const double tz0 = ((L - 1 + m.zm.off_c) * m.zm.hc) / m.zm.hf - m.zm.off_f;
    const double tz1 = ((L + 1 + m.zm.off_c) * m.zm.hc) / m.zm.hf - m.zm.off_f;
    int k0 = (int)floor(tx0) - 1, k1 = (int)ceil(tx1) + 1, l0 = (int)floor(tz0) - 1, l1 = (int)ceil(tz1) + 1;
    k0 = k0 < 0 ? 0 : k0; l0 = l0 < 0 ? 0 : l0;
    k1 = k1 > m.nxf - 1 ? m.nxf - 1 : k1; l1 = l1 > m.nzf - 1 ? m.nzf - 1 : l1;
    double s = 0.0;
    for (int l = l0; l <= l1; l++) {
        const double wz = pweight(m.zm, l, L);
        if (wz == 0.0) continue;
        for (int k = k0; k <= k1; k++) {
            const bool fr = (m.comp == 0) ? vx_free(gf, k, l) : vz_free(gf, k, l);
            if (!fr) continue;
            const double wx = pweight(m.xm, k, K);
            if (wx != 0.0) s += wx * wz * (double)rf[k + l * m.nxf];
        }
    }
    return s;
}

// b_c = P^T r_f, with ghost contributions folded back like P maps them.
template <typename T>
__global__ void k_restrict(sgrid gf, sgrid gc, comp_maps mu, comp_maps mv, const T *__restrict__ rf,
                           T *__restrict__ bc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nuc = nvx(gc), nvc = nvz(gc), nuf = nvx(gf);
    if (i < nuc) {
        const int K = i % gc.Nx, L = i / gc.Nx;
        if (!vx_free(gc, K, L)) { bc[i] = T(0); return; }
        double s = restrict_node<T>(mu, gf, rf, K, L);
        if (L == 1) s += (mu.ghost_lo == 1 ? 1.0 : -1.0) * restrict_node<T>(mu, gf, rf, K, 0);
        if (L == gc.Nz - 1) s += (mu.ghost_hi == 1 ? 1.0 : -1.0) * restrict_node<T>(mu, gf, rf, K, gc.Nz);
        bc[i] = (T)s;
    } else if (i < nuc + nvc) {
        const int j = i - nuc, K = j % (gc.Nx + 1), L = j / (gc.Nx + 1);
        if (!vz_free(gc, K, L)) { bc[i] = T(0); return; }
        const T *r = rf + nuf;
        double s = restrict_node<T>(mv, gf, r, K, L);
        if (K == 1) s += (mv.ghost_lo == 1 ? 1.0 : -1.0) * restrict_node<T>(mv, gf, r, 0, L);
        if (K == gc.Nx - 1) s += (mv.ghost_hi == 1 ? 1.0 : -1.0) * restrict_node<T>(mv, gf, r, gc.Nx, L);
        bc[i] = (T)s;
    }
}

// Coarse viscosity: geometric mean of the fine values whose positions fall
// in the coarse control volume (cells for eta_n, dual cells for eta_s).
template <typename T>
__global__ void k_coarsen_eta(sgrid gf, sgrid gc, const T *__restrict__ enf, const T *__restrict__ esf,
                              T *__restrict__ enc, T *__restrict__ esc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncc = ncell(gc), nvc = gc.Nx * gc.Nz;
    const double hx = gf.dx, hz = gf.dz, Hx = gc.dx, Hz = gc.dz;
    if (i < ncc) {
        const int K = i % (gc.Nx - 1), L = i / (gc.Nx - 1);
        // fine centres (k+0.5)hx in [K Hx, (K+1) Hx)
        int k0 = (int)ceil(K * Hx / hx - 0.5), k1 = (int)ceil((K + 1) * Hx / hx - 0.5) - 1;
        int l0 = (int)ceil(L * Hz / hz - 0.5), l1 = (int)ceil((L + 1) * Hz / hz - 0.5) - 1;
        k0 = min(max(k0, 0), gf.Nx - 2); l0 = min(max(l0, 0), gf.Nz - 2);
        k1 = min(max(k1, k0), gf.Nx - 2); l1 = min(max(l1, l0), gf.Nz - 2);
        double s = 0.0;
        int n = 0;
        for (int l = l0; l <= l1; l++)
            for (int k = k0; k <= k1; k++) { s += log((double)enf[k + l * (gf.Nx - 1)]); n++; }
        enc[i] = (T)exp(s / n);
    } else if (i < ncc + nvc) {
        const int j = i - ncc, K = j % gc.Nx, L = j / gc.Nx;
        int k0 = (int)ceil((K - 0.5) * Hx / hx), k1 = (int)ceil((K + 0.5) * Hx / hx) - 1;
        int l0 = (int)ceil((L - 0.5) * Hz / hz), l1 = (int)ceil((L + 0.5) * Hz / hz) - 1;
        k0 = min(max(k0, 0), gf.Nx - 1); l0 = min(max(l0, 0), gf.Nz - 1);
        k1 = min(max(k1, k0), gf.Nx - 1); l1 = min(max(l1, l0), gf.Nz - 1);
        double s = 0.0;
        int n = 0;
        for (int l = l0; l <= l1; l++)
            for (int k = k0; k <= k1; k++) { s += log((double)esf[k + l * gf.Nx]); n++; }
        esc[j] = (T)exp(s / n);
    }
}

// Dense coarsest solve: x[idx[i]] = sum_j Ainv[i][j] b[idx[j]]
template <typename T>
__global__ void __launch_bounds__(256) k_dense_solve(int n, const int *__restrict__ idx, const T *__restrict__ Ainv,
                                                     const T *__restrict__ b, T *__restrict__ x) {
    __shared__ T sb[1024];
    for (int j = threadIdx.x; j < n; j += blockDim.x) sb[j] = b[idx[j]];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += (double)Ainv[(size_t)i * n + j] * (double)sb[j];
        x[idx[i]] = (T)s;
    }
}

// ------------------------------------------------------- outer-level kernels
// y = K_g x for the augmented system [A_g B; C 0] (FP64, grid layout)
__global__ void __launch_bounds__(256) k_apply_Kg(sgrid g, double gamma, const double *__restrict__ x,
                                                  const double *__restrict__ en, const double *__restrict__ es,
                                                  double *__restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g), np = ncell(g);
    const double *u = x, *v = x + nu, *p = x + nu + nv;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        y[i] = vx_free(g, k, l) ? apply_vx_row<double>(g, k, l, u, v, en, es, gamma, nullptr) + grad_vx_row<double>(g, k, l, p) : 0.0;
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        y[i] = vz_free(g, k, l) ? apply_vz_row<double>(g, k, l, u, v, en, es, gamma, nullptr) + grad_vz_row<double>(g, k, l, p) : 0.0;
    } else if (i < nu + nv + np) {
        const int j = i - nu - nv, k = j % (g.Nx - 1), l = j / (g.Nx - 1);
        y[i] = div_cell<double>(g, k, l, u, v);
    }
}

// rhs_u = r_u - B z_p with z_p = -(1+gamma) eta_n/dV r_p; writes z_p (FP64)
// into zout and rhs_u (type T) into ru.
template <typename T>
__global__ void k_schur_rhs(sgrid g, double gamma, const double *__restrict__ r, const double *__restrict__ en,
                            double *__restrict__ zout, T *__restrict__ ru) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g), np = ncell(g), Ncx = g.Nx - 1;
    const double dV = g.dx * g.dz, s = -(1.0 + gamma) / dV;
    const double *rp = r + nu + nv;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        double v = 0.0;
        if (vx_free(g, k, l)) {
            const double zE = s * en[k + (l - 1) * Ncx] * rp[k + (l - 1) * Ncx];
            const double zW = s * en[k - 1 + (l - 1) * Ncx] * rp[k - 1 + (l - 1) * Ncx];
            v = r[i] - (zE - zW) * g.dz;
        }
        ru[i] = (T)v;
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        double v = 0.0;
        if (vz_free(g, k, l)) {
            const double zN = s * en[k - 1 + l * Ncx] * rp[k - 1 + l * Ncx];
            const double zS = s * en[k - 1 + (l - 1) * Ncx] * rp[k - 1 + (l - 1) * Ncx];
            v = r[i] - (zN - zS) * g.dx;
        }
        ru[i] = (T)v;
    } else if (i < nu + nv + np) {
        const int j = i - nu - nv;
        zout[i] = s * en[j] * rp[j];
    }
}

// true residual of the original system from the augmented one:
// r_u = r_g,u + B W r_p with W = gamma*eta_n/dV (grad of W r_p), r_p unchanged.
__global__ void k_true_residual(sgrid g, double gamma, const double *__restrict__ rg, const double *__restrict__ en,
                                double *__restrict__ rt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g), np = ncell(g), Ncx = g.Nx - 1;
    const double w = gamma / (g.dx * g.dz);
    const double *rp = rg + nu + nv;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        double v = 0.0;
        if (vx_free(g, k, l)) v = rg[i] + w * (en[k + (l - 1) * Ncx] * rp[k + (l - 1) * Ncx] - en[k - 1 + (l - 1) * Ncx] * rp[k - 1 + (l - 1) * Ncx]) * g.dz;
        rt[i] = v;
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        double v = 0.0;
        if (vz_free(g, k, l)) v = rg[i] + w * (en[k - 1 + l * Ncx] * rp[k - 1 + l * Ncx] - en[k - 1 + (l - 1) * Ncx] * rp[k - 1 + (l - 1) * Ncx]) * g.dx;
        rt[i] = v;
    } else if (i < nu + nv + np) {
        rt[i] = rg[i];
    }
}

// b_g,u = b_u - B W b_p  (augmented right-hand side)
__global__ void k_aug_rhs(sgrid g, double gamma, const double *__restrict__ b, const double *__restrict__ en,
                          double *__restrict__ bg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g), np = ncell(g), Ncx = g.Nx - 1;
    const double w = gamma / (g.dx * g.dz);
    const double *bp = b + nu + nv;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        double v = 0.0;
        if (vx_free(g, k, l)) v = b[i] - w * (en[k + (l - 1) * Ncx] * bp[k + (l - 1) * Ncx] - en[k - 1 + (l - 1) * Ncx] * bp[k - 1 + (l - 1) * Ncx]) * g.dz;
        bg[i] = v;
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        double v = 0.0;
        if (vz_free(g, k, l)) v = b[i] - w * (en[k - 1 + l * Ncx] * bp[k - 1 + l * Ncx] - en[k - 1 + (l - 1) * Ncx] * bp[k - 1 + (l - 1) * Ncx]) * g.dx;
        bg[i] = v;
    } else if (i < nu + nv + np) {
        bg[i] = b[i];
    }
}

// ----------------------------------------------------------- host structures
template <typename T>
struct mg_level {
    sgrid g;
    int nu = 0, nv = 0, n = 0;
    dbuf<T> en, es, dinv, x, b, r, d;
    double lmax = 1.0;
    comp_maps mu, mv;      // transfer to the next coarser level
    // coarsest level
    int nfree = 0;
    dbuf<int> idx;
    dbuf<T> Ainv;
    void release() {
        en.release(); es.release(); dinv.release(); x.release(); b.release(); r.release();
        d.release(); idx.release(); Ainv.release();
    }
};

template <typename T>
struct mg_hier {
    std::vector<mg_level<T>> L;
    dbuf<T> cg_x, cg_r, cg_z, cg_p, cg_q;   // inner CG vectors (fine velocity size)
    void release() {
        for (auto &l : L) l.release();
        L.clear();
        cg_x.release(); cg_r.release(); cg_z.release(); cg_p.release(); cg_q.release();
    }
};

struct solver_state {
    mg_hier<double> hd;
    mg_hier<float> hf;
    dbuf<double> X, R, Rt, B, Zp, V, Z;   // outer FGCR (V, Z: restart slots)
    dbuf<double> part, scal;
    int m = 0;
    size_t nfull = 0;
};

void mdg_solver_release(mdg_ctx *c) {
    if (!c || !c->solver) return;
    solver_state *s = c->solver;
    s->hd.release(); s->hf.release();
    s->X.release(); s->R.release(); s->Rt.release(); s->B.release(); s->Zp.release();
    s->V.release(); s->Z.release(); s->part.release(); s->scal.release();
    delete s;
    c->solver = nullptr;
}

static int nblk(int n) { return mdg_grid1d(n, 256); }

// dot product into a host double (synchronising)
template <typename T>
static int dot_host(mdg_ctx *c, solver_state *s, int n, const T *a, const T *b, double *out) {
    const int nb = std::min(1024, mdg_grid1d(n, RB));
    hipLaunchKernelGGL(k_dot_part<T>, dim3(nb), dim3(RB), 0, c->stream, n, a, b, s->part.p);
    hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(RB), 0, c->stream, nb, s->part.p, s->scal.p);
    HIPCHK(c, hipMemcpyAsync(out, s->scal.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

static axis_map make_axis(double hf, double hc, double off, int nc) { return axis_map{hf, hc, off, off, nc}; }

// ------------------------------------------------------------------ setup
static void host_gauss_jordan_inverse(std::vector<double> &A, int n, bool *ok) {
    std::vector<double> I((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) I[(size_t)i * n + i] = 1.0;
    *ok = true;
    for (int c0 = 0; c0 < n; c0++) {
        int piv = c0;
        for (int r = c0 + 1; r < n; r++)
            if (fabs(A[(size_t)r * n + c0]) > fabs(A[(size_t)piv * n + c0])) piv = r;
        if (fabs(A[(size_t)piv * n + c0]) < 1e-300) { *ok = false; return; }
        if (piv != c0)
            for (int j = 0; j < n; j++) {
                std::swap(A[(size_t)piv * n + j], A[(size_t)c0 * n + j]);
                std::swap(I[(size_t)piv * n + j], I[(size_t)c0 * n + j]);
            }
        const double d = A[(size_t)c0 * n + c0];
        for (int j = 0; j < n; j++) { A[(size_t)c0 * n + j] /= d; I[(size_t)c0 * n + j] /= d; }
        for (int r = 0; r < n; r++) {
            if (r == c0) continue;
            const double f = A[(size_t)r * n + c0];
            if (f == 0.0) continue;
            for (int j = 0; j < n; j++) { A[(size_t)r * n + j] -= f * A[(size_t)c0 * n + j]; I[(size_t)r * n + j] -= f * I[(size_t)c0 * n + j]; }
        }
    }
    A.swap(I);
}

template <typename T>
static int build_hierarchy(mdg_ctx *c, solver_state *s, mg_hier<T> &H, const mdg_solver_opts *o) {
    H.release();
    const double gamma = o->penalty;
    sgrid g = c->sg;
    const double Lx = c->xmax - c->xmin, Lz = c->zmax - c->zmin;
    int ncx = g.Nx - 1, ncz = g.Nz - 1;
    const int cmin = o->coarse_cells > 1 ? o->coarse_cells : 8;
    for (;;) {
        H.L.emplace_back();
        mg_level<T> &lv = H.L.back();
        lv.g = g;
        lv.nu = nvx(g); lv.nv = nvz(g); lv.n = lv.nu + lv.nv;
        HIPCHK(c, lv.en.reserve(ncell(g)));
        HIPCHK(c, lv.es.reserve((size_t)g.Nx * g.Nz));
        HIPCHK(c, lv.dinv.reserve(lv.n));
        HIPCHK(c, lv.x.reserve(lv.n)); HIPCHK(c, lv.b.reserve(lv.n));
        HIPCHK(c, lv.r.reserve(lv.n)); HIPCHK(c, lv.d.reserve(lv.n));
        if (H.L.size() == 1) {
            hipLaunchKernelGGL((k_convert<double, T>), dim3(nblk(ncell(g))), dim3(256), 0, c->stream, ncell(g), c->eta_n.p, lv.en.p);
            hipLaunchKernelGGL((k_convert<double, T>), dim3(nblk(g.Nx * g.Nz)), dim3(256), 0, c->stream, g.Nx * g.Nz, c->eta_s.p, lv.es.p);
        } else {
            mg_level<T> &fl = H.L[H.L.size() - 2];
            hipLaunchKernelGGL(k_coarsen_eta<T>, dim3(nblk(ncell(g) + g.Nx * g.Nz)), dim3(256), 0, c->stream, fl.g, g,
                               fl.en.p, fl.es.p, lv.en.p, lv.es.p);
        }
        hipLaunchKernelGGL(k_vel_diag_inv<T>, dim3(nblk(lv.n)), dim3(256), 0, c->stream, g, (T)gamma, lv.en.p, lv.es.p, lv.dinv.p);
        HIPCHK(c, hipGetLastError());
        if (ncx <= cmin || ncz <= cmin) break;
        const int ncx2 = (ncx + 1) / 2, ncz2 = (ncz + 1) / 2;
        sgrid gc = {ncx2 + 1, ncz2 + 1, Lx / ncx2, Lz / ncz2, g.ghostS, g.ghostN, g.ghostW, g.ghostE};
        // Vx: x at vertices (index * h), z at centres ((index - 0.5) * h)
        lv.mu = comp_maps{make_axis(g.dx, gc.dx, 0.0, gc.Nx), make_axis(g.dz, gc.dz, -0.5, gc.Nz + 1),
                          g.Nx, g.Nz + 1, gc.Nx, gc.Nz + 1, 0, g.ghostS, g.ghostN};
        lv.mv = comp_maps{make_axis(g.dx, gc.dx, -0.5, gc.Nx + 1), make_axis(g.dz, gc.dz, 0.0, gc.Nz),
                          g.Nx + 1, g.Nz, gc.Nx + 1, gc.Nz, 1, g.ghostW, g.ghostE};
        g = gc;
        ncx = ncx2;
        ncz = ncz2;
    }
    // Chebyshev bounds: power iteration on Dinv*A_g per level
    for (size_t li = 0; li + 1 < H.L.size(); li++) {
        mg_level<T> &lv = H.L[li];
        // start vector: the Jacobi diagonal (zero on the non-free entries)
        HIPCHK(c, hipMemcpyAsync(lv.x.p, lv.dinv.p, lv.n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        double lam = 1.0;
        for (int it = 0; it < 12; it++) {
            hipLaunchKernelGGL(k_vel_residual<T>, dim3(nblk(lv.n)), dim3(256), 0, c->stream, lv.g, (T)gamma, lv.x.p,
                               (const T *)nullptr, lv.en.p, lv.es.p, lv.r.p);     // r = -A x
            hipLaunchKernelGGL(k_scale_dinv<T>, dim3(nblk(lv.n)), dim3(256), 0, c->stream, lv.n, T(-1), lv.dinv.p, lv.r.p, lv.d.p);
            double yy, xx;
            int rc;
            if ((rc = dot_host<T>(c, s, lv.n, lv.d.p, lv.d.p, &yy)) || (rc = dot_host<T>(c, s, lv.n, lv.x.p, lv.x.p, &xx))) return rc;
            lam = sqrt(yy / xx);
            const T sc = (T)(1.0 / sqrt(yy));
            hipLaunchKernelGGL(k_axpby<T>, dim3(nblk(lv.n)), dim3(256), 0, c->stream, lv.n, sc, lv.d.p, T(0), lv.x.p);
        }
        lv.lmax = 1.1 * lam;
    }
    // coarsest: dense inverse of A_g on its free unknowns, assembled by
    // applying the operator to unit vectors on the device
    mg_level<T> &cl = H.L.back();
    std::vector<int> idx;
    for (int i = 0; i < cl.nu; i++) if (vx_free(cl.g, i % cl.g.Nx, i / cl.g.Nx)) idx.push_back(i);
    for (int j = 0; j < cl.nv; j++) if (vz_free(cl.g, j % (cl.g.Nx + 1), j / (cl.g.Nx + 1))) idx.push_back(cl.nu + j);
    const int nf = (int)idx.size();
    if (nf > 1024) return mdg_fail(c, MDG_ERR_ARG, "coarsest level too large (%d unknowns)", nf);
    cl.nfree = nf;
    HIPCHK(c, cl.idx.reserve(nf));
    HIPCHK(c, hipMemcpyAsync(cl.idx.p, idx.data(), nf * 4, hipMemcpyHostToDevice, c->stream));
    std::vector<double> A((size_t)nf * nf);
    std::vector<T> col(cl.n), e(cl.n);
    for (int j = 0; j < nf; j++) {
        std::fill(e.begin(), e.end(), T(0));
        e[idx[j]] = T(1);
        HIPCHK(c, hipMemcpyAsync(cl.x.p, e.data(), cl.n * sizeof(T), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_vel_residual<T>, dim3(nblk(cl.n)), dim3(256), 0, c->stream, cl.g, (T)gamma, cl.x.p,
                           (const T *)nullptr, cl.en.p, cl.es.p, cl.r.p);
        HIPCHK(c, hipMemcpyAsync(col.data(), cl.r.p, cl.n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < nf; i++) A[(size_t)i * nf + j] = -(double)col[idx[i]];
    }
    bool ok;
    host_gauss_jordan_inverse(A, nf, &ok);
    if (!ok) return mdg_fail(c, MDG_ERR_NOT_SPD, "singular coarsest operator");
    std::vector<T> Ai(A.size());
    for (size_t i = 0; i < A.size(); i++) Ai[i] = (T)A[i];
    HIPCHK(c, cl.Ainv.reserve(Ai.size()));
    HIPCHK(c, hipMemcpyAsync(cl.Ainv.p, Ai.data(), Ai.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    const int n0 = H.L[0].n;
    HIPCHK(c, H.cg_x.reserve(n0)); HIPCHK(c, H.cg_r.reserve(n0)); HIPCHK(c, H.cg_z.reserve(n0));
    HIPCHK(c, H.cg_p.reserve(n0)); HIPCHK(c, H.cg_q.reserve(n0));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

// ----------------------------------------------------------------- V-cycle
template <typename T>
static void chebyshev(mdg_ctx *c, mg_level<T> &lv, double gamma, int deg, bool zero_start) {
    // x (+)= Chebyshev(deg) correction for A_g x = b, using lv.r, lv.d
    const double lmax = lv.lmax, lmin = lmax / 30.0;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    double rho = 1.0 / sigma;
    const int n = lv.n;
    if (zero_start) {
        hipMemcpyAsync(lv.r.p, lv.b.p, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream);
        hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, lv.x.p, T(0));
    } else {
        hipLaunchKernelGGL(k_vel_residual<T>, dim3(nblk(n)), dim3(256), 0, c->stream, lv.g, (T)gamma, lv.x.p,
                           lv.b.p, lv.en.p, lv.es.p, lv.r.p);
    }
    hipLaunchKernelGGL(k_scale_dinv<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, (T)(1.0 / theta), lv.dinv.p, lv.r.p, lv.d.p);
    for (int k = 0; k < deg - 1; k++) {
        hipLaunchKernelGGL(k_cheb_res<T>, dim3(nblk(n)), dim3(256), 0, c->stream, lv.g, (T)gamma, lv.d.p, lv.en.p, lv.es.p, lv.r.p);
        const double rho_new = 1.0 / (2.0 * sigma - rho);
        hipLaunchKernelGGL(k_cheb_upd<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, (T)(rho_new * rho),
                           (T)(2.0 * rho_new / delta), lv.dinv.p, lv.r.p, lv.d.p, lv.x.p);
        rho = rho_new;
    }
    hipLaunchKernelGGL(k_axpby<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, T(1), lv.d.p, T(1), lv.x.p);
}

// solves approximately A_g x = b on level i (b, x in lv.b, lv.x)
template <typename T>
static void vcycle(mdg_ctx *c, mg_hier<T> &H, size_t i, double gamma, int deg) {
    mg_level<T> &lv = H.L[i];
    if (i + 1 == H.L.size()) {
        hipLaunchKernelGGL(k_fill<T>, dim3(nblk(lv.n)), dim3(256), 0, c->stream, lv.n, lv.x.p, T(0));
        hipLaunchKernelGGL(k_dense_solve<T>, dim3(1), dim3(256), 0, c->stream, lv.nfree, lv.idx.p, lv.Ainv.p, lv.b.p, lv.x.p);
        return;
    }
    mg_level<T> &cl = H.L[i + 1];
    chebyshev(c, lv, gamma, deg, true);
    hipLaunchKernelGGL(k_vel_residual<T>, dim3(nblk(lv.n)), dim3(256), 0, c->stream, lv.g, (T)gamma, lv.x.p, lv.b.p,
                       lv.en.p, lv.es.p, lv.r.p);
    hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(cl.n)), dim3(256), 0, c->stream, lv.g, cl.g, lv.mu, lv.mv, lv.r.p, cl.b.p);
    vcycle(c, H, i + 1, gamma, deg);
    hipLaunchKernelGGL(k_prolong<T>, dim3(nblk(lv.n)), dim3(256), 0, c->stream, lv.g, cl.g, lv.mu, lv.mv, cl.x.p, lv.x.p);
    chebyshev(c, lv, gamma, deg, false);
}

// Preconditioned CG on A_g x = b (b in H.cg_r on entry, x -> H.cg_x).
template <typename T>
static int inner_pcg(mdg_ctx *c, solver_state *s, mg_hier<T> &H, const mdg_solver_opts *o, int *its) {
    mg_level<T> &l0 = H.L[0];
    const int n = l0.n;
    const double gamma = o->penalty;
    const int nb = mdg_grid1d(n, RB);
    T *x = H.cg_x.p, *r = H.cg_r.p, *z = H.cg_z.p, *p = H.cg_p.p, *q = H.cg_q.p;
    HIPCHK(c, s->part.reserve((size_t)nb + 1024));   // k_vel_apply_dot writes nb partials
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, x, T(0));
    double rr0, rz, rr;
    int rc;
    if ((rc = dot_host<T>(c, s, n, r, r, &rr0))) return rc;
    *its = 0;
    if (rr0 == 0.0) return MDG_OK;
    const double tol2 = o->inner_tol * o->inner_tol * rr0;
    for (int it = 0; it < o->inner_max_its; it++) {
        // z = V(r)
        HIPCHK(c, hipMemcpyAsync(l0.b.p, r, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        vcycle(c, H, 0, gamma, o->cheb_degree);
        HIPCHK(c, hipMemcpyAsync(z, l0.x.p, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        double rz_new;
        if ((rc = dot_host<T>(c, s, n, r, z, &rz_new))) return rc;
        if (it == 0) HIPCHK(c, hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        else hipLaunchKernelGGL(k_axpby<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, T(1), z, (T)(rz_new / rz), p);
        rz = rz_new;
        hipLaunchKernelGGL(k_vel_apply_dot<T>, dim3(nb), dim3(RB), 0, c->stream, l0.g, (T)gamma, p, l0.en.p, l0.es.p, q, s->part.p);
        hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(RB), 0, c->stream, nb, s->part.p, s->scal.p);
        double pq;
        HIPCHK(c, hipMemcpyAsync(&pq, s->scal.p, 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (!(pq > 0.0)) break;   // loss of positivity: keep the current iterate
        const double alpha = rz / pq;
        hipLaunchKernelGGL(k_axpby<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, (T)alpha, p, T(1), x);
        hipLaunchKernelGGL(k_axpby<T>, dim3(nblk(n)), dim3(256), 0, c->stream, n, (T)(-alpha), q, T(1), r);
        *its = it + 1;
        if ((rc = dot_host<T>(c, s, n, r, r, &rr))) return rc;
        if (rr <= tol2) break;
    }
    return MDG_OK;
}

// z = M r (block preconditioner, FP64 in/out)
template <typename T>
static int precondition(mdg_ctx *c, solver_state *s, mg_hier<T> &H, const mdg_solver_opts *o,
                        const double *r, double *z, int *inner) {
    const sgrid g = c->sg;
    const int nu = nvx(g), nv = nvz(g), nfull = nu + nv + ncell(g);
    hipLaunchKernelGGL(k_schur_rhs<T>, dim3(nblk(nfull)), dim3(256), 0, c->stream, g, o->penalty, r, c->eta_n.p, z, H.cg_r.p);
    int its, rc;
    if ((rc = inner_pcg(c, s, H, o, &its))) return rc;
    *inner += its;
    hipLaunchKernelGGL((k_convert<T, double>), dim3(nblk(nu + nv)), dim3(256), 0, c->stream, nu + nv, H.cg_x.p, z);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

// Flexible GCR on the augmented system; b (grid layout, FP64) -> x.
template <typename T>
static int fgcr_solve(mdg_ctx *c, solver_state *s, mg_hier<T> &H, mdg_solver_opts *o, const double *b, double *x) {
    const sgrid g = c->sg;
    const int n = nvx(g) + nvz(g) + ncell(g);
    const int m = o->restart > 0 ? o->restart : 20;
    HIPCHK(c, s->R.reserve(n)); HIPCHK(c, s->Rt.reserve(n)); HIPCHK(c, s->B.reserve(n));
    HIPCHK(c, s->V.reserve((size_t)m * n)); HIPCHK(c, s->Z.reserve((size_t)m * n));
    double *R = s->R.p, *Rt = s->Rt.p;
    hipLaunchKernelGGL(k_aug_rhs, dim3(nblk(n)), dim3(256), 0, c->stream, g, o->penalty, b, c->eta_n.p, R);
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(256), 0, c->stream, n, x, 0.0);
    double bb;
    int rc;
    if ((rc = dot_host<double>(c, s, n, b, b, &bb))) return rc;
    const double bnorm = sqrt(bb);
    o->its = 0; o->inner_its = 0; o->rel_res = 0.0;
    if (bnorm == 0.0) return MDG_OK;
    int slot = 0;
    for (int it = 0; it < o->max_its; it++) {
        double *v = s->V.p + (size_t)slot * n, *z = s->Z.p + (size_t)slot * n;
        if ((rc = precondition<T>(c, s, H, o, R, z, &o->inner_its))) return rc;
        hipLaunchKernelGGL(k_apply_Kg, dim3(nblk(n)), dim3(256), 0, c->stream, g, o->penalty, z, c->eta_n.p, c->eta_s.p, v);
        for (int j = 0; j < slot; j++) {
            double a;
            if ((rc = dot_host<double>(c, s, n, s->V.p + (size_t)j * n, v, &a))) return rc;
            hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, c->stream, n, -a, s->V.p + (size_t)j * n, 1.0, v);
            hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, c->stream, n, -a, s->Z.p + (size_t)j * n, 1.0, z);
        }
        double vv, vr;
        if ((rc = dot_host<double>(c, s, n, v, v, &vv))) return rc;
        if (!(vv > 0.0)) return mdg_fail(c, MDG_ERR_DIVERGED, "GCR breakdown");
        const double nv = sqrt(vv);
        hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, c->stream, n, 0.0, v, 1.0 / nv, v);
        hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, c->stream, n, 0.0, z, 1.0 / nv, z);
        if ((rc = dot_host<double>(c, s, n, v, R, &vr))) return rc;
        hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, c->stream, n, vr, z, 1.0, x);
        hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, c->stream, n, -vr, v, 1.0, R);
        // residual of the original (non-augmented) system
        hipLaunchKernelGGL(k_true_residual, dim3(nblk(n)), dim3(256), 0, c->stream, g, o->penalty, R, c->eta_n.p, Rt);
        double rr;
        if ((rc = dot_host<double>(c, s, n, Rt, Rt, &rr))) return rc;
        o->its = it + 1;
        o->rel_res = sqrt(rr) / bnorm;
        if (std::isnan(o->rel_res)) return mdg_fail(c, MDG_ERR_NAN, "NaN in the Stokes solve");
        if (o->rel_res <= o->rel_tol) return MDG_OK;
        slot = (slot + 1 == m) ? 0 : slot + 1;
        if (slot == 0) { /* restart: discard the stored directions */ }
    }
    return mdg_fail(c, MDG_ERR_DIVERGED, "no convergence in %d iterations (rel. residual %.3e)", o->max_its, o->rel_res);
}

static void default_opts(mdg_solver_opts *o) {
    if (o->penalty <= 0) o->penalty = 1.0;
    if (o->rel_tol <= 0) o->rel_tol = 1e-8;
    if (o->max_its <= 0) o->max_its = 500;
    if (o->restart <= 0) o->restart = 20;
    if (o->inner_tol <= 0) o->inner_tol = 1e-3;
    if (o->inner_max_its <= 0) o->inner_max_its = 60;
    if (o->cheb_degree <= 0) o->cheb_degree = 3;
    if (o->coarse_cells <= 0) o->coarse_cells = 8;
}

// Grid-layout solve: b (u|v|p) -> x (u|v|p), device arrays of nvx+nvz+ncell.
int mdg_solve_grid(mdg_ctx *c, mdg_solver_opts *o, const double *b, double *x) {
    if (!c->solver) c->solver = new solver_state();
    solver_state *s = c->solver;
    HIPCHK(c, s->part.reserve(4096));
    HIPCHK(c, s->scal.reserve(16));
    default_opts(o);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int rc;
    if (o->precision == MDG_PREC_MIXED) {
        if ((rc = build_hierarchy<float>(c, s, s->hf, o))) return rc;
        rc = fgcr_solve<float>(c, s, s->hf, o, b, x);
    } else {
        if ((rc = build_hierarchy<double>(c, s, s->hd, o))) return rc;
        rc = fgcr_solve<double>(c, s, s->hd, o, b, x);
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    o->time_ms = ms;
    o->levels = c->solver ? (int)(o->precision == MDG_PREC_MIXED ? s->hf.L.size() : s->hd.L.size()) : 0;
    o->status = rc;
    return rc;
}

extern "C" int mdg_killer_solver(mdg_ctx *c, mdg_solver_opts *o, const double *D11_n,
                                 const double *D22_n, const double *D33_s, const double *rhs_mom,
                                 const double *rhs_cont, double *sol) {
    if (!c || !o || !rhs_mom || !rhs_cont || !sol) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    int rc;
    if ((rc = mdg_stage_coeffs(c, D11_n, D22_n, D33_s)) || (rc = mdg_reserve_grid_vectors(c))) return rc;
    const sgrid g = c->sg;
    const size_t n = (size_t)nvx(g) + nvz(g) + ncell(g);
    // F (equation numbering) -> grid layout
    HIPCHK(c, c->eqbuf.reserve((size_t)c->neq));
    HIPCHK(c, hipMemcpyAsync(c->eqbuf.p, rhs_mom, (size_t)c->neq_mom * 8, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->eqbuf.p + c->neq_mom, rhs_cont, (size_t)c->neq_cont * 8, hipMemcpyDefault, c->stream));
    double *b = c->gx.p, *x = c->gy.p;
    (void)n;
    if ((rc = mdg_eqn_to_grid(c, c->eqbuf.p, b, b + nvx(g), b + nvx(g) + nvz(g), 1.0))) return rc;
    rc = mdg_solve_grid(c, o, b, x);
    if (rc && rc != MDG_ERR_DIVERGED) return rc;
    const int rc2 = mdg_grid_to_eqn(c, x, x + nvx(g), x + nvx(g) + nvz(g), c->eqbuf.p);
    if (rc2) return rc2;
    HIPCHK(c, hipMemcpyAsync(sol, c->eqbuf.p, (size_t)c->neq * 8, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return rc;
}
