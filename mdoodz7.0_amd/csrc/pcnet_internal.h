// Solver-side hooks of the learned preconditioner (defined in pcnet.hip).
#pragma once
#include "mdg_internal.h"

// True once network parameters are loaded.
int mdg_pcnet_ready(mdg_ctx *c);
// Network initial guess x0 for A_gamma x = b on the fine velocity grid g:
// input channels (log10 eta_n / 3, D^-1 b / s on the W and S faces of each
// cell), output scaled back by s.  dinv = 1/diag(A_gamma).
template <typename V, typename T>
int mdg_pcnet_guess(mdg_ctx *c, const sgrid &g, const V *b, const T *dinv, const T *en, V *x0);
// Training-pair capture around a plain inner solve: the right-hand side
// before the solve, the solution after it.  No-ops unless capture is on.
// V is the type of the solver vectors (b, x), T that of the multigrid
// level arrays (dinv, en).
template <typename V, typename T>
int mdg_pcnet_capture_in(mdg_ctx *c, const sgrid &g, const V *b, const T *dinv, const T *en);
template <typename T>
int mdg_pcnet_capture_out(mdg_ctx *c, const sgrid &g, const T *x);

// Outer (FGCR) level.  From the augmented residual R (grid layout, FP64) the
// network gives two search directions: zu (velocity, zero pressure) and zp
// (pressure, zero velocity).  Input channels: log10(eta_n)/3, the
// Jacobi-scaled momentum residual on the W and S faces, the Schur-scaled
// continuity residual (1+gamma) eta_n/dV R_p, each part over its RMS.
template <typename T>
int mdg_pcnet_outer_dirs(mdg_ctx *c, const sgrid &g, double gamma, const double *R, const T *dinv,
                         const double *en, double *zu, double *zp);
// Outer training pairs: input from R_k with the iterate x_k kept at each
// iteration; once the solve has converged to x*, targets x* - x_k.
int mdg_pcnet_outer_capturing(mdg_ctx *c);
template <typename T>
int mdg_pcnet_outer_capture_in(mdg_ctx *c, const sgrid &g, double gamma, const double *R, const T *dinv,
                               const double *en, const double *x);
int mdg_pcnet_outer_capture_done(mdg_ctx *c, const sgrid &g, const double *xstar, int converged);

// y = K_gamma x for the augmented Picard system on the context's current
// viscosities (grid layout, FP64); defined with the solver.
int mdg_apply_Kg_grid(mdg_ctx *c, double gamma, const double *x, double *y);
