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
