// Declarations shared by the Stokes operator, residual, rheology and solver sources.
#pragma once
#include <vector>

#include "mdg_internal.h"

int mdg_build_numbering(mdg_ctx *c, const std::vector<signed char> &tu,
                        const std::vector<signed char> &tv, const std::vector<signed char> &tp);
int mdg_eqn_to_grid(mdg_ctx *c, const double *x, double *u, double *v, double *p, double scale);
int mdg_grid_to_eqn(mdg_ctx *c, const double *u, const double *v, const double *p, double *x);
int mdg_apply_K(mdg_ctx *c, const double *u, const double *v, const double *p, const double *en,
                const double *es, double *yu, double *yv, double *yp);
int mdg_stage_coeffs(mdg_ctx *c, const double *D11, const double *D22, const double *D33);
int mdg_reserve_grid_vectors(mdg_ctx *c);
// Looks up this device's copy of the static scratch of the allocation-free
// entry points (hipGetSymbolAddress, once per device) for the context.
int mdg_static_scratch_init(mdg_ctx *c);
