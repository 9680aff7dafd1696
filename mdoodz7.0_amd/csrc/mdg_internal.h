// Internal state of an mdg_ctx and helpers shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mdgpu.h"
#include "stencil.h"

// Device buffer with capacity growth; freed with the context.
template <typename T>
struct dbuf {
    T     *p   = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, (n ? n : 1) * sizeof(T));
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// A dbuf for staging inside one call: released on every return path.
template <typename T>
struct scoped_dbuf : dbuf<T> {
    scoped_dbuf() = default;
    scoped_dbuf(const scoped_dbuf &) = delete;
    scoped_dbuf &operator=(const scoped_dbuf &) = delete;
    ~scoped_dbuf() { this->release(); }
};

struct mdg_ctx {
    // grid
    int    Nx = 0, Nz = 0, Ncx = 0, Ncz = 0;
    double dx = 0, dz = 0, xmin = 0, xmax = 0, zmin = 0, zmax = 0;
    std::vector<double> h_xg, h_zg, h_xc, h_zc, h_xvz, h_zvx;
    dbuf<double> xg, zg, xc, zc, xvz, zvx;
    // boundary tags / values
    dbuf<signed char> BCu_t, BCv_t, BCp_t, BCg_t;
    dbuf<double>      BCu_v, BCv_v;
    bool              have_bc = false;
    // markers, sorted by (cell, quarter cell)
    int               n_markers = 0;
    dbuf<double>      mx, mz;        // sorted coordinates
    dbuf<int>         mphase;        // sorted phase
    dbuf<int>         perm;          // sorted slot -> caller index
    dbuf<unsigned>    qoff;          // 4*ncells+1 quarter-cell offsets
    // per-marker P2G weights of the 1-cell stencil for the four node sets
    // (centres, vertices, Vx, Vz), computed on first use after mdg_markers_set;
    // state 0 not computed, 1 valid, 2 some marker's nearest node differs from
    // the one its sort key implies (then the kernels compute weights inline)
    dbuf<double>      mw[4];
    int               mw_state[4] = {0, 0, 0, 0};
    double            weights_ms = 0.0;
    dbuf<unsigned>    keys, keys2;   // sort scratch
    dbuf<int>         idx;           // sort scratch
    dbuf<unsigned char> sort_tmp;
    // P2G scratch
    dbuf<double>      field_tmp;     // per-marker field in sorted order
    dbuf<double>      node_tmp;      // device node arrays for host callers
    dbuf<double>      table_tmp;     // per-phase tables
    dbuf<signed char> ntype_tmp;
    dbuf<int>         flag_dev;      // stray / status flags
    // Stokes: grid descriptor for the stencils, DOF numbering, coefficients
    sgrid             sg = {};
    int               neq = 0, neq_mom = 0, neq_cont = 0, ndof_norm[2] = {0, 0};
    dbuf<int>         eqn_u, eqn_v, eqn_p;
    dbuf<double>      eta_n, eta_s;          // Picard coefficients (D11/2, D33)
    dbuf<double>      tmp_a, tmp_b, tmp_c, tmp_d;
    dbuf<double>      gx, gy, eqbuf;         // grid-layout work vectors
    dbuf<double>      part;                  // reduction partials
    struct solver_state *solver = nullptr;
    struct rheo_scratch *rheo = nullptr;
    struct pcnet_state *pcnet = nullptr;
    struct newton_state *newton = nullptr;
    struct advect_scratch *advect = nullptr;
    struct g2p_scratch *g2p = nullptr;
    struct thermal_scratch *thermal = nullptr;
    struct reseed_scratch *reseed = nullptr;
    double      *static_scratch = nullptr;   // this device's static scratch (mdg_stokes_residual_static)
    // runtime
    hipStream_t  stream = nullptr;
    hipEvent_t   ev0 = nullptr, ev1 = nullptr;
    double       last_ms = 0.0;
    std::string  err;
};

#define MDG_MAX_DEVICES 64
int mdg_fail(mdg_ctx *c, int code, const char *fmt, ...);
#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return mdg_fail((ctx), MDG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// True when p is a device (or managed) allocation visible to the GPU.
bool mdg_is_device_ptr(const void *p);

// Copy n elements of a possibly-host array into a device buffer (or return
// the device pointer itself when it already is one).
template <typename T>
int mdg_stage_in(mdg_ctx *c, const T *src, size_t n, dbuf<T> &tmp, const T **out) {
    if (!src) { *out = nullptr; return MDG_OK; }
    if (mdg_is_device_ptr(src)) { *out = src; return MDG_OK; }
    HIPCHK(c, tmp.reserve(n));
    HIPCHK(c, hipMemcpyAsync(tmp.p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    *out = tmp.p;
    return MDG_OK;
}

static inline int mdg_grid1d(long n, int block) { return (int)((n + block - 1) / block); }

// Node-set geometry for P2Mastah's `centroid` argument (ParticleRoutines.c:2190-2213).
struct node_set {
    int nx, nz;
    const double *X, *Z;     // device coordinate vectors
    double X0, Z0;           // first coordinates (host copies)
};
int mdg_node_set(mdg_ctx *c, int centroid, node_set *ns);

// Frees the solver's device state (defined with the solver).
void mdg_solver_release(mdg_ctx *c);
void mdg_rheo_release(mdg_ctx *c);
void mdg_pcnet_release(mdg_ctx *c);
void mdg_newton_release(mdg_ctx *c);
void mdg_advect_release(mdg_ctx *c);
void mdg_g2p_release(mdg_ctx *c);
void mdg_thermal_release(mdg_ctx *c);
void mdg_reseed_release(mdg_ctx *c);
// Newton operator staged by mdg_newton_solver: y = J_gamma x (grid layout).
bool mdg_newton_active(const mdg_ctx *c);
int mdg_newton_apply_grid(mdg_ctx *c, double gamma, const double *x, double *y);
int mdg_solve_grid(mdg_ctx *c, mdg_solver_opts *o, const double *b, double *x);
