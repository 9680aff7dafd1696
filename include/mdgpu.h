/*
 * mdgpu.h — C ABI of the MI355X (gfx950) hot path for MDOODZ 7.0.
 *
 * One context owns all device buffers for one grid.  Every call is
 * synchronous on return (the context's stream is synchronised) and returns
 * MDG_OK or a negative MDG_ERR_* code; mdg_last_error() gives the message.
 * Array arguments may be host pointers or device pointers: the library
 * detects device allocations and then works in place without copies.
 *
 * Layouts are MDOODZ's own (row-major, x index fastest):
 *   vertices Nx*Nz, centres (Nx-1)*(Nz-1), Vx nodes Nx*(Nz+1),
 *   Vz nodes (Nx+1)*Nz;   Nx, Nz count vertices.
 * Boundary tags are MDOODZ's (Setup.c:206-230): -1 free, 0 and 11 Dirichlet,
 * 13 free slip, 2 stress, -2/-12 periodic, 30 air, 31 surface pressure.
 * References below are to files of the MDOODZ 7.0 source tree (MDLIB/).
 */
#ifndef MDGPU_H
#define MDGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define MDG_OK               0
#define MDG_ERR_ARG         -1   /* bad argument or size                    */
#define MDG_ERR_HIP         -2   /* HIP runtime / allocation failure        */
#define MDG_ERR_UNSUPPORTED -3   /* option outside the implemented subset   */
#define MDG_ERR_STATE       -4   /* call order (e.g. no markers set)        */
#define MDG_ERR_NOT_SPD     -5   /* non-positive diagonal (Solvers.c:1246)  */
#define MDG_ERR_DIVERGED    -6   /* solver did not reach its tolerance      */
#define MDG_ERR_NAN         -7   /* NaN residual (StokesRoutines.c:947-953) */
#define MDG_ERR_INCOMPRESSIBLE -8 /* continuity residual above KillerSolver's
                                    a-posteriori bound (Solvers.c:1333-1336)  */
#define MDG_ERR_NO_MARKERS  -9   /* re-seeding found no marker to copy (the
                                    reference exits, ParticleRoutines.c:3812) */

typedef struct mdg_ctx mdg_ctx;

/* Grid description: replaces the coordinate set-up of SetGridCoordinates
 * (GridRoutines.c:246-330); coordinates are built as the same running sums. */
typedef struct {
    int    Nx, Nz;                 /* vertices                               */
    double xmin, xmax, zmin, zmax;
    int    periodic_x;             /* must be 0 (periodic x not implemented) */
} mdg_grid_desc;

/* Context lifetime; device 0 of the calling process. */
int         mdg_create(const mdg_grid_desc *desc, mdg_ctx **out);
void        mdg_destroy(mdg_ctx *ctx);
const char *mdg_last_error(const mdg_ctx *ctx);

/* Device memory owned by the caller but allocated through the library's HIP
 * runtime, for callers that keep marker and grid arrays resident on the GPU
 * between calls (the reference has no counterpart: its arrays are host
 * arrays).  mdg_memcpy copies in any direction and is synchronous. */
int mdg_dev_alloc(mdg_ctx *ctx, unsigned long long bytes, void **ptr);
int mdg_dev_free(mdg_ctx *ctx, void *ptr);
int mdg_memcpy(mdg_ctx *ctx, void *dst, const void *src, unsigned long long bytes);
/* y = a*x on device arrays of n doubles (e.g. D11_n = 2*eta_n). */
int mdg_dev_scale(mdg_ctx *ctx, long long n, double a, const double *x, double *y);

/* Coordinates as computed by the library (host copies), lengths Nx, Nz,
 * Nx-1, Nz-1, Nx+1, Nz+1.  Mirrors mesh->xg_coord ... (GridRoutines.c:246-330). */
int mdg_grid_coords(const mdg_ctx *ctx, double *xg, double *zg, double *xc, double *zc,
                    double *xvz, double *zvx);

/* Boundary tags and values on the four node sets (mesh->BCu/BCv/BCp/BCg,
 * filled by SetBCs, Setup.c:206-~784).  BCu_v/BCv_v hold 2*value for tag 11
 * as SetBCs stores them. */
int mdg_set_bc(mdg_ctx *ctx, const signed char *BCu_t, const double *BCu_v,
               const signed char *BCv_t, const double *BCv_v, const signed char *BCp_t,
               const signed char *BCg_t);

/* ------------------------------------------------------------------ markers
 * Replaces the marker arrays handed to P2Mastah (the `markers` struct of
 * mdoodz-private.h:30-49: x, z, phase, Nb_part).  The library keeps a
 * device copy sorted by (cell, quarter cell); phase == -1 marks inactive
 * markers (ParticleRoutines.c:2292).  Capacity grows as needed.           */
int mdg_markers_set(mdg_ctx *ctx, int n, const double *x, const double *z, const int *phase);

/* Marker -> node index of ParticleRoutines.c:2300-2310 for node set
 * `centroid` (1 centres, 0 vertices, -1 Vx, -2 Vz), in the caller's marker
 * order.  Bit-exact with the reference expression ceil(d/dx + 0.5) - 1. */
int mdg_marker_index(mdg_ctx *ctx, int centroid, int *ip, int *jp);

/* Device copy of a per-marker field in the library's sorted order, for
 * repeated P2G calls on device-resident data (MDG_SRC_SORTED below).
 * src is in the caller's marker order; dst must be a device array of n. */
int mdg_markers_gather(mdg_ctx *ctx, const double *src, double *dst_sorted);

/* ---------------------------------------------------------------------- P2G
 * P2Mastah (ParticleRoutines.c:2170-2665, prototype mdoodz-private.h:350):
 *   flag 0: mat_prop is a per-phase table (mat_prop[phase]);
 *   flag 1: mat_prop is a per-marker array in the caller's marker order
 *           (or, with flag MDG_SRC_SORTED added, in the library's sorted
 *           order, device resident);
 *   avg 0/1/2: arithmetic / harmonic / geometric weighted mean;
 *   prop 1: phase fractions into phase_perc (nb_phases node arrays,
 *           centroid 0 or 1), node_field unused;
 *   centroid 1/0/-1/-2, interp_stencil 1 or 9.
 * Nodes with total weight < 1e-30 or tagged 30/31 in node_type keep their
 * previous value.  node_type may be NULL (no tag test).  Results agree with
 * the reference to round-off (summation order differs), not bit for bit. */
#define MDG_SRC_SORTED 0x100
int mdg_p2mastah(mdg_ctx *ctx, const double *mat_prop, int n_mat_prop, double *node_field,
                 const signed char *node_type, int flag, int avg, int prop, int centroid,
                 int interp_stencil, int nb_phases, double *const *phase_perc);

/* P2Mastah_Fused (ParticleRoutines.c:1895-2163, field descriptor
 * mdoodz-private.h:311-319): nfields (<= 16) fields, one marker pass. */
typedef struct {
    const double      *src;        /* table (flag 0) or per-marker (flag 1) */
    int                n_src;      /* entries in src                        */
    double            *dst;        /* node array                            */
    const signed char *node_type;  /* may be NULL                           */
    int                flag;       /* 0, 1, or 1|MDG_SRC_SORTED             */
    int                avg;        /* 0, 1, 2                               */
} mdg_p2g_field;
int mdg_p2mastah_fused(mdg_ctx *ctx, int centroid, int interp_stencil,
                       const mdg_p2g_field *fields, int nfields);

/* ------------------------------------------------------------ Stokes operator
 * Picard operator K = [A B; C 0] of the isotropic viscous Stokes equations
 * (StokesAssemblyDecoupled.c:62-212, 562-725, 1090-1145 with Assemble = 1),
 * applied matrix-free: y = K x with x, y in MDOODZ's equation numbering
 * (StokesRoutines.c:1162-1220; Vx rows, Vz rows, then cells).  D11_n, D22_n
 * (cells) and D33_s (vertices) are RheologicalOperators' Picard coefficients
 * (StokesRoutines.c:113-122,195-198: D11 = D22 = 2 eta_n, D33 = eta_s).  The
 * supported boundary layout is MDOODZ's closed box: walls tag 0 (normal
 * velocity), ghost rows/columns tag 13 (free slip) or 11 (no slip). */
int mdg_eqn_count(const mdg_ctx *ctx, int *neq, int *neq_mom, int *neq_cont);
int mdg_stokes_apply(mdg_ctx *ctx, const double *D11_n, const double *D22_n,
                     const double *D33_s, const double *x, double *y);

/* Ghost velocities from the tags (ApplyBC, StokesRoutines.c:415-500; prototype
 * mdoodz-private.h:540), in place
 * on the Vx (Nx*(Nz+1)) and Vz ((Nx+1)*Nz) arrays. */
int mdg_apply_bc(mdg_ctx *ctx, double *u, double *v);

/* ExtractSolutions2 (StokesRoutines.c:536-576, called by the defect step
 * :858 and the line search :650, :680; file-local, no prototype in
 * mdoodz-private.h): u_in, v_in, p_in += alpha * sol (equation numbering,
 * [momentum | continuity]) on every node that carries an unknown, then
 * ApplyBC.  The non-linear loop's defect step (alpha = -1) and its line-search
 * trials; host or device arrays. */
int mdg_extract_solution(mdg_ctx *ctx, const double *sol, double alpha, double *u, double *v, double *p);

/* StrainRateComponents (RheologyDensity.c:2001-2097, incompressible;
 * prototype mdoodz-private.h:426): div, exxd, ezzd on cells and exz on
 * vertices from u_in, v_in. */
int mdg_strain_rates(mdg_ctx *ctx, const double *u, const double *v, double *div,
                     double *exxd, double *ezzd, double *exz);

/* ExtractDiagonalScale (StokesRoutines.c:1260-1300; prototype
 * mdoodz-private.h:604): the scaling vector d
 * (equation numbering) = diagonal of the Picard block A on momentum rows and
 * 1 on continuity rows.  Residual norms with diag_scaling = 1 are taken on
 * d .* F (mdg_stokes_residual's d argument). */
int mdg_stokes_diag(mdg_ctx *ctx, const double *D11_n, const double *D22_n, const double *D33_s,
                    double *d);

/* InterpCentroidsToVerticesDouble (GridRoutines.c:108-212; prototype
 * mdoodz-private.h:294): vertex value =
 * 1/4 of the four surrounding centre values on the edge-copied padded array;
 * tag-30/31 centres contribute zero (divisor stays 4), tag-30 vertices get 0.
 * periodic_x selects the periodic column copy. */
int mdg_c2v(mdg_ctx *ctx, int periodic_x, const double *cent, double *vert);

/* The node-averaging tail of StrainRateComponents (RheologyDensity.c:2066-2094):
 * exz_n = 1/4 of the non-air vertex exz around each non-air cell, and exxd,
 * ezzd, div carried to the vertices by mdg_c2v.  Any output may be NULL. */
int mdg_strain_rates_nodes(mdg_ctx *ctx, const double *exxd, const double *ezzd,
                           const double *exz, const double *div, double *exz_n,
                           double *exxd_s, double *ezzd_s, double *div_s);

/* ---- rheology (NonNewtonianViscosityGrid) -------------------------------- */

/* One phase of mat_prop (mdoodz.h:158-170) for the implemented creep laws:
 * constant viscosity (cstv), dislocation power law (pwlv), diffusion creep
 * (linv), grain-boundary sliding (gbsv), regularised Peierls creep (expv),
 * Drucker-Prager plasticity with viscoplastic regularisation (plast = 1);
 * values in the model's scaled units, as ReadDataPowerLaw / ReadDataLinear /
 * ReadDataGBS / ReadDataExponential (FlowLaws.c:46-842) leave them.  beta is
 * the compressibility giving K = 1/beta.  A zero-filled tail (linv onward)
 * reproduces the results of the laws before it exactly: the grain-size,
 * fugacity and melt factors d^(m/n) f^(-r/n) exp(-a phi/n) are then 1. */
typedef struct {
    int    cstv, pwlv, plast;
    double eta0;
    double pref_pwl, F_pwl, A_pwl, n_pwl, Q_pwl, V_pwl;
    double C, fric, dil, Slim, eta_vp, n_vp;
    double beta;
    int    linv, gbsv, expv, reserved;
    double m_pwl, r_pwl, f_pwl, a_pwl;                         /* mpwl rpwl fpwl apwl */
    double F_lin, A_lin, n_lin, m_lin, r_lin, f_lin, a_lin, Q_lin, V_lin;
    double F_gbs, A_gbs, n_gbs, m_gbs, r_gbs, f_gbs, a_gbs, Q_gbs, V_gbs;
    double G_exp, S_exp, Q_exp, E_exp, n_exp;                  /* Gexp Sexp Qexp Eexp nexp */
    double gs_ref;                                             /* grain size of post-processing */
} mdg_phase_prop;

/* Fields of params / scale used by ViscosityConcise.  average is
 * model->eta_average: 0 arithmetic, 1 harmonic, 2 geometric.  T_peierls is
 * TmaxPeierls, (1200 + zeroC) / scaling->T (RheologyDensity.c:393): Peierls
 * creep acts below it.  d_min is 1e-13 m / scaling->L: grain-size dependent
 * diffusion creep refuses grain sizes below it (RheologyDensity.c:506-509). */
typedef struct {
    int    step, elastic, average;
    double dt, R, zeroC, min_eta, max_eta, eta_tol, Tiimin, scale_E, scale_eta;
    double T_peierls, d_min;
} mdg_rheo_params;

/* Node arrays of one node set (centres or vertices).  Inputs: perc is
 * nph x n (phase_perc_n / phase_perc_s, phase-major); G (mu), T, P (p_in or
 * P_s), the strain rates, the old stresses (NULL = 0), and the per-node
 * fric, C, dil (NULL = the phase's own value) and pref (a multiplier of
 * pref_pwl, NULL = 1; no counterpart in the reference), grain size d
 * (d0_n / d0_s) and melt fraction phi (phi0_n / phi0_s), NULL = 0.  Outputs:
 * eta (eta_n / eta_s, required), eta_phys, the stresses, p_corr, eII_pl
 * (each may be NULL). */
typedef struct {
    const double *perc, *G, *T, *P, *exx, *ezz, *exz, *sxx0, *szz0, *sxz0;
    const double *fric, *C, *dil, *pref;
    double *eta, *eta_phys, *sxx, *szz, *sxz, *p_corr, *eII_pl;
    const double *d, *phi;
} mdg_rheo_nodes;

#define MDG_MAX_PHASES 64

/* Per-phase shear modulus, compressibility and softening parameters
 * (mat_prop G, bet, C, C_end, phi, phi_end, psi, psi_end, pls_start,
 * pls_end, H_c, phi_soft, psi_soft, coh_soft). */
typedef struct {
    double G, beta, C, C_end, fric, fric_end, dil, dil_end, pls_start, pls_end, H_c;
    int    fric_soft, dil_soft, coh_soft;
} mdg_phase_soft;

/* ShearModCompExpGrid (RheologyDensity.c:1634-1730) and
 * CohesionFrictionDilationGrid (:1539-1616, Softening :1432-1535) on the
 * centres (centroid = 1) or vertices (0): mu and bet are harmonic phase
 * averages of G and beta; with strain_pl (the accumulated plastic strain on
 * the nodes, P2G of the markers' strain_pl) C, fric and dil are arithmetic
 * averages of the softened phase values (style 0 linear, 1 erfc, 2 linear
 * hardening).  Air nodes get zeros.  Outputs may be NULL. */
int mdg_material_grid(mdg_ctx *ctx, const mdg_phase_soft *phases, int nph, int centroid,
                      int style, const double *perc, const double *strain_pl, double *mu,
                      double *bet, double *C, double *fric, double *dil);

/* NonNewtonianViscosityGrid (RheologyDensity.c:1131-1426; prototype
 * mdoodz-private.h:423) on the centres
 * (centroid = 1, loop :1157-1316) or the vertices (centroid = 0, :1320-1407):
 * per node the elastic effective strain rate E + tau0/(2 G dt), per active
 * phase (fraction > 1e-13) ViscosityConcise (:385-1080), the phase average
 * of the effective and dissipative viscosities, and the stresses 2 eta E.
 * Air nodes (tags 30, and 31 on centres) get zeros.  Returns
 * MDG_ERR_DIVERGED where the reference exits on a failed local iteration
 * and MDG_ERR_ARG where it exits on a grain size below d_min in a phase with
 * grain-size dependent diffusion creep (m_lin > 0). */
int mdg_viscosity_grid(mdg_ctx *ctx, const mdg_phase_prop *phases, int nph,
                       const mdg_rheo_params *prm, int centroid, const mdg_rheo_nodes *f);

/* EvaluateRHS body force (StokesRoutines.c:961-1160; prototype
 * mdoodz-private.h:357): roger_x = gx * mean of
 * the two vertex densities, roger_z = gz * mean of the two cell densities. */
int mdg_eval_rhs(mdg_ctx *ctx, double gx, double gz, const double *rho_n, const double *rho_s,
                 double *roger_x, double *roger_z);

/* Residual of EvaluateStokesResidualDecoupled (StokesRoutines.c:867-955;
 * prototype mdoodz-private.h:497)
 * from the grid stresses: F (equation numbering) and the interior RMS norms
 * norms[0..2] = resx, resz, resp.  d (equation numbering) is the diagonal
 * scaling of ExtractDiagonalScale, or NULL for none. */
int mdg_stokes_residual(mdg_ctx *ctx, const double *sxxd, const double *szzd, const double *sxz,
                        const double *p_corr, const double *div, const double *roger_x,
                        const double *roger_z, const double *d, double *F, double norms[3]);

/* mdg_stokes_residual without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops): no counterpart in the reference.  Same arguments and results as
 * mdg_stokes_residual, but every array must be device memory (norms is a
 * host array of 3) and the reduction runs in a static device scratch of
 * MDG_STATIC_SCRATCH_DOUBLES doubles that each GPU holds until the process
 * ends.  All calls on one GPU share it: calls that overlap on one GPU race.
 * Grids needing more than the scratch (3 partial sums per 256 unknowns,
 * about 3800 x 3800 cells) are refused with MDG_ERR_ARG. */
#define MDG_STATIC_SCRATCH_DOUBLES (1 << 19)
int mdg_stokes_residual_static(mdg_ctx *ctx, const double *sxxd, const double *szzd,
                               const double *sxz, const double *p_corr, const double *div,
                               const double *roger_x, const double *roger_z, const double *d,
                               double *F, double norms[3]);

/* --------------------------------------------------------------- Stokes solve
 * Replaces KillerSolver (Solvers.c:1054-1454; prototype mdoodz-private.h:494;
 * called from SolveStokesDefectDecoupled, StokesRoutines.c:845-848) for the Picard
 * operator: solves K sol = [rhs_mom; rhs_cont] (equation numbering) to a
 * relative 2-norm residual rel_tol.  Outer flexible GCR (FP64) on the
 * augmented system with a viscosity-weighted penalty gamma*eta_n/dV
 * (Powell-Hestenes term of KillerSolver with Dcm = penalty/dV), inner CG on the
 * augmented velocity block (FP64) preconditioned by a staggered geometric
 * multigrid V-cycle computed in FP64 or FP32 (precision).  Zero fields take
 * defaults. */
#define MDG_PREC_FP64  0
#define MDG_PREC_MIXED 1
typedef struct {
    double penalty;        /* gamma of the viscosity-weighted penalty (10)     */
    double rel_tol;        /* ||r||/||r0|| target (1e-8)                       */
    int    max_its;        /* outer iterations (500)                           */
    int    restart;        /* GCR restart length (100)                         */
    double inner_tol;      /* relative tolerance of the velocity CG (1e-6)     */
    int    inner_max_its;  /* (80)                                             */
    int    cheb_degree;    /* Chebyshev smoothing degree on the finest level (8) */
    int    coarse_cells;   /* coarsen until <= this many cells per side (8)    */
    int    precision;      /* MDG_PREC_FP64 or MDG_PREC_MIXED                  */
    int    precond;        /* 0 multigrid; 1 multigrid CG started from the
                              learned network's guess; 2 two more outer
                              directions per iteration from the network
                              applied to the residual; 3 the network's
                              correction first, multigrid applied to what it
                              leaves (mdg_pcnet_*)                             */
    int    cycle;          /* 1 V-cycle (1; other values are refused: the
                              W-cycle stalled on the ellipse workloads)        */
    int    eta_avg;        /* coarse viscosity: 0 geometric, 1 arithmetic,
                              2 harmonic, 3 maximum (0)                        */
    int    check_every;    /* inner CG convergence test period (1)             */
    int    schur;          /* Schur complement approximation: 0 viscosity-
                              weighted mass, 1 scaled BFBt (needs precision =
                              MDG_PREC_FP64), 2 and 3 the mass weighted by
                              the largest viscosity over the 3x3 and 5x5 cells
                              around each cell (0)                              */
    double poisson_tol;    /* BFBt pressure-Poisson CG tolerance (1e-3)        */
    int    poisson_max_its;/* (30)                                             */
    int    smoother;       /* 0 fused three-term Chebyshev kernels, 1 the
                              residual/update two-kernel form, 2 the two-kernel
                              form preconditioned by vertex-patch additive
                              Schwarz (exact solves on the four faces around
                              each vertex, which hold the divergence-free
                              modes the penalty term leaves to the smoother)
                              instead of the diagonal, 3 the same patch
                              preconditioning fused into one kernel per step
                              with the patch systems solved in the kernel (0) */
    int    cheb_degree_coarse; /* smoothing degree on every level below the
                              finest (16)                                      */
    double cheb_range;     /* the Chebyshev smoother targets eigenvalues of
                              D^-1 A in [lmax / cheb_range, lmax] (30)         */
    int    coarsen4;       /* levels with at most this many cells per side
                              coarsen by 4 instead of 2 (0: never) (0)         */
    int    graphs;         /* 0 replay each inner CG iteration as a captured
                              HIP graph, 1 launch its kernels one by one (for
                              profilers that do not trace graph nodes), 2 as 1
                              with HIP events around every fine-level
                              smoother step (smoother_ms below) (0)           */
    /* outputs */
    int    its, inner_its, poisson_its, levels, status;
    double rel_res, time_ms;
    int    guess_used;     /* precond 1: inner solves started from the
                              network's guess; 2: network directions used      */
    double guess_gain;     /* precond 1: mean ||r0||^2/||b||^2 after the
                              guesses; 2: mean ||R||^2 after/before each
                              iteration's network directions; 3: mean
                              ||R - K z0||^2/||R||^2 of the network part       */
    /* KillerSolver's own iteration and tolerances (Solvers.c:1054-1454,
     * defaults InputOutput.c:1174-1198).  method 0 is the solve above;
     * method 1 runs KillerSolver's Powell-Hestenes loop with the constant
     * penalty gamma/dV (penalty field) around GCR(6) with the reference's
     * stopping rules, the Cholesky factor of Jts = A - B (gamma/dV) C being
     * replaced by a multigrid-preconditioned CG solve of Jts to the relative
     * tolerance pc_tol.                                                     */
    int    method;         /* 0 or 1 (0)                                       */
    int    max_its_PH;     /* method 1 (10)                                    */
    double rel_tol_KSP;    /* method 1 (1e-4)                                  */
    int    max_its_KSP;    /* method 1 (100)                                   */
    double lin_abs_mom, lin_rel_mom;   /* method 1 PH stop (1e-9, 1e-5)        */
    /* continuity tolerances: the PH stop of method 1 and, for both methods,
     * the a-posteriori check of Solvers.c:1333: the call returns
     * MDG_ERR_INCOMPRESSIBLE (the solution is still written) when
     * |max_div| > 10 lin_abs_div and max_div / max_div0 > lin_rel_div.
     * Method 0 takes both as 0 = no check unless set (method 1: 1e-9, 1e-5). */
    double lin_abs_div, lin_rel_div;
    double pc_tol;         /* method 1 (1e-10)                                 */
    /* outputs: signed maxima (MinMaxArrayVal, MiscFunctions.c:342-375) of the
     * momentum and continuity residuals of the returned solution, and of the
     * first Powell-Hestenes iterate (method 1) or of the right-hand side
     * (method 0, which starts from zero and has no Powell-Hestenes iterate) */
    double max_mom, max_mom0, max_div, max_div0;
    int    its_PH, its_KSP;/* method 1: PH iterations, GCR iterations summed
                              as kspgcr counts them (Solvers.c:208-211)       */
    double smoother_ms;    /* graphs = 2: mean duration of the fine-level
                              smoother steps (k_cheb3; k_pcheb with smoother
                              3) in this solve                                */
    int    smoother_launches;
    /* input: smoothers 2 and 3 use vertex patches on the finest patch_levels
     * levels and the diagonal below them (0: every level, the default; the
     * coarse levels need the patches too: 2 levels of patches take 3x the
     * V-cycles of all levels at 2047^2)                                      */
    int    patch_levels;
} mdg_solver_opts;
int mdg_killer_solver(mdg_ctx *ctx, mdg_solver_opts *opts, const double *D11_n,
                      const double *D22_n, const double *D33_s, const double *rhs_mom,
                      const double *rhs_cont, double *sol);

/* ---------------------------------------------------------------- advection
 * V2P (ParticleRoutines.c:63-256): the grid velocity at the markers. u on
 * the Vx nodes (Nx x (Nz+1), mesh->u_in), v on the Vz nodes ((Nx+1) x Nz,
 * mesh->v_in); conserv_interp = model.conserv_interp (Taras' correction);
 * inactive markers (phase -1) get 0. The reference's air-node (tag 30)
 * substitutions are kept, though mdg_set_bc does not accept free-surface
 * tags yet. Arrays may be host or device memory. */
int mdg_v2p(mdg_ctx *ctx, const double *u, const double *v, int conserv_interp, int n, const double *x,
            const double *z, const int *phase, double *vx, double *vz);

/* RogerGuntherII (AdvectionRoutines.c:347-410; prototype
 * mdoodz-private.h:593; called at Main_DOODZ.c:1304) with isoutPart
 * (AdvectionRoutines.c:320-341, non-periodic box): fourth-order Runge-Kutta
 * with V2P at each stage. x, z, phase are updated in place (markers leaving
 * the box get phase -1 and keep their exit position); vx, vz (may be NULL:
 * particles->Vx/Vz) receive the RK4 velocity of the markers still active
 * and keep their values otherwise. The sorted copy of mdg_markers_set is
 * not touched: call it again with the new positions before the next P2G. */
int mdg_advect_rk4(mdg_ctx *ctx, const double *u, const double *v, double dt, int conserv_interp, int n,
                   double *x, double *z, int *phase, double *vx, double *vz);

/* ---------------------------------------------------------------- grid -> markers
 * Interp_Grid2P_centroids2 (ParticleRoutines.c:1332-1580; non-periodic): a
 * centre field F ((Nx-1) x (Nz-1)) at the markers, bilinear between the
 * centres, copied along the box edges, BCp tag 30/31 cells replaced by their
 * vertical neighbour, normalised by the weight sum; inactive markers get 0. */
int mdg_g2p_centroids(mdg_ctx *ctx, const double *F, int n, const double *x, const double *z,
                      const int *phase, double *out);

/* Interp_Grid2P (ParticleRoutines.c:1223-1325) on the vertices (xg, zg,
 * BCg tags): a vertex field F (Nx x Nz) at the markers, bilinear in the
 * clamped cell, tag 30/31 vertices replaced by their vertical neighbour, not
 * normalised (as the reference); inactive markers get 0. */
int mdg_g2p_vertices(mdg_ctx *ctx, const double *F, int n, const double *x, const double *z,
                     const int *phase, double *out);

/* UpdateParticleStress (RheologyParticles.c:1725-2082) for elastic runs
 * with model.subgrid_diffusion 0 or 1 and model.stress_rotation = 0
 * (:1924-1960): sxxd - sxxd0, szzd - szzd0 on the cells not tagged 30/31 and
 * sxz - sxz0 on the vertices not tagged 30, interpolated as
 * mdg_g2p_centroids / mdg_g2p_vertices and added to the marker stresses
 * (particles->sxxd, szzd, sxz) in place. The stress rotation step
 * (stress_rotation 1, 2) is not part of this entry point. */
int mdg_update_particle_stress(mdg_ctx *ctx, const double *sxxd, const double *szzd, const double *sxz,
                               const double *sxxd0, const double *szzd0, const double *sxz0, int n,
                               const double *x, const double *z, const int *phase, double *msxxd,
                               double *mszzd, double *msxz);

/* UpdateParticleStress with model.subgrid_diffusion = 2 (RheologyParticles.c:
 * 1811-1895): the old grid stresses (sxxd0, szzd0 on the cells, sxz0 on the
 * vertices) and the vertex viscosity eta_phys_s at the markers; on active
 * markers the subgrid relaxation dtau_s = -(tau_m - tau0_m)
 * (1 - exp(-dt / (eta_m / G[phase]))); dtau_s onto the grid (P2Mastah,
 * flag 1, arithmetic, interp_stencil, BCp / BCg node tags); the remaining
 * grid increments (tau - tau0) - dtau_s on non-air nodes back at the
 * markers; tau_m += dtau_s + dtau_r in place.  G_phase holds mat_prop G of
 * the nph phases.  The markers must be the ones set by mdg_markers_set. */
int mdg_update_particle_stress_subgrid(mdg_ctx *ctx, double dt, const double *G_phase, int nph, int interp_stencil,
                                       const double *sxxd, const double *szzd, const double *sxz,
                                       const double *sxxd0, const double *szzd0, const double *sxz0,
                                       const double *eta_phys_s, int n, const double *x, const double *z,
                                       const int *phase, double *msxxd, double *mszzd, double *msxz);

/* UpdateParticlePressure (RheologyParticles.c:1621-1704; called after the
 * solve) with model.subgrid_diffusion < 2: dp = p_in - p0_n on the cells not
 * tagged 30 (mesh->dp, returned when dp is not NULL), interpolated as above
 * and added to the marker pressure P (particles->P) in place. */
int mdg_update_particle_pressure(mdg_ctx *ctx, const double *p_in, const double *p0_n, double *dp, int n,
                                 const double *x, const double *z, const int *phase, double *P);

/* UpdateParticlePressure with model.subgrid_diffusion >= 2
 * (RheologyParticles.c:1621-1695): dp = p_in - p0_n on non-air cells (dp may
 * be NULL); eta_phys_n and p0_n at the markers; dP_s = -(P_m - P0_m)
 * (1 - exp(-dt / (eta_m / beta[phase]))) on active markers (the reference's
 * relaxation time, kept as written); dP_s onto the centres (P2Mastah, flag
 * 1, 1-cell stencil as the reference hard-codes), the remaining dp - dP_s
 * back at the markers, P += dP_s + dP_r in place.  beta_phase holds mat_prop
 * bet of the nph phases.  The reference leaves BCt-30 cells out of the old
 * pressure; the BCp tags used here mark the same air cells. */
int mdg_update_particle_pressure_subgrid(mdg_ctx *ctx, double dt, const double *beta_phase, int nph,
                                         const double *p_in, const double *p0_n, const double *eta_phys_n,
                                         double *dp, int n, const double *x, const double *z, const int *phase,
                                         double *P);

/* AccumulatedStrainII (RheologyParticles.c:346-649; called at
 * Main_DOODZ.c:1251): increments dt * eII on the centres (the total being
 * the sum of the seven mechanisms), interpolated to each active marker from
 * the four centres around it (BCp 30/31 centres left out, weights
 * renormalised) and added to its accumulated strains.  eII holds 7 centre
 * arrays: eII_pl, eII_pwl, eII_exp, eII_lin, eII_gbs, eII_el, eII_cst (an
 * entry may be NULL = 0); div_u_pl may be NULL; strain holds 8 marker arrays
 * updated in place: strain, strain_el, strain_pl, strain_pl_vol,
 * strain_pwl, strain_exp, strain_lin, strain_gbs (an entry may be NULL =
 * skipped).  A negative total increment returns MDG_ERR_ARG (the reference
 * exits). */
int mdg_accumulated_strain(mdg_ctx *ctx, double dt, const double *const *eII, const double *div_u_pl, int n,
                           const double *x, const double *z, const int *phase, double *const *strain);

/* The stress rotation of UpdateParticleStress for elastic runs
 * (RheologyParticles.c:1965-1988): with the markers' spin wxzm, mode 1
 * (stress_rotation = 1) the Jaumann rate, mode 2 the rotation by the angle
 * dt * wxzm, mode 0 nothing; active markers only, in place.  The reference
 * fills wxzm with P2Mastah(wxzm, mesh->wxz, flag 1) at :1971, which
 * interpolates the zero-initialised marker array onto the vertex spin
 * instead of the spin onto the markers: there wxzm stays 0, the stresses
 * are unchanged and mesh->wxz is overwritten with zeros on every vertex
 * with markers.  The same sequence here is mdg_p2mastah of a zero marker
 * array onto the vertices (centroid 0, flag 1) followed by this call with
 * wxzm = 0. */
int mdg_rotate_particle_stress(mdg_ctx *ctx, int mode, double dt, int n, const int *phase, const double *wxzm,
                               double *sxxd, double *szzd, double *sxz);

/* The re-seeding of CountPartCell_OLD (ParticleRoutines.c:3224-4050,
 * reseed_mode = 0, reseed_markers = 1; prototype mdoodz-private.h; called at
 * Main_DOODZ.c:1401) for a box without free surface or periodic x: fine
 * cells of dx/2 x dz/2 (the box plus one coarse cell of margin) count the
 * active markers nearest to their centres; every interior fine cell of a
 * non-air coarse cell (cell_tag, the reference's BCt.type on the cells:
 * 30 = air; NULL = no air) holding fewer than floor(min_part_cell / 4) markers
 * gets one new marker at its centre that copies the phase of the closest
 * marker of the 5 x 5 fine cells around it (FindClosestPhase2, ties to the
 * reference's list order).  New markers, in x-major fine-cell order, take
 * the inactive slots (phase -1) in increasing order, then slots n, n+1, ...
 * x, z, phase hold capacity entries (n used on entry); new_slot and new_src
 * (at least 4 Ncx Ncz entries) receive each new marker's slot and source
 * marker, so the caller copies its other carried fields (AssignMarker-
 * Properties, ParticleRoutines.c:695-766) and interpolates the grid fields to
 * the new slots (mdg_g2p_centroids / mdg_g2p_vertices).  *n_out is the new
 * marker count.  The phase fractions the reference computes first are
 * mdg_p2mastah with prop = 1.  Errors: MDG_ERR_NO_MARKERS (empty
 * neighbourhood), MDG_ERR_ARG (capacity exceeded); nothing is written then. */
int mdg_count_part_cell(mdg_ctx *ctx, int min_part_cell, const signed char *cell_tag, int n, int capacity, double *x,
                        double *z, int *phase, int *new_slot, int *new_src, int *n_new, int *n_out);

/* ---------------------------------------------------------------- energy
 * One implicit energy step: the matrix of EnergyDirectSolve
 * (ThermalRoutines.c:74-470; prototype mdoodz-private.h, called from
 * Main_DOODZ.c's time loop) for a box without air cells (theta = 1,
 * transient), solved with SolveThermalPCG (ThermalSolver.c:338-480,
 * model.thermal_solver = 1, prototype mdoodz-private.h:471): Jacobi PCG from
 * x0 = T, stopping when ||r|| / ||r0|| < rel_tol.
 * Cells (Ncx*Ncz, x fastest): rho_n, Cp, T (in/out) and the optional source
 * terms Qr, Wdiss, Ha (NULL: absent); kx on the Vx nodes (Nx*(Nz+1)), kz on
 * the Vz nodes ((Nx+1)*Nz); BCT_t/BCT_v on the cells with a ghost ring
 * ((Ncx+2)*(Ncz+2)): type 1 Dirichlet at the wall face, 0 insulated (the
 * south ghost adds the reference's -val/dx flux term).  *its receives the
 * reference's return value: iterations (> 0), 0 (zero initial residual),
 * -1 (max_its reached, T holds the last iterate), -2 / -3 (breakdown); the
 * reference then falls back to its direct solver, which is the caller's
 * choice here. */
int mdg_energy_step(mdg_ctx *ctx, double dt, const double *rho_n, const double *Cp, const double *Qr,
                    const double *Wdiss, const double *Ha, const double *kx, const double *kz,
                    const signed char *BCT_t, const double *BCT_v, int max_its, double rel_tol, double *T, int *its);

/* ---------------------------------------------------------------- Newton
 * RheologicalOperators with Jacobian = 1 (StokesRoutines.c:250-375;
 * prototype mdoodz-private.h:607), isotropic and incompressible: the
 * tangent coefficients from central finite differences of the phase-averaged
 * effective viscosity (DerivativesOnTheFly_n/_s, FD_Jacobian.c:48-323).
 * D (nrow x n, row-major) receives D11..D14, D21..D24 on the centres
 * (centroid = 1, nrow = 8) or D31..D34 on the vertices (centroid = 0,
 * nrow = 4).  Inputs as for mdg_viscosity_grid (outputs unused). */
int mdg_tangent_grid(mdg_ctx *ctx, const mdg_phase_prop *phases, int nph,
                     const mdg_rheo_params *prm, int centroid, const mdg_rheo_nodes *f, double *D);

/* y = J x for the Newton Jacobian (BuildJacobianOperatorDecoupled,
 * StokesAssemblyDecoupled.c:1786-2304, applied matrix-free) in equation
 * numbering; Dn = 8 x cells, Ds = 4 x vertices from mdg_tangent_grid. */
int mdg_jacobian_apply(mdg_ctx *ctx, const double *Dn, const double *Ds, const double *x,
                       double *y);

/* KillerSolver on the Newton system J dx = F (KillerSolver with the
 * Jacobian blocks, Solvers.c:1054-1454): FGCR on J_gamma, preconditioned by
 * the Picard block preconditioner built from D11_n = 2 eta_n and D33_s =
 * eta_s (as the reference preconditions with the Picard factor). */
int mdg_newton_solver(mdg_ctx *ctx, mdg_solver_opts *opts, const double *Dn, const double *Ds,
                      const double *D11_n, const double *D33_s, const double *rhs_mom,
                      const double *rhs_cont, double *sol);

/* ------------------------------------------------ learned preconditioner
 * A 4-level U-Net of 3x3 convolutions (4 -> 32 -> 32 -> 64 -> 64 channels
 * down, 64 -> 32 -> 32 -> 3 up, 2x2 mean pooling, nearest upsampling with
 * additive skips, ReLU) computed with BF16 MFMA and FP32 accumulation.  No
 * counterpart in the reference: it is the learned part of the solve.
 * Images are H x W x 4 FP32 in NHWC order (H, W multiples of 8); outputs use
 * channels 0..2 of an H x W x 4 array.  Parameters are one FP32 vector:
 * per layer W[tap][ci][co] (tap = 3*(dy+1) + (dx+1)) then b[co]. */
int mdg_pcnet_param_count(mdg_ctx *ctx, long long *n);
int mdg_pcnet_set_params(mdg_ctx *ctx, const float *params, long long n);
/* params and/or grads (the last mdg_pcnet_grad result) may be NULL. */
int mdg_pcnet_get_params(mdg_ctx *ctx, float *params, float *grads);
int mdg_pcnet_forward(mdg_ctx *ctx, int H, int W, const float *x, float *y);
/* Loss 0.5 * sum(mask (y - t)^2) / sum(mask) over channels 0..2 and its
 * gradient (added to the previous one when accumulate != 0). */
int mdg_pcnet_grad(mdg_ctx *ctx, int H, int W, const float *x, const float *t, const float *mask,
                   int accumulate, double *loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the stored gradient. */
int mdg_pcnet_adam(mdg_ctx *ctx, float lr);
/* Training pairs generated by the solver itself: while max_pairs > 0,
 * every inner multigrid-CG solve of the velocity block A_gamma y = b in
 * mdg_killer_solver (precond = 0) stores the input image (log10(eta_n)/3 and
 * D^-1 b / s on each cell's W and S faces, s = RMS of D^-1 b) and the target
 * y / s with its mask, in a ring of max_pairs device slots (one grid size).
 * mdg_pcnet_capture(ctx, 0) stops capturing; a new call clears the ring. */
int mdg_pcnet_capture(mdg_ctx *ctx, int max_pairs);
int mdg_pcnet_captured(mdg_ctx *ctx, int *n, int *H, int *W);
int mdg_pcnet_get_pair(mdg_ctx *ctx, int i, float *x, float *t, float *mask);
/* As mdg_pcnet_capture, for the outer level: the next solves record, at
 * every outer iteration k, the network input built from the residual R_k and,
 * once the solve has converged to x*, the target x* - x_k (velocity and
 * pressure parts each over its RMS).  Pairs of a solve that does not
 * converge are dropped.  These pairs train the network for precond = 2. */
int mdg_pcnet_capture_outer(mdg_ctx *ctx, int max_pairs);
/* steps Adam updates on the captured pairs, batch pairs per step (chosen
 * by a counter hash of seed), learning rate lr; mean loss per step. */
int mdg_pcnet_train(mdg_ctx *ctx, int steps, int batch, float lr, unsigned long long seed,
                    double *losses);
/* Training for precond = 2 on outer pairs with the loss the solver sees:
 * L = min over a, b of ||R - a K z_u - b K z_p||^2 / ||R||^2, z_u and z_p the
 * network's velocity and pressure directions for the captured residual R and
 * K the augmented operator of that pair's viscosities.  Same sampling and
 * Adam as mdg_pcnet_train; losses (may be NULL) get the mean L per step. */
int mdg_pcnet_train_residual(mdg_ctx *ctx, int steps, int batch, float lr, unsigned long long seed,
                             double *losses);

/* Times the dominant kernel of the solve: the fine-level fused Chebyshev
 * smoother step (FP64) of the last FP64 solve (the vertex-patch step when
 * that solve used smoother 3, the Jacobi step otherwise), averaged over reps
 * launches with HIP events on the context stream.  Returns ms per launch and
 * the algorithmic (compulsory) bytes per launch.  No counterpart in the
 * reference; it feeds the roofline figure of the benchmark. */
int mdg_time_smoother(mdg_ctx *ctx, int reps, double *ms_per_launch, double *bytes_per_launch);

/* Time of the last timed call (P2G kernel, G2P and advection kernels,
 * energy step: ms, HIP events on the context stream). */
int mdg_last_kernel_ms(const mdg_ctx *ctx, double *ms);
/* Time of the last computation of the cached per-marker P2G weights (ms):
 * the first 1-cell-stencil P2G call on a node set after mdg_markers_set
 * computes each marker's nearest node and weight once; later calls on that
 * set read them.  No counterpart in the reference (which recomputes them in
 * every P2Mastah call); reported beside the P2G kernel times. */
int mdg_last_weights_ms(const mdg_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif
