/*
 * mdoracle — CPU FP64 restatement of the MDOODZ 7.0 per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * (libmdgpu.so).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the MDOODZ7.0 source tree).  Grid layouts are the reference's own:
 *   vertices   Nx*Nz        index k + l*Nx          (xg, zg)
 *   centres    (Nx-1)*(Nz-1) index k + l*(Nx-1)      (xc, zc)
 *   Vx nodes   Nx*(Nz+1)    index k + l*Nx          (xg, zvx)
 *   Vz nodes   (Nx+1)*Nz    index k + l*(Nx+1)      (xvz, zg)
 * BC tags follow Setup.c:206-230 (-1 free, 0/11 Dirichlet, 13 free slip,
 * 2 stress, -2/-12 periodic, 30 air, 31 surface pressure).
 */
#ifndef MDORACLE_H
#define MDORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int Nx, Nz;                 /* vertices (MDOODZ convention)          */
    double dx, dz;
    double *xg, *zg;            /* Nx, Nz                                */
    double *xc, *zc;            /* Nx-1, Nz-1                            */
    double *xvz, *zvx;          /* Nx+1, Nz+1                            */
    signed char *BCu_t, *BCv_t, *BCp_t, *BCg_t;
    double *BCu_v, *BCv_v;
} orc_grid;

/* GridRoutines.c:246-330 (running-sum coordinates) */
void orc_grid_coords(int Nx, int Nz, double xmin, double zmin, double xmax, double zmax,
                     double *dx, double *dz, double *xg, double *zg, double *xc, double *zc,
                     double *xvz, double *zvx);

/* Setup.c:206-600 with SetPureShearBCVx/Vz (Setup.c:907-981) */
void orc_bc_pure_shear(orc_grid *g, double bkg_strain_rate);

/* StokesRoutines.c:1162-1220 ; returns neq, writes neq_mom / neq_cont */
int orc_eqn_numbering(const orc_grid *g, int *eqn_u, int *eqn_v, int *eqn_p,
                      int *neq_mom, int *neq_cont);

/* ParticleRoutines.c:2300-2310 marker -> node index for one node set */
void orc_marker_index(const orc_grid *g, int centroid, int n, const double *x,
                      const double *z, int *ip, int *jp);

/* ParticleRoutines.c:2170-2665 (P2Mastah).  mode 0: thread-private buffers
 * and serial-order reduction; mode 3: atomic scatter.  phase_perc is
 * nb_phases contiguous node arrays (used when prop==1). */
void orc_p2mastah(const orc_grid *g, int n, const double *x, const double *z, const int *phase,
                  const double *mat_prop, int flag, int avg, int prop, int centroid,
                  int stencil, int periodic_x, const signed char *node_type,
                  double *node_field, int nb_phases, double *phase_perc, int mode,
                  int nthreads);

/* StokesAssemblyDecoupled.c:219-556 (Xjacobian_InnerNodesDecoupled3),
 * :732-1085 (Zjacobian_InnerNodesDecoupled3), :1786-2304 (row loops): the
 * Newton momentum blocks AJ (neq_mom x neq_mom) and BJ (neq_mom x neq_cont)
 * as CSR, for inner nodes of a box without periodicity or stress walls.
 * Dn = D11,D12,D13,D14,D21,D22,D23,D24 on the centres (8 x cells), Ds =
 * D31..D34 on the vertices (4 x vertices); co = comp*out_of_plane (the
 * reference fixes comp = 1 in this assembly).  Capacity: 21 entries per
 * row in AJ, 6 in BJ. */
void orc_assemble_jacobian(const orc_grid *g, const int *eqn_u, const int *eqn_v, const int *eqn_p,
                           int neq_mom, int neq_cont, const double *Dn, const double *Ds, double co,
                           int *Ap, int *Aj, double *Ax, int *Bp, int *Bj, double *Bx);

/* GridRoutines.c:108-212 */
void orc_c2v(const orc_grid *g, int periodic_x, const double *cent, double *vert);

/* ParticleRoutines.c:63-256 (V2P): grid velocity at the markers; u on the
 * Vx nodes, v on the Vz nodes; conserv = model.conserv_interp (Taras'
 * correction); inactive markers (phase -1) get 0. */
void orc_v2p(const orc_grid *g, const double *u, const double *v, int conserv, int n, const double *x,
             const double *z, const int *phase, double *vx, double *vz);

/* AdvectionRoutines.c:347-410 (RogerGuntherII, RK4 with V2P at each stage)
 * and :320-341 (isoutPart, non-periodic): x, z, phase in place; vx, vz (may
 * be NULL) receive the RK4 velocity of the markers still active. */
void orc_advect_rk4(const orc_grid *g, const double *u, const double *v, double dt, int conserv,
                    double xmin, double xmax, double zmin, double zmax, int n, double *x, double *z,
                    int *phase, double *vx, double *vz);

/* ParticleRoutines.c:1332-1580 (Interp_Grid2P_centroids2, non-periodic):
 * a centre field at the markers; inactive markers get 0. */
void orc_g2p_centroids(const orc_grid *g, const double *F, int n, const double *x, const double *z,
                       const int *phase, double *out);

/* ParticleRoutines.c:1223-1325 (Interp_Grid2P) on the vertices: a vertex
 * field at the markers, unnormalised; inactive markers get 0. */
void orc_g2p_vertices(const orc_grid *g, const double *F, int n, const double *x, const double *z,
                      const int *phase, double *out);

/* RheologyParticles.c:1725-2082 (UpdateParticleStress: elastic, no subgrid
 * stress diffusion, no rotation): grid stress increments added to the
 * marker stresses in place. */
void orc_update_particle_stress(const orc_grid *g, const double *sxxd, const double *szzd, const double *sxz,
                                const double *sxxd0, const double *szzd0, const double *sxz0, int n,
                                const double *x, const double *z, const int *phase, double *msxxd,
                                double *mszzd, double *msxz);

/* RheologyParticles.c:1621-1704 (UpdateParticlePressure, no subgrid
 * diffusion): dp = p_in - p0_n off air cells, interpolated and added to P. */
void orc_update_particle_pressure(const orc_grid *g, const double *p_in, const double *p0_n, double *dp,
                                  int n, const double *x, const double *z, const int *phase, double *P);

/* UpdateParticleStress / UpdateParticlePressure with subgrid diffusion
 * (RheologyParticles.c:1811-1895, :1621-1695): per-phase G / beta tables,
 * vertex / centre physical viscosity, P2Mastah interp_stencil (stress). */
void orc_update_particle_stress_subgrid(const orc_grid *g, double dt, const double *G_phase, int stencil,
                                        const double *sxxd, const double *szzd, const double *sxz,
                                        const double *sxxd0, const double *szzd0, const double *sxz0,
                                        const double *eta_phys_s, int n, const double *x, const double *z,
                                        const int *phase, double *msxxd, double *mszzd, double *msxz);
void orc_update_particle_pressure_subgrid(const orc_grid *g, double dt, const double *beta_phase,
                                          const double *p_in, const double *p0_n, const double *eta_phys_n,
                                          double *dp, int n, const double *x, const double *z, const int *phase,
                                          double *P);

/* StokesRoutines.c:415-~500 (ghost values from BC tags) */
void orc_apply_bc(const orc_grid *g, double *u, double *v);

/* RheologyDensity.c:2001-2097 */
void orc_strain_rates(const orc_grid *g, int compressible, const double *u, const double *v,
                      double *div, double *exxd, double *ezzd, double *exz, double *wxz,
                      double *exxd_s, double *ezzd_s, double *div_s);

/* RheologyDensity.c:2066-2085 (exz averaged to the cells) */
void orc_exz_n(const orc_grid *g, const double *exz, double *exz_n);

/* StokesRoutines.c:961-1160 */
void orc_eval_rhs(const orc_grid *g, double gx, double gz, const double *rho_n,
                  const double *rho_s, double *roger_x, double *roger_z);

/* StokesAssemblyDecoupled.c:184-211,688-724,1128-1144 (Assemble=0) and
 * StokesRoutines.c:867-955 (scaling by d and interior RMS norms).
 * F is in equation numbering (neq entries); d may be NULL (=1). */
void orc_residual(const orc_grid *g, const int *eqn_u, const int *eqn_v, const int *eqn_p,
                  int neq_mom, const double *sxxd, const double *szzd, const double *sxz,
                  const double *p_corr, const double *div, const double *roger_x,
                  const double *roger_z, const double *rhs_p, const double *d, double *F,
                  double norms[3]);

/* StokesAssemblyDecoupled.c:62-212,562-725,1090-1145,1295-1565 (Assemble=1),
 * incompressible, no stabilisation.  CSR (row pointers, cols, vals), columns in
 * block-local numbering.  Capacity: A 9 per row, B/C 4 per row. Returns nnz. */
void orc_assemble_picard(const orc_grid *g, const int *eqn_u, const int *eqn_v,
                         const int *eqn_p, int neq_mom, int neq_cont, const double *D11,
                         const double *D22, const double *D33,
                         int *Ap, int *Aj, double *Ax, int *Bp, int *Bj, double *Bx,
                         int *Cp, int *Cj, double *Cx);

typedef struct {
    double penalty;            /* model.penalty (gamma)                  */
    double celvol;             /* dx*dz                                  */
    int    max_its_PH;         /* InputOutput.c:1184 default 10          */
    double rel_tol_KSP;        /* default 1e-4                           */
    int    max_its_KSP;        /* default 1000                           */
    double lin_abs_div, lin_rel_div, lin_abs_mom, lin_rel_mom;
    /* outputs */
    int    its_PH, its_KSP_tot;
    double maxdiv, maxdiv0, maxru, maxru0;
    int    status;             /* 0 ok, -1 not SPD, -2 PH abort          */
} orc_killer_opts;

/* Solvers.c:124-230 (kspgcr) and Solvers.c:1054-1454 (KillerSolver) for the
 * incompressible case.  The CHOLMOD factor of Jts = A - B*(penalty*I)*C is
 * replaced by an exact banded Cholesky factor of the same matrix. */
int orc_killer_solve(int neq_mom, int neq_cont,
                     const int *Ap, const int *Aj, const double *Ax,
                     const int *Bp, const int *Bj, const double *Bx,
                     const int *Cp, const int *Cj, const double *Cx,
                     const double *rhs_mom, const double *rhs_cont, double *du, double *dp,
                     orc_killer_opts *opts);

/* The same on the Newton system (Solvers.c:1176-1193,1276-1314): GCR iterates
 * on Js = JA - JB (penalty/dV) C and the residuals use JA, JB, while the
 * factor stays that of the Picard Jts = A - B (penalty/dV) C
 * (preconditioner = 0, the default, InputOutput.c:1177). */
int orc_killer_solve_newton(int neq_mom, int neq_cont,
                            const int *Ap, const int *Aj, const double *Ax,
                            const int *Bp, const int *Bj, const double *Bx,
                            const int *JAp, const int *JAj, const double *JAx,
                            const int *JBp, const int *JBj, const double *JBx,
                            const int *Cp, const int *Cj, const double *Cx,
                            const double *rhs_mom, const double *rhs_cont, double *du, double *dp,
                            orc_killer_opts *opts);

/* AccumulatedStrainII (RheologyParticles.c:346-649): eII[0..6] = eII_pl,
 * eII_pwl, eII_exp, eII_lin, eII_gbs, eII_el, eII_cst on the centres (NULL =
 * 0), strain[0..7] = strain, strain_el, strain_pl, strain_pl_vol,
 * strain_pwl, strain_exp, strain_lin, strain_gbs of the markers (NULL =
 * skipped).  Returns -1, or the first cell with a negative increment. */
int orc_accumulated_strain(const orc_grid *g, double dt, const double *const *eII, const double *div_u_pl, int n,
                           const double *x, const double *z, const int *phase, double *const *strain);

/* Stress rotation of UpdateParticleStress (RheologyParticles.c:1965-1988):
 * mode 1 Jaumann rate, mode 2 analytical rotation, with the markers' spin. */
void orc_rotate_stress(int mode, double dt, int n, const int *phase, const double *wxzm, double *sxxd, double *szzd,
                       double *sxz);

/* CountPartCell_OLD re-seeding (ParticleRoutines.c:3224-4050), no free
 * surface, no periodic x: returns the number of new markers, -1 for an
 * empty 5 x 5 neighbourhood, -2 when capacity is exceeded. */
int orc_reseed(const orc_grid *g, double xmin, double xmax, int min_part_cell, int n, int capacity, double *x,
               double *z, int *phase, int *new_slot, int *new_src, int *n_out);

/* EnergyDirectSolve (ThermalRoutines.c:74-470) matrix for a box without
 * air cells: CSR rows in cell order (<= 5 entries), boundary rhs bbc. */
void orc_energy_matrix(const orc_grid *g, double dt, const double *rho_n, const double *Cp, const double *kx,
                       const double *kz, const signed char *BCT_t, const double *BCT_v, int *Ap, int *Aj, double *Ax,
                       double *bbc);
/* One implicit energy step with SolveThermalPCG (ThermalSolver.c:338-480):
 * T updated in place; returns the PCG iterations (or the reference's
 * negative failure codes). */
int orc_energy_step(const orc_grid *g, double dt, const double *rho_n, const double *Cp, const double *Qr,
                    const double *Wdiss, const double *Ha, const double *kx, const double *kz, const signed char *BCT_t,
                    const double *BCT_v, int max_its, double rel_tol, double *T);

/* Material parameters of one phase (mat_prop of mdoodz.h:158-170, the
 * members ViscosityConcise reads for the laws below; units are the model's
 * scaled units). */
typedef struct {
    int    cstv, pwlv, plast;
    double eta0;                                     /* constant viscosity   */
    double pref_pwl, F_pwl, A_pwl, n_pwl, Q_pwl, V_pwl; /* power-law creep   */
    double C, fric, dil, Slim, eta_vp, n_vp;         /* Drucker-Prager       */
    double beta_for_K;                               /* compressibility beta */
    int    linv, gbsv, expv, reserved;
    double m_pwl, r_pwl, f_pwl, a_pwl;               /* power-law d, f, phi  */
    double F_lin, A_lin, n_lin, m_lin, r_lin, f_lin, a_lin, Q_lin, V_lin; /* diffusion */
    double F_gbs, A_gbs, n_gbs, m_gbs, r_gbs, f_gbs, a_gbs, Q_gbs, V_gbs; /* GBS */
    double G_exp, S_exp, Q_exp, E_exp, n_exp;        /* Peierls (Gexp ...)   */
    double gs_ref;                                   /* reference grain size */
} orc_phase;

typedef struct {
    int    step, elastic;
    double dt, R, zeroC, min_eta, max_eta, eta_tol, Tiimin, scale_E, scale_eta;
    double T_peierls, d_min;     /* TmaxPeierls, 1e-13 m scaled */
} orc_rheo_model;

/* RheologyDensity.c:385-1080 (constant, power-law, diffusion, GBS and
 * Peierls creep, elasticity, Drucker-Prager plast = 1) at grain size d and
 * melt fraction phi.  pref_mult multiplies the power-law prefactor (1 in
 * the reference).  *status gets -1 where the reference exits on failed
 * local iterations, -2 where it exits on a grain size below d_min. */
double orc_viscosity_concise(const orc_phase *m, const orc_rheo_model *md, double G, double T,
                             double P, double d, double phi, double Exx, double Ezz, double Exz,
                             double fric, double C, double dil, double pref_mult, double *etaVE,
                             double *Pcorr, double *Eii_pl, int *status);

/* RheologyDensity.c:1157-1316 / 1320-1407 on one node set.  Nullable
 * arrays: fric, C, dil (phase values), pref (1), d, phi (0), outputs
 * except eta_out. */
int orc_rheology_nodes(const orc_rheo_model *md, const orc_phase *phases, int nph, int n,
                       const double *perc, const double *G, const double *T, const double *P,
                       const double *exx, const double *ezz, const double *exz,
                       const double *sxx0, const double *szz0, const double *sxz0,
                       const double *fric, const double *C, const double *dil,
                       const double *pref, const double *d, const double *phi,
                       const signed char *tag, int centroid, int average,
                       double *eta_out, double *eta_phys, double *sxx, double *szz, double *sxz,
                       double *p_corr, double *eII_pl);

/* Newton tangent coefficients by finite differences of the averaged
 * viscosity (StokesRoutines.c:250-375, FD_Jacobian.c:48-323). */
int orc_tangent_nodes(const orc_rheo_model *md, const orc_phase *phases, int nph, int n,
                      const double *perc, const double *G, const double *T, const double *P,
                      const double *exx, const double *ezz, const double *exz,
                      const double *sxx0, const double *szz0, const double *sxz0,
                      const double *fric, const double *C, const double *dil,
                      const double *pref, const double *d, const double *phi,
                      const signed char *tag, int centroid, int average, double *D);

/* Per-phase elastic moduli and softening parameters (mat_prop: G, bet, C,
 * C_end, phi, phi_end, psi, psi_end, pls_start, pls_end, H_c, *_soft). */
typedef struct {
    double G, bet, C, C_end, phi, phi_end, psi, psi_end, pls_start, pls_end, H_c;
    int    phi_soft, psi_soft, coh_soft;
} orc_phase_soft;

/* ShearModCompExpGrid + CohesionFrictionDilationGrid on one node set. */
void orc_material_grid(const orc_phase_soft *ph, int nph, int n, const double *perc,
                       const double *strain_pl, const signed char *tag, int centroid, int style,
                       double *mu, double *bet, double *C, double *fric, double *dil);

/* Row-interleaved momentum DOF order (Vx row l, then Vz row l) used to
 * give the exact factor of orc_killer_solve a narrow band. */
void orc_set_band_order(const orc_grid *g, const int *eqn_u, const int *eqn_v, int neq_mom);

#ifdef __cplusplus
}
#endif
#endif
