"""Host mirror of MDOODZ's non-linear Picard loop on the GPU entry points.

The loop body follows Main_DOODZ.c:878-1062 with Newton = 0 and
line_search = 0, calling the library where the reference calls its CPU
routines:
  UpdateNonLinearity (StokesRoutines.c:55)
      ApplyBC                          -> Context.apply_bc
      StrainRateComponents             -> Context.strain_rates + strain_rates_nodes
      InterpCentroidsToVerticesDouble  -> Context.c2v (T, p_in to the vertices)
      NonNewtonianViscosityGrid        -> Context.viscosity_grid (centres, vertices)
  RheologicalOperators (Picard)        D11 = D22 = 2 eta_n, D33 = eta_s
  ExtractDiagonalScale                 -> Context.stokes_diag
  EvaluateStokesResidualDecoupled      -> Context.stokes_residual
  convergence test                     Main_DOODZ.c:1003
  SolveStokesDefectDecoupled           -> Context.killer_solver, alpha = -1
  ExtractSolutions2 + ApplyBC          scatter below + Context.apply_bc
"""
import numpy as np

from . import RheoParams


def scatter(eqn_u, eqn_v, eqn_p, dx, u, v, p, alpha=-1.0):
    """ExtractSolutions2 (StokesRoutines.c:536-576) for free unknowns."""
    m = eqn_u >= 0
    u[m] += alpha * dx[eqn_u[m]]
    m = eqn_v >= 0
    v[m] += alpha * dx[eqn_v[m]]
    m = eqn_p >= 0
    p[m] += alpha * dx[eqn_p[m]]


class PicardLoop:
    """State and loop of one time step's non-linear iterations.

    st holds host arrays: u, v, p (grid layout), T, mu_n, mu_s, perc_n
    (nph x cells), perc_s (nph x vertices), the old stresses sxxd0, szzd0,
    sxz0_n (centres) and sxxd0_s, szzd0_s, sxz0 (vertices), rho_n, rho_s."""

    def __init__(self, ctx, eqn_u, eqn_v, eqn_p, phases, prm: RheoParams):
        self.ctx, self.phases, self.prm = ctx, phases, prm
        self.eqn = (eqn_u, eqn_v, eqn_p)

    def update_nonlinearity(self, st):
        c = self.ctx
        c.apply_bc(st["u"], st["v"])
        sr = c.strain_rates(st["u"], st["v"])
        sr.update(c.strain_rates_nodes(sr["exxd"], sr["ezzd"], sr["exz"], sr["div"]))
        T_s, P_s = c.c2v(st["T"]), c.c2v(st["p"])
        cen = c.viscosity_grid(self.phases, self.prm, 1, perc=st["perc_n"], G=st["mu_n"],
                               T=st["T"], P=st["p"], exx=sr["exxd"], ezz=sr["ezzd"],
                               exz=sr["exz_n"], sxx0=st["sxxd0"], szz0=st["szzd0"],
                               sxz0=st["sxz0_n"])
        ver = c.viscosity_grid(self.phases, self.prm, 0, perc=st["perc_s"], G=st["mu_s"],
                               T=T_s, P=P_s, exx=sr["exxd_s"], ezz=sr["ezzd_s"], exz=sr["exz"],
                               sxx0=st["sxxd0_s"], szz0=st["szzd0_s"], sxz0=st["sxz0"])
        return sr, cen, ver

    def run(self, st, nit_max, tol, diag_scaling=1, gx=0.0, gz=0.0, **solver_opts):
        """Iterates in place on st; tol = (abs_mom, rel_mom, abs_div, rel_div).
        Returns (residual norm history, number of linear solves, solver stats)."""
        c = self.ctx
        rx, rz = c.eval_rhs(gx, gz, st["rho_n"], st["rho_s"])
        abs_mom, rel_mom, abs_div, rel_div = tol
        hist, stats, res0 = [], [], None
        nm = c.eqn_count()[1]
        for nit in range(nit_max + 1):
            sr, cen, ver = self.update_nonlinearity(st)
            D11, D33 = 2.0 * cen["eta"], ver["eta"]
            d = c.stokes_diag(D11, D11, D33) if diag_scaling else None
            _, norms = c.stokes_residual(cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"],
                                         sr["div"], rx, rz, d=d)
            norms = np.asarray(norms, dtype=float)
            if res0 is None:
                res0 = norms.copy()
            hist.append(norms.copy())
            r = norms / res0
            if ((norms[0] < abs_mom or r[0] < rel_mom) and (norms[1] < abs_mom or r[1] < rel_mom)
                    and (norms[2] < abs_div or r[2] < rel_div)):
                break
            F, _ = c.stokes_residual(cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"],
                                     sr["div"], rx, rz)
            dx, o = c.killer_solver(D11, D11, D33, F[:nm], F[nm:], **solver_opts)
            stats.append(o)
            scatter(*self.eqn, dx, st["u"], st["v"], st["p"])
            c.apply_bc(st["u"], st["v"])
        st["eta_n"], st["eta_s"] = cen["eta"], ver["eta"]
        return hist, len(stats), stats
