"""Oracle non-linear (Picard) loop for visco-elasto-plastic Stokes.
TEST INFRASTRUCTURE ONLY.

Restates the order of operations of the non-linear loop of
Main_DOODZ.c:878-1062 with Newton = 0 and line_search = 0:
  UpdateNonLinearity (StokesRoutines.c:55):
    ApplyBC -> StrainRateComponents (RheologyDensity.c:2001-2097, with exz_n
    and the vertex copies of exxd/ezzd) -> NonNewtonianViscosityGrid
    (RheologyDensity.c:1131-1426; T and p_in carried to the vertices by
    InterpCentroidsToVerticesDouble at :1150-1151)
  RheologicalOperators (Picard: D11 = 2 eta_n, D33 = eta_s)
  ExtractDiagonalScale when diag_scaling (StokesRoutines.c:1260-1300)
  EvaluateStokesResidualDecoupled (norms of d .* F) and the convergence test
  of Main_DOODZ.c:1003
  SolveStokesDefectDecoupled with alpha = -1 (StokesRoutines.c:836-861).
"""
import numpy as np

from . import (apply_bc, assemble_picard, c2v, eval_rhs, exz_n, killer_solve, lib,
               residual, rheology_nodes)
from .driver import scatter_solution

RHEO_KEYS = ("eta", "eta_phys", "sxx", "szz", "sxz", "p_corr", "eII_pl")


def strain_rate_components(g, u, v):
    nc, nv = (g.Nx - 1) * (g.Nz - 1), g.Nx * g.Nz
    out = {k: np.zeros(nc) for k in ("div", "exxd", "ezzd")}
    out.update({k: np.zeros(nv) for k in ("exz", "wxz", "exxd_s", "ezzd_s", "div_s")})
    lib().orc_strain_rates(g.c, 0, u, v, out["div"], out["exxd"], out["ezzd"], out["exz"],
                           out["wxz"].ctypes.data, out["exxd_s"].ctypes.data,
                           out["ezzd_s"].ctypes.data, out["div_s"].ctypes.data)
    out["exz_n"] = exz_n(g, out["exz"])
    return out


def update_nonlinearity(g, st, md, phases, average):
    """Strain rates then the grid rheology on centres and vertices; returns
    (strain rates, centre outputs, vertex outputs)."""
    apply_bc(g, st["u"], st["v"])
    sr = strain_rate_components(g, st["u"], st["v"])
    T_s = c2v(g, st["T"])
    P_s = c2v(g, st["p"])
    cen = rheology_nodes(md, phases, st["perc_n"], st["mu_n"], st["T"], st["p"], sr["exxd"],
                         sr["ezzd"], sr["exz_n"], st["sxxd0"], st["szzd0"], st["sxz0_n"],
                         pref=st.get("pref_n"), tag=g.BCp_t, centroid=1, average=average,
                         d=st.get("d0_n"), phi=st.get("phi0_n"))
    ver = rheology_nodes(md, phases, st["perc_s"], st["mu_s"], T_s, P_s, sr["exxd_s"],
                         sr["ezzd_s"], sr["exz"], st["sxxd0_s"], st["szzd0_s"], st["sxz0"],
                         pref=st.get("pref_s"), tag=g.BCg_t, centroid=0, average=average,
                         d=st.get("d0_s"), phi=st.get("phi0_s"))
    if cen["status"] or ver["status"]:
        raise RuntimeError("visco-elastic local iterations failed")
    return sr, cen, ver


def picard_diag(g, num, eta_n, eta_s):
    """ExtractDiagonalScale: diag(A) on momentum rows, 1 on continuity rows."""
    A, _, _ = assemble_picard(g, num, eta_n, eta_s)
    return np.concatenate([A.diagonal(), np.ones(num[4])])


def line_search(g, num, st, dx, md, phases, average, d, rx, rz, res_f, min_step=0.0,
                let_res_grow=0, pc=False):
    """LineSearchDecoupled (StokesRoutines.c:582-807) for Picard steps: six
    trial steps alpha in [-2.5, 0] (evenly spaced from the top), each
    followed by UpdateNonLinearity and the residual, then the acceptance
    cascade on the combined, xz, x and z norms.  With pc (pc_line_search,
    :636-668) the longest step is tried first and taken without the scan if
    it lowers the combined norm.  Returns (alpha, stagnated); u, v, p are
    left unchanged."""
    ntry, minalpha, maxalpha = 6, -2.5, -0.0
    saved = [st[k].copy() for k in ("u", "v", "p")]

    def trial(alpha):
        st["u"][:], st["v"][:], st["p"][:] = saved
        scatter_solution(num, dx, st["u"], st["v"], st["p"], alpha)
        apply_bc(g, st["u"], st["v"])
        sr, cen, ver = update_nonlinearity(g, st, md, phases, average)
        _, norms = residual(g, num, cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"],
                            sr["div"], rx, rz, d)
        return norms.copy()

    frac = 1.0
    fxzp0 = np.sqrt((np.asarray(res_f) ** 2).sum())
    if pc:
        r_pc = trial(minalpha)
        st["u"][:], st["v"][:], st["p"][:] = saved
        if np.sqrt((r_pc ** 2).sum()) < frac * fxzp0:
            return minalpha, False
    dalpha = abs(maxalpha - minalpha) / (ntry - 1)
    alphav, R = [], []
    alpha = maxalpha
    for kk in range(ntry):
        if kk > 0:
            alpha -= dalpha
        alphav.append(alpha)
        R.append(trial(alpha))
    st["u"][:], st["v"][:], st["p"][:] = saved
    R = np.array(R)
    fxzp = np.sqrt((R ** 2).sum(1))
    fxz = np.sqrt(R[:, 0] ** 2 + R[:, 1] ** 2)
    # first index of the minimum, as the strict "<" scan of the reference
    ixzp, ixz, ix, iz = (int(np.argmin(a)) for a in (fxzp, fxz, R[:, 0], R[:, 1]))
    success = False
    if fxzp[ixzp] < frac * fxzp0:
        alpha, success = alphav[ixzp], True
    if not success and R[ixz, 0] < frac * res_f[0] and R[ixz, 1] < frac * res_f[1]:
        alpha, success = alphav[ixz], True
    if not success and R[ix, 0] < frac * res_f[0]:
        alpha, success = alphav[ix], True
    if not success and R[iz, 1] < frac * res_f[1]:
        alpha, success = alphav[iz], True
    stagnated = False
    if abs(alpha) < 1e-13 or not success:
        if let_res_grow:
            alpha = min_step
        else:
            alpha, stagnated = 0.0, True
    return alpha, stagnated


def picard_loop(g, st, md, phases, nit_max, tol, average=0, diag_scaling=1, gx=0.0, gz=0.0,
                line_search_on=False, line_search_min=0.0, let_res_grow=0, pc_line_search=0,
                safe_mode=0, safe_dt_div=5.0, max_num_stag=3, _line_search=None, **killer_kw):
    """Runs the loop in place on st (u, v, p updated).  tol = (abs_mom,
    rel_mom, abs_div, rel_div).  Returns the list of residual norm triples
    (one per evaluated iteration) and the number of linear solves; with
    line_search_on the chosen step lengths are in st["alphas"].

    A stagnated line search ends the loop, unless safe_mode = 1 with
    elasticity (Main_DOODZ.c:1021-1052): then md.dt is divided by
    safe_dt_div, u, v, p return to their values at the call and the
    iterations restart from 0; the max_num_stag-th stagnation raises (the
    reference exits).  st["safe_restarts"] lists the iteration numbers that
    stagnated.  _line_search replaces line_search (tests of the control
    flow)."""
    num = g.numbering()
    nm = num[3]
    rx, rz = eval_rhs(g, gx, gz, st["rho_n"], st["rho_s"])
    abs_mom, rel_mom, abs_div, rel_div = tol
    hist, nsolve, res0 = [], 0, None
    start = [st[k].copy() for k in ("u", "v", "p")]
    ls = _line_search or line_search
    st["safe_restarts"] = []
    nstag, nit = 0, 0
    while nit <= nit_max:
        sr, cen, ver = update_nonlinearity(g, st, md, phases, average)
        d = picard_diag(g, num, cen["eta"], ver["eta"]) if diag_scaling else None
        F, norms = residual(g, num, cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"],
                            sr["div"], rx, rz, d)
        if res0 is None:
            res0 = norms.copy()
        hist.append(norms.copy())
        r = norms / res0
        if ((norms[0] < abs_mom or r[0] < rel_mom) and (norms[1] < abs_mom or r[1] < rel_mom)
                and (norms[2] < abs_div or r[2] < rel_div)):
            break
        Fu, _ = residual(g, num, cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"], sr["div"],
                         rx, rz)
        A, B, Cm = assemble_picard(g, num, cen["eta"], ver["eta"])
        du, dp, stats, rc = killer_solve(g, num, A, B, Cm, Fu[:nm], Fu[nm:], **killer_kw)
        if rc != 0:
            raise RuntimeError("KillerSolver status %d at iteration %d (eta_n in [%g, %g], "
                               "eta_s in [%g, %g])" % (rc, nit, cen["eta"].min(), cen["eta"].max(),
                                                       ver["eta"].min(), ver["eta"].max()))
        dx = np.concatenate([du, dp])
        alpha, stagnated = -1.0, False
        if line_search_on:
            alpha, stagnated = ls(g, num, st, dx, md, phases, average, d, rx, rz, norms,
                                  min_step=line_search_min, let_res_grow=let_res_grow,
                                  pc=bool(pc_line_search))
            st.setdefault("alphas", []).append(alpha)
        scatter_solution(num, dx, st["u"], st["v"], st["p"], alpha)
        apply_bc(g, st["u"], st["v"])
        nsolve += 1
        if stagnated and safe_mode <= 0:
            break
        if stagnated and md.elastic == 1 and safe_mode == 1:
            st["safe_restarts"].append(nit)
            md.dt /= safe_dt_div
            st["u"][:], st["v"][:], st["p"][:] = start
            nstag += 1
            if nstag == max_num_stag:
                raise RuntimeError("non-linear iterations stagnated %d times in safe mode" % nstag)
            res0, nit = None, 0
            continue
        nit += 1
    st["eta_n"], st["eta_s"] = cen["eta"], ver["eta"]
    return hist, nsolve
