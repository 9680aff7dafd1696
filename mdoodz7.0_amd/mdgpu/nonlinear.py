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
                                          or LineSearchDecoupled's step
  ExtractSolutions2 + ApplyBC          scatter below + Context.apply_bc, or
                                          Context.extract_solution on the device

With device=True the state is uploaded once per run() and every field of the
loop stays in device memory (DeviceArray): velocity, pressure, strain rates,
the rheology outputs, tangents, residual, scaling vector and correction.  The
host sees three residual norms per evaluation and the solver statistics.
"""
import time
import warnings

import numpy as np

from . import MDG_ERR_DIVERGED, RHEO_OUT, DeviceArray, RheoParams

# fields of the state that the loop reads (uploaded once in device mode)
_STATE_IN = ("u", "v", "p", "T", "mu_n", "mu_s", "perc_n", "perc_s", "sxxd0", "szzd0", "sxz0_n",
             "sxxd0_s", "szzd0_s", "sxz0", "rho_n", "rho_s", "pref_n", "pref_s", "d0_n", "d0_s", "phi0_n",
             "phi0_s")


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

    def __init__(self, ctx, eqn_u, eqn_v, eqn_p, phases, prm: RheoParams, device=False):
        self.ctx, self.phases, self.prm = ctx, phases, prm
        self.eqn = (eqn_u, eqn_v, eqn_p)
        self.device = device
        self.W = None          # device work arrays (device mode)

    def _work(self):
        """Device work arrays, allocated once per loop."""
        if self.W is not None:
            return self.W
        c = self.ctx
        nc, nv = (c.Nx - 1) * (c.Nz - 1), c.Nx * c.Nz
        neq = c.eqn_count()[0]
        nu, nw = c.Nx * (c.Nz + 1), (c.Nx + 1) * c.Nz
        sizes = dict(div=nc, exxd=nc, ezzd=nc, exz=nv, exz_n=nc, exxd_s=nv, ezzd_s=nv, div_s=nv,
                     T_s=nv, P_s=nv, D11=nc, D33=nv, eta_n=nc, F=neq, d=neq, sol=neq, rx=nu, rz=nw, u0=nu,
                     v0=nw, p0=nc, u_start=nu, v_start=nw, p_start=nc, Dn=8 * nc, Ds=4 * nv)
        W = {k: DeviceArray(c, n) for k, n in sizes.items()}
        W["cen"] = {k: DeviceArray(c, nc) for k in RHEO_OUT}
        W["ver"] = {k: DeviceArray(c, nv) for k in RHEO_OUT}
        self.W = W
        return W

    def node_inputs(self, st, sr):
        """Rheology inputs on centres and vertices (NonNewtonianViscosityGrid
        arguments; T and p_in carried to the vertices, RheologyDensity.c:1150-1151)."""
        c, W = self.ctx, self.W if self.device else None
        T_s = c.c2v(st["T"], W["T_s"] if W else None)
        P_s = c.c2v(st["p"], W["P_s"] if W else None)
        cin = dict(perc=st["perc_n"], G=st["mu_n"], T=st["T"], P=st["p"], exx=sr["exxd"],
                   ezz=sr["ezzd"], exz=sr["exz_n"], sxx0=st["sxxd0"], szz0=st["szzd0"],
                   sxz0=st["sxz0_n"])
        vin = dict(perc=st["perc_s"], G=st["mu_s"], T=T_s, P=P_s, exx=sr["exxd_s"],
                   ezz=sr["ezzd_s"], exz=sr["exz"], sxx0=st["sxxd0_s"], szz0=st["szzd0_s"],
                   sxz0=st["sxz0"])
        if "pref_n" in st:   # per-node multiplier of the power-law prefactor
            cin["pref"], vin["pref"] = st["pref_n"], st["pref_s"]
        if "d0_n" in st:     # grain size (d0_n / d0_s of the reference's grid)
            cin["d"], vin["d"] = st["d0_n"], st["d0_s"]
        if "phi0_n" in st:   # melt fraction (phi0_n / phi0_s)
            cin["phi"], vin["phi"] = st["phi0_n"], st["phi0_s"]
        return cin, vin

    def update_nonlinearity(self, st):
        c, W = self.ctx, self.W if self.device else None
        c.apply_bc(st["u"], st["v"])
        if W:
            sr = c.strain_rates(st["u"], st["v"], out={k: W[k] for k in ("div", "exxd", "ezzd", "exz")})
            sr.update(c.strain_rates_nodes(sr["exxd"], sr["ezzd"], sr["exz"], sr["div"],
                                           out={k: W[k] for k in ("exz_n", "exxd_s", "ezzd_s", "div_s")}))
        else:
            sr = c.strain_rates(st["u"], st["v"])
            sr.update(c.strain_rates_nodes(sr["exxd"], sr["ezzd"], sr["exz"], sr["div"]))
        cin, vin = self.node_inputs(st, sr)
        cen = c.viscosity_grid(self.phases, self.prm, 1, **cin, **(W["cen"] if W else {}))
        ver = c.viscosity_grid(self.phases, self.prm, 0, **vin, **(W["ver"] if W else {}))
        self.last_inputs = (cin, vin)
        return sr, cen, ver

    def _residual(self, cen, ver, sr, rx, rz, d=None):
        F = self.W["F"] if self.device else None
        return self.ctx.stokes_residual(cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"], sr["div"], rx, rz,
                                        d=d, F=F)

    def _step(self, st, dx, alpha):
        """ExtractSolutions2 + ApplyBC on the state."""
        if self.device:
            self.ctx.extract_solution(dx, alpha, st["u"], st["v"], st["p"])
        else:
            scatter(*self.eqn, dx, st["u"], st["v"], st["p"], alpha)
            self.ctx.apply_bc(st["u"], st["v"])

    def _save(self, st, slot="0"):
        """u, v, p into the work slot ("0": the line search's, "_start":
        the time step's start for safe-mode restarts)."""
        if self.device:
            W = self.W
            for k in ("u", "v", "p"):
                W[k + slot].copy_from(st[k])
            return tuple(W[k + slot] for k in ("u", "v", "p"))
        return [st[k].copy() for k in ("u", "v", "p")]

    def _restore(self, st, saved):
        for k, a in zip(("u", "v", "p"), saved):
            if self.device:
                st[k].copy_from(a)
            else:
                st[k][:] = a

    def residual_norms(self, st, d, rx, rz):
        sr, cen, ver = self.update_nonlinearity(st)
        _, norms = self._residual(cen, ver, sr, rx, rz, d=d)
        return np.asarray(norms, dtype=float)

    def line_search(self, st, dx, d, rx, rz, res_f, min_step=0.0, let_res_grow=0, newton=False, pc=False):
        """LineSearchDecoupled (StokesRoutines.c:582-807): six trial steps
        from alpha = maxalpha down to minalpha, each followed by
        UpdateNonLinearity and the residual norms, then the acceptance cascade
        on the combined, xz, x and z norms.  With pc (pc_line_search,
        :636-668) the step minalpha is tried first and taken without the scan
        if it lowers the combined norm.  Returns (alpha, stagnated) and
        leaves u, v, p as they were."""
        # Newton steps search [-1, -min_step], Picard steps [-2.5, 0] (:603-613)
        ntry = 6
        minalpha, maxalpha = (-1.0, -min_step) if newton else (-2.5, -0.0)
        saved = self._save(st)
        f0 = np.asarray(res_f, dtype=float)
        if pc:
            self._step(st, dx, minalpha)
            r_pc = self.residual_norms(st, d, rx, rz)
            self._restore(st, saved)
            if np.sqrt((r_pc ** 2).sum()) < np.sqrt((f0 ** 2).sum()):
                return minalpha, False
        dalpha = abs(maxalpha - minalpha) / (ntry - 1)
        trials, rows = [], []
        alpha = maxalpha
        for kk in range(ntry):
            if kk > 0:
                alpha -= dalpha
            trials.append(alpha)
            self._restore(st, saved)
            self._step(st, dx, alpha)
            rows.append(self.residual_norms(st, d, rx, rz))
        self._restore(st, saved)
        R = np.array(rows)
        comb = np.sqrt((R ** 2).sum(1))
        xz = np.hypot(R[:, 0], R[:, 1])
        best = [int(np.argmin(a)) for a in (comb, xz, R[:, 0], R[:, 1])]
        choice = None
        if comb[best[0]] < np.sqrt((f0 ** 2).sum()):
            choice = best[0]
        elif R[best[1], 0] < f0[0] and R[best[1], 1] < f0[1]:
            choice = best[1]
        elif R[best[2], 0] < f0[0]:
            choice = best[2]
        elif R[best[3], 1] < f0[1]:
            choice = best[3]
        alpha = trials[choice] if choice is not None else 0.0
        if abs(alpha) < 1e-13 or choice is None:
            return (min_step, False) if let_res_grow else (0.0, True)
        return alpha, False

    def run(self, st, nit_max, tol, diag_scaling=1, gx=0.0, gz=0.0, line_search=False,
            newton=False, picard2newton=True, picard2newton_tol=1e-1, max_its_pic=10,
            line_search_min=0.0, let_res_grow=0, pc_line_search=0, safe_mode=0, safe_dt_div=5.0,
            max_num_stag=3, **solver_opts):
        """Iterates in place on st; tol = (abs_mom, rel_mom, abs_div, rel_div).
        Returns (residual norm history, number of linear solves, solver stats);
        with line_search the chosen steps are in st["alphas"].  With newton,
        the loop switches from Picard to Newton once a relative residual of
        the previous iteration drops below picard2newton_tol or after
        max_its_pic iterations (Main_DOODZ.c:884-896); with picard2newton
        False (the reference's default Picard2Newton = 0, InputOutput.c:1193)
        every iteration is a Newton step.  Newton steps use the tangent
        coefficients and the line search (InputOutput.c:1334).
        st["newton_steps"] lists the iteration numbers solved with Newton,
        st["iter_ms"] the wall time of each iteration (rheology, residual,
        solve, line search and update).  In device mode st's fields are
        uploaded once (st["upload_ms"]) and u, v, p, eta_n, eta_s come back
        at the end.

        A stagnated line search ends the loop, unless safe_mode = 1 with
        elasticity (Main_DOODZ.c:1021-1052): then the time step self.prm.dt
        is divided by safe_dt_div, u, v, p return to their values at the call
        and the iterations restart from 0 (st["safe_restarts"] lists the
        iterations that stagnated); the max_num_stag-th stagnation raises
        (the reference exits).  line_search_min, let_res_grow and
        pc_line_search are the reference's keys of the same names."""
        c = self.ctx
        # the reference's continuity tolerances (InputOutput.c:1179-1180)
        # for KillerSolver's a-posteriori incompressibility check
        solver_opts.setdefault("lin_abs_div", 1e-9)
        solver_opts.setdefault("lin_rel_div", 1e-5)
        W = None
        S = st
        if self.device:
            t0 = time.perf_counter()
            W = self._work()
            S = {k: DeviceArray(c, np.size(st[k]), host=st[k]) for k in _STATE_IN if k in st}
            st["upload_ms"] = 1e3 * (time.perf_counter() - t0)
        rx, rz = c.eval_rhs(gx, gz, S["rho_n"], S["rho_s"], *((W["rx"], W["rz"]) if W else ()))
        abs_mom, rel_mom, abs_div, rel_div = tol
        hist, stats, res0 = [], [], None
        nm = c.eqn_count()[1]
        is_newton = bool(newton and not picard2newton)
        st["newton_steps"] = []
        st["diverged_solves"] = []
        st["iter_ms"] = []
        st["safe_restarts"] = []
        start = self._save(S, "_start") if safe_mode == 1 else None
        nstag, nit = 0, 0
        while nit <= nit_max:
            t_it = time.perf_counter()
            if newton and nit > 0 and not is_newton:
                rr = hist[-1] / res0
                if (rr < picard2newton_tol).any() or nit >= max_its_pic:
                    is_newton = True
            sr, cen, ver = self.update_nonlinearity(S)
            if W:
                # the line-search trials overwrite the rheology outputs: keep
                # this evaluation's viscosities
                D11, D33 = W["D11"], W["D33"].copy_from(ver["eta"])
                W["eta_n"].copy_from(cen["eta"])
                c.scale_into(cen["eta"], 2.0, D11)
            else:
                D11, D33 = 2.0 * cen["eta"], ver["eta"]
            d = c.stokes_diag(D11, D11, D33, *((W["d"],) if W else ())) if diag_scaling else None
            _, norms = self._residual(cen, ver, sr, rx, rz, d=d)
            norms = np.asarray(norms, dtype=float)
            if res0 is None:
                res0 = norms.copy()
            hist.append(norms.copy())
            r = norms / res0
            if ((norms[0] < abs_mom or r[0] < rel_mom) and (norms[1] < abs_mom or r[1] < rel_mom)
                    and (norms[2] < abs_div or r[2] < rel_div)):
                st["iter_ms"].append(1e3 * (time.perf_counter() - t_it))
                break
            F, _ = self._residual(cen, ver, sr, rx, rz)
            Fm, Fc = (F.view(0, nm), F.view(nm, len(F) - nm)) if W else (F[:nm], F[nm:])
            sol = W["sol"] if W else None
            if is_newton:
                cin, vin = self.last_inputs
                Dn = c.tangent_grid(self.phases, self.prm, 1, D=W["Dn"] if W else None, **cin)
                Ds = c.tangent_grid(self.phases, self.prm, 0, D=W["Ds"] if W else None, **vin)
                dx, o = c.newton_solver(Dn, Ds, D11, D33, Fm, Fc, accept_diverged=True, sol=sol,
                                        **solver_opts)
                st["newton_steps"].append(nit)
            else:
                dx, o = c.killer_solver(D11, D11, D33, Fm, Fc, accept_diverged=True, sol=sol,
                                        **solver_opts)
            if o.status == MDG_ERR_DIVERGED:
                # A missed momentum tolerance: KillerSolver carries on with the
                # correction after max_its (the line search or alpha decides what
                # it is worth).  A continuity residual above its a-posteriori
                # bound (Solvers.c:1333-1336) is not carried on with: the solver
                # returns MDG_ERR_INCOMPRESSIBLE for it, which raises above.
                warnings.warn("linear solve of non-linear iteration %d stopped at relative "
                              "residual %.3e" % (nit, o.rel_res))
                st["diverged_solves"].append(nit)
            stats.append(o)
            alpha, stagnated = -1.0, False
            if line_search or is_newton:
                alpha, stagnated = self.line_search(S, dx, d, rx, rz, norms, min_step=line_search_min,
                                                    let_res_grow=let_res_grow, newton=is_newton,
                                                    pc=bool(pc_line_search))
                st.setdefault("alphas", []).append(alpha)
            self._step(S, dx, alpha)
            st["iter_ms"].append(1e3 * (time.perf_counter() - t_it))
            if stagnated and safe_mode <= 0:
                break
            if stagnated and self.prm.elastic == 1 and safe_mode == 1:
                st["safe_restarts"].append(nit)
                self.prm.dt /= safe_dt_div
                self._restore(S, start)
                nstag += 1
                if nstag == max_num_stag:
                    raise RuntimeError("non-linear iterations stagnated %d times in safe mode" % nstag)
                res0, nit = None, 0
                continue
            nit += 1
        if W:
            for k in ("u", "v", "p"):
                st[k][:] = S[k].numpy()
            st["eta_n"], st["eta_s"] = W["eta_n"].numpy(), W["D33"].numpy()
            for a in S.values():
                a.free()
        else:
            st["eta_n"], st["eta_s"] = cen["eta"], ver["eta"]
        return hist, len(stats), stats
