"""Control flow of the non-linear loop around the line search, on the
oracle loop (CPU): the predictor step of pc_line_search
(StokesRoutines.c:636-668) and the safe-mode restarts of Main_DOODZ.c:
1021-1052.  The GPU loop runs the same control flow against these
(test_gpu_nonlinear.py)."""
import numpy as np
import pytest

import oracle as O
from oracle import nonlinear as NL
from mdgpu import setups as S

TOL = (1e-12, 1e-8, 1e-12, 1e-8)
KILLER = dict(penalty=1e3, max_its_PH=50, rel_tol_KSP=1e-12, max_its_KSP=200, lin_abs_div=1e-14,
              lin_rel_div=1e-12, lin_abs_mom=1e-14, lin_rel_mom=1e-12)


def _setup(dt=1.0, elastic=1, C=2.0, n=15):
    from test_gpu_nonlinear import _problem
    g, st = _problem(O, n, seed=4)
    phs = []
    for d in S.vep_phases(C=C):
        d = dict(d)
        b = d.pop("beta")
        phs.append(O.OrcPhase(beta_for_K=b, **d))
    md = O.OrcRheoModel(step=1, elastic=elastic, dt=dt, R=1.0, zeroC=0.0, min_eta=1e-4, max_eta=1e4,
                        eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
    return g, st, md, phs


def _run(nit_max=6, dt=1.0, elastic=1, penalty=1e3, **kw):
    g, st, md, phs = _setup(dt=dt, elastic=elastic)
    hist, ns = NL.picard_loop(g, st, md, phs, nit_max, TOL, diag_scaling=1, gz=-1.0, line_search_on=True,
                              **dict(KILLER, penalty=penalty), **kw)
    return st, md, hist, ns


def _scripted(stagnate_on):
    """A line search that reports stagnation on the given call numbers (1-based)
    and otherwise runs the oracle's."""
    calls = []

    def ls(*a, **k):
        calls.append(1)
        if len(calls) in stagnate_on:
            return 0.0, True
        return NL.line_search(*a, **k)
    return ls


def test_predictor_step_rule():
    """With pc_line_search the longest step (-2.5 for Picard) is taken when
    its combined residual is below the current one; otherwise the scan's
    choice stands."""
    seen = []

    def both(g, num, st, dx, md, phases, average, d, rx, rz, res_f, **k):
        a_pc = NL.line_search(g, num, st, dx, md, phases, average, d, rx, rz, res_f, pc=True)
        a_scan = NL.line_search(g, num, st, dx, md, phases, average, d, rx, rz, res_f, pc=False)
        saved = [st[q].copy() for q in ("u", "v", "p")]
        NL.scatter_solution(num, dx, st["u"], st["v"], st["p"], -2.5)
        NL.apply_bc(g, st["u"], st["v"])
        sr, cen, ver = NL.update_nonlinearity(g, st, md, phases, average)
        _, r = NL.residual(g, num, cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"], sr["div"], rx, rz, d)
        st["u"][:], st["v"][:], st["p"][:] = saved
        takes = np.sqrt((np.asarray(r) ** 2).sum()) < np.sqrt((np.asarray(res_f) ** 2).sum())
        seen.append((takes, a_pc, a_scan))
        return a_pc
    _run(_line_search=both)
    assert len(seen) >= 3
    for takes, a_pc, a_scan in seen:
        assert a_pc == (-2.5, False) if takes else a_pc == a_scan
    # on this box the predictor is taken where the scan would pick a shorter step
    assert any(takes and a_scan[0] != -2.5 for takes, _, a_scan in seen)


def test_safe_mode_restart_is_a_fresh_run_at_the_reduced_step():
    st1, md1, hist1, ns1 = _run(safe_mode=1, _line_search=_scripted({2}))
    assert st1["safe_restarts"] == [1] and md1.dt == 0.2
    st2, md2, hist2, ns2 = _run(dt=0.2)
    assert st2["safe_restarts"] == []
    for k in ("u", "v", "p", "eta_n"):
        assert np.array_equal(st1[k], st2[k]), k
    assert np.array_equal(np.array(hist1[-len(hist2):]), np.array(hist2))
    assert ns1 == ns2 + 2


def test_safe_mode_gives_up_after_max_num_stag():
    with pytest.raises(RuntimeError, match="stagnated 3 times"):
        _run(safe_mode=1, max_num_stag=3, _line_search=_scripted(set(range(1, 100))))


def test_stagnation_without_safe_mode_ends_the_loop():
    st, md, hist, ns = _run(_line_search=_scripted({2}))
    assert ns == 2 and len(hist) == 2 and md.dt == 1.0 and st["alphas"][-1] == 0.0


def test_safe_mode_needs_elasticity():
    # safe_mode = 1 without elasticity neither restarts nor stops: the
    # iterations go on from the unchanged state
    st, md, hist, ns = _run(elastic=0, safe_mode=1, nit_max=3, penalty=1e5, _line_search=_scripted({1}))
    assert st["safe_restarts"] == [] and md.dt == 1.0 and st["alphas"][0] == 0.0 and ns >= 2
    assert np.array_equal(hist[1], hist[0])     # the stagnated step changed nothing
