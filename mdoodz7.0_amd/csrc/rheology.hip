// Grid rheology: NonNewtonianViscosityGrid / ViscosityConcise for constant,
// power-law, Maxwell-elastic and Drucker-Prager (plast = 1) phases, and the
// centre-to-vertex averaging used around it.
//
// One thread per node.  The phase table sits in LDS (a few hundred bytes);
// phase fractions are phase-major (perc[p*n + i]) so that for a given phase
// a wavefront reads 64 consecutive doubles.  The per-phase work is dominated
// by FP64 pow/exp inside the two local Newton iterations, so the kernel is
// bounded by FP64 issue rather than by its ~20 doubles of traffic per node.
#include "mdg_internal.h"

namespace {

struct rheo_ctl {
    int    step, elastic, average, nph;
    double dt, R, zeroC, min_eta, max_eta, eta_tol, Tiimin, scale_E, scale_eta;
};

// LocalIterationViscoElastic (RheologyDensity.c:152-197): Newton iterations on
// eta_ve so that E = el*T/(2 eta_el) + T/(2 eta_cst) + C_pwl T^n, T = 2 eta_ve E.
// Returns false when the reference would exit (20 iterations, res > tol).
__device__ bool local_ve(double Eii, double &eta, bool elastic, double eta_el, bool constant,
                         double eta_cst, bool dislocation, double C_pwl, double n_pwl,
                         double &Eii_pwl) {
    const double tol = 1.0e-11;
    for (int it = 0; it < 20; it++) {
        const double Tii = 2.0 * eta * Eii;
        const double e_cst = constant ? Tii / 2.0 / eta_cst : 0.0;
        const double e_pwl = dislocation ? C_pwl * pow(Tii, n_pwl) : 0.0;
        Eii_pwl = e_pwl;
        const double r = Eii - (elastic ? 1.0 : 0.0) * Tii / (2.0 * eta_el) - (e_pwl + e_cst);
        const double res = fabs(r / Eii);
        if (res < tol / 100) return true;
        if (it == 19 && res > tol) return false;
        double dfdeta = 0.0;
        if (elastic) dfdeta += -Eii / eta_el;
        if (dislocation) dfdeta += -e_pwl * n_pwl / eta;
        if (constant) dfdeta += -Eii / eta_cst;
        eta -= r / dfdeta;
    }
    return true;
}

// ViscosityConcise (RheologyDensity.c:385-1080) for one (node, phase):
// returns the dissipative viscosity; etaVE, Pcorr, Eii_pl by reference.
__device__ double viscosity_concise(const mdg_phase_prop &m, const rheo_ctl &md, double G, double T,
                                    double P, double Exx, double Ezz, double Exz, double fric,
                                    double C, double dil, double pref_mult, double &etaVE,
                                    double &Pcorr, double &Eii_pl, bool &ok) {
    const double dt = (md.step == 0) ? 1e100 : md.dt;
    bool elastic = md.elastic != 0, plastic = (m.plast == 1);
    if (md.step == 0) { elastic = false; plastic = false; }          // :483
    const bool constant = m.cstv != 0, dislocation = m.pwlv != 0;
    double K = 1.0 / m.beta;
    const double Eyy = -(Exx + Ezz);
    double Eii = sqrt(0.5 * (Exx * Exx + Ezz * Ezz + Eyy * Eyy) + Exz * Exz);   // :457-459
    if (Eii * md.scale_E < 1e-30) Eii = 1e-30 / md.scale_E;
    Pcorr = P;
    if (!elastic) { G = 1e1; dil = 0.0; K = 1e10; }                   // :493
    if (T < md.zeroC) T = md.zeroC;
    double B_pwl = 0.0, C_pwl = 0.0;
    if (dislocation) {                                                 // :501-504
        B_pwl = m.pref_pwl * pref_mult * m.F_pwl * pow(m.A_pwl, -1.0 / m.n_pwl) *
                exp((m.Q_pwl + P * m.V_pwl) / md.R / m.n_pwl / T);
        C_pwl = pow(2.0 * B_pwl, -m.n_pwl);
    }
    const double sin_fric = sin(fric), cos_fric = cos(fric), sin_dil = sin(dil);
    double Tyield = 0.0;
    if (plastic) {                                                     // :530-539
        Tyield = C * cos_fric + P * sin_fric;
        if (m.Slim < Tyield) Tyield = m.Slim;
    }
    const double eta_el = G * dt;
    const double eta_cst = constant ? m.eta0 : 0.0;
    double eta_up = 1.0e100 / md.scale_eta;                            // :585-592
    if (constant) eta_up = fmin(eta_up, eta_cst);
    if (dislocation) eta_up = fmin(eta_up, B_pwl * pow(Eii, 1.0 / m.n_pwl - 1.0));
    if (elastic) eta_up = fmin(eta_up, eta_el);
    double eta_ve = eta_up, Eii_pwl = 0.0;
    if (!local_ve(Eii, eta_ve, elastic, eta_el, constant, eta_cst, dislocation, C_pwl, m.n_pwl,
                  Eii_pwl))
        ok = false;
    double Tii = 2.0 * eta_ve * Eii;
    bool is_pl = false;
    double eta_vep = eta_ve, gdot = 0.0;
    if (plastic) {                                                     // :655-707
        double F_trial = Tii - Tyield;
        if (F_trial > 1e-17) {
            is_pl = true;
            const double F0 = F_trial, tol = md.eta_tol;
            const double Kdt_ss = K * dt * sin_fric * sin_dil;
            double eta_vp = m.eta_vp * pow(Eii, 1.0 / m.n_vp - 1);
            gdot = F_trial / (eta_ve + eta_vp + Kdt_ss);
            double Pc = P, Tiic = Tii;
            for (int it = 0; it < 20; it++) {
                Pc = P + K * dt * (gdot * sin_dil);
                eta_vp = m.eta_vp * pow(fabs(gdot), 1.0 / m.n_vp - 1.0);
                Tyield = Tyield + gdot * eta_vp;    // accumulates, as in the reference
                Tiic = Tii - eta_ve * gdot;
                F_trial = Tiic - Tyield;
                const double res = fabs(F_trial);
                if (res < tol || res / F0 < tol) break;
                gdot -= F_trial / (-eta_ve - eta_vp / m.n_vp - Kdt_ss);
            }
            if (Tiic < 0.0) {                                          // :694-701
                gdot = (Tii - md.Tiimin) / eta_ve;
                Tiic = md.Tiimin;
                Pc = P;
            }
            Pcorr = Pc;
            eta_vep = Tiic / (2.0 * Eii);
            Tii = Tiic;
        }
    }
    etaVE = is_pl ? eta_vep : eta_ve;                                  // :1293-1301
    Eii_pl = is_pl ? gdot / 2.0 : 0.0;
    double inv = 0.0;                                                  // :1024-1030
    if (dislocation) inv += 1.0 / (B_pwl * pow(Eii_pwl, 1.0 / m.n_pwl - 1.0));
    if (constant) inv += 1.0 / eta_cst;
    if (is_pl) inv += 1.0 / (Tii / gdot);
    double eta = 1.0 / inv;
    etaVE = fmax(fmin(etaVE, md.max_eta), md.min_eta);                 // :1058-1073
    if (eta > md.max_eta) eta = md.max_eta;
    if (eta < md.min_eta) eta = md.min_eta;
    return eta;
}

__global__ void __launch_bounds__(256)
k_viscosity_grid(rheo_ctl md, const mdg_phase_prop *__restrict__ phases, int n,
                 const signed char *__restrict__ tag, int centroid, mdg_rheo_nodes f,
                 int *__restrict__ fail) {
    __shared__ mdg_phase_prop sph[MDG_MAX_PHASES];
    for (int p = threadIdx.x; p < md.nph; p += blockDim.x) sph[p] = phases[p];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const signed char t = tag[i];
    if (t == 30 || (centroid && t == 31)) {
        f.eta[i] = 0.0;
        if (f.eta_phys) f.eta_phys[i] = 0.0;
        if (f.sxx) f.sxx[i] = 0.0;
        if (f.szz) f.szz[i] = 0.0;
        if (f.sxz) f.sxz[i] = 0.0;
        if (f.p_corr) f.p_corr[i] = 0.0;
        if (f.eII_pl) f.eII_pl[i] = 0.0;
        return;
    }
    const double el = md.elastic ? 1.0 : 0.0;
    const double G = f.G[i];
    const double eta_e = md.elastic ? md.dt * G : 1.0;                 // :1197-1200
    const double Exx = f.exx[i] + (f.sxx0 ? el * f.sxx0[i] / eta_e / 2.0 : 0.0);
    const double Ezz = f.ezz[i] + (f.szz0 ? el * f.szz0[i] / eta_e / 2.0 : 0.0);
    const double Exz = f.exz[i] + (f.sxz0 ? el * f.sxz0[i] / eta_e / 2.0 : 0.0);
    const double T = f.T[i], P = f.P[i];
    const double pref = f.pref ? f.pref[i] : 1.0;
    double e = 0.0, ep = 0.0, pc = 0.0, epl = 0.0;
    bool ok = true;
    for (int p = 0; p < md.nph; p++) {
        const double w = f.perc[(size_t)p * n + i];
        if (!(fabs(w) > 1e-13)) continue;                              // :1212-1215
        const mdg_phase_prop &m = sph[p];
        double etaVE, Pcorr, Eii_pl;
        const double eta = viscosity_concise(m, md, G, T, P, Exx, Ezz, Exz,
                                             f.fric ? f.fric[i] : m.fric, f.C ? f.C[i] : m.C,
                                             f.dil ? f.dil[i] : m.dil, pref, etaVE, Pcorr,
                                             Eii_pl, ok);
        if (md.average == 0) { e += w * etaVE; ep += w * eta; }
        else if (md.average == 1) { e += w / etaVE; ep += w / eta; }
        else { e += w * log(etaVE); ep += w * log(eta); }
        pc += w * Pcorr;
        epl += w * Eii_pl;
    }
    if (md.average == 1) { e = 1.0 / e; ep = 1.0 / ep; }
    else if (md.average == 2) { e = exp(e); ep = exp(ep); }
    f.eta[i] = e;
    if (f.eta_phys) f.eta_phys[i] = ep;
    if (f.sxx) f.sxx[i] = 2.0 * e * Exx;                               // :1311-1313
    if (f.szz) f.szz[i] = 2.0 * e * Ezz;
    if (f.sxz) f.sxz[i] = 2.0 * e * Exz;
    if (f.p_corr) f.p_corr[i] = pc;
    if (f.eII_pl) f.eII_pl[i] = epl;
    if (!ok) atomicOr(fail, 1);
}

// InterpCentroidsToVerticesDouble (GridRoutines.c:108-212) without the padded
// copy: the padded index of a vertex's neighbour cell is clamped (edge copy)
// or wrapped (periodic) onto the interior cells directly.
__global__ void __launch_bounds__(256)
k_c2v(int Nx, int Nz, int per, const double *__restrict__ cent, const signed char *__restrict__ tp,
      const signed char *__restrict__ tg, double *__restrict__ vert) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nx * Nz) return;
    const int Ncx = Nx - 1, Ncz = Nz - 1;
    const int k = i % Nx, l = i / Nx;
    if (tg[i] == 30) { vert[i] = 0.0; return; }
    double s = 0.0;
    for (int dl = -1; dl <= 0; dl++) {
        // padded row l+dl+1 in [0, Ncz+1]; rows 0 and Ncz+1 copy rows 1 and Ncz
        const int lc = min(max(l + dl, 0), Ncz - 1);
        for (int dk = -1; dk <= 0; dk++) {
            int kc = k + dk;
            if (kc < 0) kc = per ? Ncx - 1 : 0;
            else if (kc > Ncx - 1) kc = per ? 0 : Ncx - 1;
            const int c = kc + lc * Ncx;
            const signed char t = tp[c];
            if (t != 30 && t != 31) s += cent[c];
        }
    }
    vert[i] = 0.25 * s;
}

// exz_n = 1/4 sum of the non-air vertex values (RheologyDensity.c:2066-2085)
__global__ void __launch_bounds__(256)
k_v2c_exz(int Nx, int Nz, const double *__restrict__ exz, const signed char *__restrict__ tp,
          const signed char *__restrict__ tg, double *__restrict__ exz_n) {
    const int Ncx = Nx - 1, Ncz = Nz - 1;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= Ncx * Ncz) return;
    const int k = c % Ncx, l = c / Ncx, v = k + l * Nx;
    double s = 0.0;
    if (tp[c] != 30 && tp[c] != 31) {
        if (tg[v] != 30) s += 0.25 * exz[v];
        if (tg[v + 1] != 30) s += 0.25 * exz[v + 1];
        if (tg[v + Nx] != 30) s += 0.25 * exz[v + Nx];
        if (tg[v + Nx + 1] != 30) s += 0.25 * exz[v + Nx + 1];
    }
    exz_n[c] = s;
}

// Output buffer: the caller's device array, or a scratch buffer copied back.
int out_dev(mdg_ctx *c, double *user, size_t n, dbuf<double> &tmp, double **o) {
    if (!user) { *o = nullptr; return MDG_OK; }
    if (mdg_is_device_ptr(user)) { *o = user; return MDG_OK; }
    HIPCHK(c, tmp.reserve(n));
    *o = tmp.p;
    return MDG_OK;
}

}  // namespace

struct rheo_scratch {
    dbuf<double> in[14], out[7];
    dbuf<mdg_phase_prop> ph;
};

namespace {

rheo_scratch *scratch_of(mdg_ctx *c) {
    static_assert(sizeof(mdg_phase_prop) % 8 == 0, "phase table packing");
    if (!c->rheo) c->rheo = new rheo_scratch;
    return c->rheo;
}

}  // namespace

void mdg_rheo_release(mdg_ctx *c) {
    if (!c->rheo) return;
    for (auto &b : c->rheo->in) b.release();
    for (auto &b : c->rheo->out) b.release();
    c->rheo->ph.release();
    delete c->rheo;
    c->rheo = nullptr;
}

extern "C" int mdg_viscosity_grid(mdg_ctx *c, const mdg_phase_prop *phases, int nph,
                                  const mdg_rheo_params *prm, int centroid,
                                  const mdg_rheo_nodes *f) {
    if (!c || !phases || !prm || !f) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (nph < 1 || nph > MDG_MAX_PHASES) return mdg_fail(c, MDG_ERR_ARG, "nph = %d outside [1, %d]", nph, MDG_MAX_PHASES);
    if (!f->perc || !f->G || !f->T || !f->P || !f->exx || !f->ezz || !f->exz || !f->eta)
        return mdg_fail(c, MDG_ERR_ARG, "perc, G, T, P, exx, ezz, exz and eta are required");
    if (prm->average < 0 || prm->average > 2) return mdg_fail(c, MDG_ERR_ARG, "average must be 0, 1 or 2");
    for (int p = 0; p < nph; p++) {
        const mdg_phase_prop &m = phases[p];
        if (m.plast != 0 && m.plast != 1)
            return mdg_fail(c, MDG_ERR_UNSUPPORTED, "phase %d: plast = %d (only 0 and 1 are implemented)", p, m.plast);
        if (!m.cstv && !m.pwlv && !prm->elastic)
            return mdg_fail(c, MDG_ERR_UNSUPPORTED, "phase %d has no creep law", p);
    }
    const int n = centroid ? (c->Nx - 1) * (c->Nz - 1) : c->Nx * c->Nz;
    const signed char *tag = centroid ? c->BCp_t.p : c->BCg_t.p;
    rheo_scratch *s = scratch_of(c);
    int rc;
    mdg_rheo_nodes d = {};
    const double *const *ins[14] = {&f->perc, &f->G, &f->T, &f->P, &f->exx, &f->ezz, &f->exz,
                                    &f->sxx0, &f->szz0, &f->sxz0, &f->fric, &f->C, &f->dil, &f->pref};
    const double **din[14] = {&d.perc, &d.G, &d.T, &d.P, &d.exx, &d.ezz, &d.exz,
                              &d.sxx0, &d.szz0, &d.sxz0, &d.fric, &d.C, &d.dil, &d.pref};
    for (int q = 0; q < 14; q++)
        if ((rc = mdg_stage_in(c, *ins[q], (size_t)n * (q == 0 ? nph : 1), s->in[q], din[q]))) return rc;
    double *const outs[7] = {f->eta, f->eta_phys, f->sxx, f->szz, f->sxz, f->p_corr, f->eII_pl};
    double **dout[7] = {&d.eta, &d.eta_phys, &d.sxx, &d.szz, &d.sxz, &d.p_corr, &d.eII_pl};
    for (int q = 0; q < 7; q++)
        if ((rc = out_dev(c, outs[q], (size_t)n, s->out[q], dout[q]))) return rc;
    HIPCHK(c, s->ph.reserve(nph));
    HIPCHK(c, hipMemcpyAsync(s->ph.p, phases, nph * sizeof(mdg_phase_prop), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, c->flag_dev.reserve(4));
    HIPCHK(c, hipMemsetAsync(c->flag_dev.p, 0, sizeof(int), c->stream));
    rheo_ctl md = {prm->step, prm->elastic, prm->average, nph, prm->dt, prm->R, prm->zeroC,
                   prm->min_eta, prm->max_eta, prm->eta_tol, prm->Tiimin, prm->scale_E, prm->scale_eta};
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_viscosity_grid, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, md,
                       (const mdg_phase_prop *)s->ph.p, n, tag, centroid, d, c->flag_dev.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    for (int q = 0; q < 7; q++)
        if (outs[q] && *dout[q] != outs[q])
            HIPCHK(c, hipMemcpyAsync(outs[q], *dout[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    int fail = 0;
    HIPCHK(c, hipMemcpyAsync(&fail, c->flag_dev.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (fail) return mdg_fail(c, MDG_ERR_DIVERGED, "visco-elastic local iterations failed to converge");
    return MDG_OK;
}

extern "C" int mdg_c2v(mdg_ctx *c, int periodic_x, const double *cent, double *vert) {
    if (!c || !cent || !vert) return mdg_fail(c, MDG_ERR_ARG, "null array");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    const int nc = (c->Nx - 1) * (c->Nz - 1), nv = c->Nx * c->Nz;
    rheo_scratch *s = scratch_of(c);
    const double *dc;
    double *dv;
    int rc;
    if ((rc = mdg_stage_in(c, cent, (size_t)nc, s->in[0], &dc)) || (rc = out_dev(c, vert, nv, s->out[0], &dv)))
        return rc;
    hipLaunchKernelGGL(k_c2v, dim3(mdg_grid1d(nv, 256)), dim3(256), 0, c->stream, c->Nx, c->Nz,
                       periodic_x ? 1 : 0, dc, c->BCp_t.p, c->BCg_t.p, dv);
    HIPCHK(c, hipGetLastError());
    if (dv != vert) HIPCHK(c, hipMemcpyAsync(vert, dv, (size_t)nv * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_strain_rates_nodes(mdg_ctx *c, const double *exxd, const double *ezzd,
                                      const double *exz, const double *div, double *exz_n,
                                      double *exxd_s, double *ezzd_s, double *div_s) {
    if (!c) return MDG_ERR_ARG;
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    const int nc = (c->Nx - 1) * (c->Nz - 1), nv = c->Nx * c->Nz;
    rheo_scratch *s = scratch_of(c);
    int rc;
    if (exz_n) {
        if (!exz) return mdg_fail(c, MDG_ERR_ARG, "exz_n needs exz");
        const double *de;
        double *dn;
        if ((rc = mdg_stage_in(c, exz, (size_t)nv, s->in[1], &de)) || (rc = out_dev(c, exz_n, nc, s->out[1], &dn)))
            return rc;
        hipLaunchKernelGGL(k_v2c_exz, dim3(mdg_grid1d(nc, 256)), dim3(256), 0, c->stream, c->Nx,
                           c->Nz, de, c->BCp_t.p, c->BCg_t.p, dn);
        HIPCHK(c, hipGetLastError());
        if (dn != exz_n) HIPCHK(c, hipMemcpyAsync(exz_n, dn, (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    const double *srcs[3] = {exxd, ezzd, div};
    double *dsts[3] = {exxd_s, ezzd_s, div_s};
    for (int q = 0; q < 3; q++) {
        if (!dsts[q]) continue;
        if (!srcs[q]) return mdg_fail(c, MDG_ERR_ARG, "vertex output %d needs its centre input", q);
        if ((rc = mdg_c2v(c, 0, srcs[q], dsts[q]))) return rc;
    }
    return MDG_OK;
}
