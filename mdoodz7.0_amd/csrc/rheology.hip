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

#include "rheo_device.h"

using rheo::rheo_ctl;
using rheo::viscosity_concise;

namespace {

template <bool EXT>
__global__ void __launch_bounds__(256)
k_viscosity_grid(rheo_ctl md, const mdg_phase_prop *__restrict__ phases, int n,
                 const signed char *__restrict__ tag, int centroid, mdg_rheo_nodes f,
                 int *__restrict__ fail) {
    extern __shared__ double sph_raw[];     // nph phase records
    mdg_phase_prop *sph = reinterpret_cast<mdg_phase_prop *>(sph_raw);
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
    const double dg = f.d ? f.d[i] : 0.0, phi = f.phi ? f.phi[i] : 0.0;
    double e = 0.0, ep = 0.0, pc = 0.0, epl = 0.0;
    int bad = 0;
    for (int p = 0; p < md.nph; p++) {
        const double w = f.perc[(size_t)p * n + i];
        if (!(fabs(w) > 1e-13)) continue;                              // :1212-1215
        const mdg_phase_prop &m = sph[p];
        double etaVE, Pcorr, Eii_pl;
        const double eta = viscosity_concise<EXT>(m, md, G, T, P, dg, phi, Exx, Ezz, Exz,
                                             f.fric ? f.fric[i] : m.fric, f.C ? f.C[i] : m.C,
                                             f.dil ? f.dil[i] : m.dil, pref, etaVE, Pcorr,
                                             Eii_pl, bad);
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
    if (bad) atomicOr(fail, bad);
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

// Softening (RheologyDensity.c:1432-1535) of one phase at accumulated
// plastic strain e: style 0 piecewise linear (slope times e, as the
// reference), 1 erfc, 2 linear hardening of the cohesion.
__device__ void soften(const mdg_phase_soft &m, double e, int style, double &fric, double &dil, double &C) {
    fric = m.fric; dil = m.dil; C = m.C;
    const double ds = m.pls_end - m.pls_start;
    if (style == 1) {
        const double mu = 0.5 * (m.pls_end + m.pls_start);
        if (m.fric_soft == 1) {
            const double d = m.fric - m.fric_end;
            const double f = m.fric - d / 2.0 * erfc(-(e - mu) / ds) - m.fric_end;
            const double f0 = m.fric - d / 2.0 * erfc(-(0.0 - mu) / ds) - m.fric_end;
            fric = f * d / f0 + m.fric_end;
        }
        if (m.dil_soft == 1) {
            const double d = m.dil - m.dil_end;
            const double f = m.dil - d / 2.0 * erfc(-(e - mu) / ds) - m.dil_end;
            const double f0 = m.dil - d / 2.0 * erfc(-(0.0 - mu) / ds) - m.dil_end;
            dil = f * d / f0 + m.dil_end;
        }
        if (m.coh_soft == 1) {
            const double d = m.C - m.C_end;
            const double f = m.C - d / 2.0 * erfc(-(e - mu) / ds) - m.C_end;
            const double f0 = m.C - d / 2.0 * erfc(-(0.0 - mu) / ds) - m.C_end;
            C = f * d / f0 + m.C_end;
        }
    } else if (style == 2) {
        C = fmax(m.C + m.H_c * e, m.C_end);
    } else {
        if (e < m.pls_start) {
            // initial values (already set)
        } else if (e >= m.pls_end) {
            if (m.fric_soft == 1) fric = m.fric_end;
            if (m.dil_soft == 1) dil = m.dil_end;
            if (m.coh_soft == 1) C = m.C_end;
        } else {
            if (m.fric_soft == 1) fric = m.fric + (m.fric_end - m.fric) / ds * e;
            if (m.dil_soft == 1) dil = m.dil + (m.dil_end - m.dil) / ds * e;
            if (m.coh_soft == 1) C = m.C + (m.C_end - m.C) / ds * e;
        }
    }
}

// ShearModCompExpGrid (RheologyDensity.c:1634-1730: harmonic G and beta)
// and CohesionFrictionDilationGrid (:1539-1616: arithmetic average of the
// softened friction, dilation and cohesion) on one node set.
__global__ void __launch_bounds__(256)
k_material_grid(int n, int nph, int style, int centroid, const mdg_phase_soft *__restrict__ phases,
                const signed char *__restrict__ tag, const double *__restrict__ perc,
                const double *__restrict__ epl, double *__restrict__ mu, double *__restrict__ bet,
                double *__restrict__ C, double *__restrict__ fric, double *__restrict__ dil) {
    __shared__ mdg_phase_soft sph[MDG_MAX_PHASES];
    for (int p = threadIdx.x; p < nph; p += blockDim.x) sph[p] = phases[p];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const signed char t = tag[i];
    double a = 0.0, b = 0.0, cc = 0.0, ff = 0.0, dd = 0.0;
    if (!(t == 30 || (centroid && t == 31))) {
        for (int p = 0; p < nph; p++) {
            const double w = perc[(size_t)p * n + i];
            a += w * 1.0 / sph[p].G;
            b += w * 1.0 / sph[p].beta;
            if (epl) {
                double f, d, c;
                soften(sph[p], epl[i], style, f, d, c);
                ff += w * f; dd += w * d; cc += w * c;
            }
        }
        a = 1.0 / a;
        b = 1.0 / b;
    }
    if (mu) mu[i] = a;
    if (bet) bet[i] = b;
    if (epl) {
        if (C) C[i] = cc;
        if (fric) fric[i] = ff;
        if (dil) dil[i] = dd;
    }
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
    dbuf<double> in[16], out[7];
    dbuf<mdg_phase_prop> ph;
    dbuf<mdg_phase_soft> soft;
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
    c->rheo->soft.release();
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
        if (!m.cstv && !m.pwlv && !m.linv && !m.expv && !prm->elastic)
            return mdg_fail(c, MDG_ERR_UNSUPPORTED, "phase %d has no creep law", p);
    }
    const int n = centroid ? (c->Nx - 1) * (c->Nz - 1) : c->Nx * c->Nz;
    const signed char *tag = centroid ? c->BCp_t.p : c->BCg_t.p;
    rheo_scratch *s = scratch_of(c);
    int rc;
    mdg_rheo_nodes d = {};
    const double *const *ins[16] = {&f->perc, &f->G, &f->T, &f->P, &f->exx, &f->ezz, &f->exz, &f->sxx0,
                                    &f->szz0, &f->sxz0, &f->fric, &f->C, &f->dil, &f->pref, &f->d, &f->phi};
    const double **din[16] = {&d.perc, &d.G, &d.T, &d.P, &d.exx, &d.ezz, &d.exz, &d.sxx0,
                              &d.szz0, &d.sxz0, &d.fric, &d.C, &d.dil, &d.pref, &d.d, &d.phi};
    for (int q = 0; q < 16; q++)
        if ((rc = mdg_stage_in(c, *ins[q], (size_t)n * (q == 0 ? nph : 1), s->in[q], din[q]))) return rc;
    double *const outs[7] = {f->eta, f->eta_phys, f->sxx, f->szz, f->sxz, f->p_corr, f->eII_pl};
    double **dout[7] = {&d.eta, &d.eta_phys, &d.sxx, &d.szz, &d.sxz, &d.p_corr, &d.eII_pl};
    for (int q = 0; q < 7; q++)
        if ((rc = out_dev(c, outs[q], (size_t)n, s->out[q], dout[q]))) return rc;
    HIPCHK(c, s->ph.reserve(nph));
    HIPCHK(c, hipMemcpyAsync(s->ph.p, phases, nph * sizeof(mdg_phase_prop), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, c->flag_dev.reserve(4));
    HIPCHK(c, hipMemsetAsync(c->flag_dev.p, 0, sizeof(int), c->stream));
    const rheo_ctl md = rheo::make_ctl(prm, nph);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(rheo::needs_ext(phases, nph) ? k_viscosity_grid<true> : k_viscosity_grid<false>,
                       dim3(mdg_grid1d(n, 256)), dim3(256), nph * sizeof(mdg_phase_prop),
                       c->stream, md,
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
    return rheo::fail_status(c, fail);
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

extern "C" int mdg_material_grid(mdg_ctx *c, const mdg_phase_soft *phases, int nph, int centroid,
                                 int style, const double *perc, const double *strain_pl,
                                 double *mu, double *bet, double *C, double *fric, double *dil) {
    if (!c || !phases || !perc) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (nph < 1 || nph > MDG_MAX_PHASES) return mdg_fail(c, MDG_ERR_ARG, "nph = %d outside [1, %d]", nph, MDG_MAX_PHASES);
    if (style < 0 || style > 2) return mdg_fail(c, MDG_ERR_ARG, "softening style must be 0, 1 or 2");
    const int n = centroid ? (c->Nx - 1) * (c->Nz - 1) : c->Nx * c->Nz;
    rheo_scratch *s = scratch_of(c);
    const double *dperc, *depl;
    int rc;
    if ((rc = mdg_stage_in(c, perc, (size_t)n * nph, s->in[0], &dperc)) ||
        (rc = mdg_stage_in(c, strain_pl, (size_t)n, s->in[1], &depl)))
        return rc;
    double *const outs[5] = {mu, bet, C, fric, dil};
    double *d[5];
    for (int q = 0; q < 5; q++)
        if ((rc = out_dev(c, outs[q], (size_t)n, s->out[q], &d[q]))) return rc;
    HIPCHK(c, s->soft.reserve(nph));
    HIPCHK(c, hipMemcpyAsync(s->soft.p, phases, nph * sizeof(mdg_phase_soft), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_material_grid, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n, nph, style,
                       centroid ? 1 : 0, (const mdg_phase_soft *)s->soft.p, centroid ? c->BCp_t.p : c->BCg_t.p,
                       dperc, depl, d[0], d[1], d[2], d[3], d[4]);
    HIPCHK(c, hipGetLastError());
    for (int q = 0; q < 5; q++)
        if (outs[q] && d[q] != outs[q])
            HIPCHK(c, hipMemcpyAsync(outs[q], d[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}
