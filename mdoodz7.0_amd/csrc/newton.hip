// Newton linearisation: tangent coefficients and the matrix-free Jacobian.
//
// Tangent (RheologicalOperators with Jacobian = 1, StokesRoutines.c:250-375,
// isotropic and incompressible): the derivatives of the phase-averaged
// effective viscosity w.r.t. Exx, Ezz, gxz = 2 Exz and P are central finite
// differences of ViscosityConcise (DerivativesOnTheFly_n/_s and
// PhaseRheologyLoop_v1, FD_Jacobian.c:48-323), eight extra evaluations per
// active phase and node; then
//   D11 = 2 eta + 2 eta_xx Exx, D12 = 2 eta_zz Exx, D13 = 2 eta_g Exx, D14 = 2 eta_p Exx
//   D21 = 2 eta_xx Ezz, D22 = 2 eta + 2 eta_zz Ezz, D23 = 2 eta_g Ezz, D24 = 2 eta_p Ezz
//   D31 = 2 eta_xx Exz, D32 = 2 eta_zz Exz, D33 = eta + 2 eta_g Exz, D34 = 2 eta_p Exz.
//
// Jacobian-vector product: the rows of the reference's assembled Newton
// operator (BuildJacobianOperatorDecoupled, StokesAssemblyDecoupled.c:
// 1786-2304), applied without assembling it (k_jac_rows below).
#include "mdg_internal.h"
#include "rheo_device.h"
#include "stokes_internal.h"

using rheo::rheo_ctl;
using rheo::viscosity_concise;

namespace {

template <bool EXT>
__global__ void __launch_bounds__(256)
k_tangent_grid(rheo_ctl md, const mdg_phase_prop *__restrict__ phases, int n,
               const signed char *__restrict__ tag, int centroid, mdg_rheo_nodes f,
               double *__restrict__ D, int *__restrict__ fail) {
    extern __shared__ double sph_raw[];     // nph phase records
    mdg_phase_prop *sph = reinterpret_cast<mdg_phase_prop *>(sph_raw);
    for (int p = threadIdx.x; p < md.nph; p += blockDim.x) sph[p] = phases[p];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nrow = centroid ? 8 : 4;
    const signed char t = tag[i];
    if (t == 30 || (centroid && t == 31)) {
        for (int r = 0; r < nrow; r++) D[(size_t)r * n + i] = 0.0;
        return;
    }
    const double el = md.elastic ? 1.0 : 0.0;
    const double G = f.G[i];
    const double eta_e = md.elastic ? md.dt * G : 1.0;
    // EffectiveStrainRate (RheologyDensity.c:116-146), isotropic
    const double Exx = f.exx[i] + (f.sxx0 ? el * f.sxx0[i] / eta_e / 2.0 : 0.0);
    const double Ezz = f.ezz[i] + (f.szz0 ? el * f.szz0[i] / eta_e / 2.0 : 0.0);
    const double Exz = f.exz[i] + (f.sxz0 ? el * f.sxz0[i] / eta_e / 2.0 : 0.0);
    const double T = f.T[i], P = f.P[i];
    const double pref = f.pref ? f.pref[i] : 1.0;
    const double tol = centroid ? 1e-5 : 1e-7;
    const double px = tol * fabs(Exx) + tol / 1e10, pz = tol * fabs(Ezz) + tol / 1e10;
    const double pg = tol * fabs(Exz) + tol / 1e10, pp = tol * fabs(P) + tol / 1e10;
    const double dg = f.d ? f.d[i] : 0.0, phi = f.phi ? f.phi[i] : 0.0;
    int bad = 0;
    double e = 0.0;
    for (int p = 0; p < md.nph; p++) {
        const double w = f.perc[(size_t)p * n + i];
        if (!(fabs(w) > 1e-13)) continue;
        const mdg_phase_prop &m = sph[p];
        double etaVE, Pc, Epl;
        viscosity_concise<EXT>(m, md, G, T, P, dg, phi, Exx, Ezz, Exz, f.fric ? f.fric[i] : m.fric,
                          f.C ? f.C[i] : m.C, f.dil ? f.dil[i] : m.dil, pref, etaVE, Pc, Epl, bad);
        if (md.average == 0) e += w * etaVE;
        else if (md.average == 1) e += w / etaVE;
        else e += w * log(etaVE);
    }
    if (md.average == 1) e = 1.0 / e;
    else if (md.average == 2) e = exp(e);
    // Central differences per phase: each phase's +h and -h values are
    // differenced before the 1/(2h) scaling (the reference sums the scaled
    // +h and -h terms of all phases, FD_Jacobian.c:66-78, leaving round-off
    // of size ulp(eta)/h in two-phase cells: O(1) at the h = 1e-15 floor of
    // a zero strain-rate component).  Same central difference, no cancellation.
    double der[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < 4; q++) {
        const double h = (q == 0) ? px : (q == 1) ? pz : (q == 2) ? pg : pp;
        const double den = (q == 2) ? 4.0 * h : 2.0 * h;
        for (int p = 0; p < md.nph; p++) {
            const double w = f.perc[(size_t)p * n + i];
            if (!(fabs(w) > 1e-13)) continue;
            const mdg_phase_prop &m = sph[p];
            const double fr = f.fric ? f.fric[i] : m.fric, co = f.C ? f.C[i] : m.C, di = f.dil ? f.dil[i] : m.dil;
            double eb = 0.0, Pc, Epl;
            if (md.average != 0)   // the phase's own value at the base state
                viscosity_concise<EXT>(m, md, G, T, P, dg, phi, Exx, Ezz, Exz, fr, co, di, pref, eb, Pc, Epl, bad);
            double epm[2];
            for (int sgn = 0; sgn < 2; sgn++) {
                const double d = sgn ? -h : h;
                viscosity_concise<EXT>(m, md, G, T, P + (q == 3 ? d : 0.0), dg, phi, Exx + (q == 0 ? d : 0.0),
                                  Ezz + (q == 1 ? d : 0.0), Exz + (q == 2 ? d : 0.0), fr, co, di, pref, epm[sgn],
                                  Pc, Epl, bad);
            }
            double v = w * (epm[0] - epm[1]) / den;
            if (md.average == 1) v /= eb * eb;
            else if (md.average == 2) v /= eb;
            der[q] += v;
        }
    }
    for (int q = 0; q < 4; q++) {
        if (md.average == 1) der[q] *= e * e;
        else if (md.average == 2) der[q] *= e;
    }
    if (centroid) {
        D[0 * (size_t)n + i] = 2.0 * e + 2.0 * der[0] * Exx;
        D[1 * (size_t)n + i] = 2.0 * der[1] * Exx;
        D[2 * (size_t)n + i] = 2.0 * der[2] * Exx;
        D[3 * (size_t)n + i] = 2.0 * der[3] * Exx;
        D[4 * (size_t)n + i] = 2.0 * der[0] * Ezz;
        D[5 * (size_t)n + i] = 2.0 * e + 2.0 * der[1] * Ezz;
        D[6 * (size_t)n + i] = 2.0 * der[2] * Ezz;
        D[7 * (size_t)n + i] = 2.0 * der[3] * Ezz;
    } else {
        D[0 * (size_t)n + i] = 2.0 * der[0] * Exz;
        D[1 * (size_t)n + i] = 2.0 * der[1] * Exz;
        D[2 * (size_t)n + i] = e + 2.0 * der[2] * Exz;
        D[3 * (size_t)n + i] = 2.0 * der[3] * Exz;
    }
    if (bad) atomicOr(fail, bad);
}

// ---------------------------------------------------------------------------
// The Jacobian rows of BuildJacobianOperatorDecoupled (Xjacobian /
// Zjacobian_InnerNodesDecoupled3, StokesAssemblyDecoupled.c:219-1085), applied
// matrix-free: one thread per row computes the row's coefficients from the
// tangent D of the two centres and two vertices it touches, the node flags of
// the tags and the vertex-averaging weights, and sums them against x.  The
// coefficients are written in terms of the tangent stresses' responses to the
// deviatoric normal rates exx_d = exx - co (exx + ezz)/3 (co = 1 in the
// reference's assembly, cm = co - 3):
//   centres  a = (cm D11 + co D12)/3, b = (co D11 + cm D12)/3 (x rows),
//            a2 = (co D21 + cm D22)/3, b2 = (cm D21 + co D22)/3 (z rows);
//   vertices s = (cm D31 + co D32)/3, t = (co D31 + cm D32)/3.
// Non-free entries of x are zero, so a coefficient of a Dirichlet, ghost or
// wall neighbour contributes nothing (AddCoeff3 sends them to the right-hand
// side); a no-slip ghost's coefficient is first folded into its interior
// neighbour's, as the reference does.
struct jac_tags {
    const signed char *u, *v, *p, *g;   // BCu, BCv, BCp, BCg types
};

__device__ inline double fl(bool c) { return c ? 1.0 : 0.0; }

__device__ double jac_x_row(const sgrid &g, const jac_tags &t, const double *__restrict__ Dn,
                            const double *__restrict__ Ds, double co, const double *__restrict__ u,
                            const double *__restrict__ v, const double *__restrict__ p, int k, int l) {
    const int nx = g.Nx, nz = g.Nz, nxvz = nx + 1, ncx = nx - 1, nzvx = nz + 1;
    const size_t nc = (size_t)ncx * (nz - 1), nvt = (size_t)nx * nz;
    const double dx = g.dx, dz = g.dz, dx2 = dx * dx, dz2 = dz * dz, dxz = dx * dz, cm = co - 3.0;
    const int c1 = k + l * nx, c2 = (k - 1) + (l - 1) * ncx, c3 = k + l * nxvz;
    const int S = c1 - nx, N = c1 + nx;                 // Vx south / north
    const int vSWi = c3 - nxvz, vSEi = vSWi + 1, vNWi = c3, vNEi = c3 + 1;
    const int pW = c2, pE = c2 + 1;
    const int xyS = c1 - nx, xyN = c1;                  // vertices below / above
    // flags
    const double inW = fl(t.p[pW] == -1), inE = fl(t.p[pE] == -1);
    const double inS = fl(t.u[S] != 30 && t.u[S] != 13), inN = fl(t.u[N] != 30 && t.u[N] != 13);
    const double inSv = fl(t.g[xyS] != 30), inNv = fl(t.g[xyN] != 30);
    const double cSW = (l > 1) ? fl(t.p[pW - ncx] == -1) : 0.0, cSE = (l > 1) ? fl(t.p[pE - ncx] == -1) : 0.0;
    const double cNW = (l < nzvx - 2) ? fl(t.p[pW + ncx] == -1) : 0.0, cNE = (l < nzvx - 2) ? fl(t.p[pE + ncx] == -1) : 0.0;
    const double vSWv = fl(t.g[xyS - 1] != 30), vNWv = fl(t.g[xyN - 1] != 30);
    const double vSEv = fl(t.g[xyS + 1] != 30), vNEv = fl(t.g[xyN + 1] != 30);
    const double fSWW = fl(t.v[vSWi - 1] == -1), fNWW = fl(t.v[vNWi - 1] == -1);
    const double fSEE = fl(t.v[vSEi + 1] == -1), fNEE = fl(t.v[vNEi + 1] == -1);
    double wS = inW + inE + cSW + cSE, wN = inW + inE + cNW + cNE;
    if (wS > 1.0) wS = 1.0 / wS;
    if (wN > 1.0) wN = 1.0 / wN;
    // tangent responses
    const double d11W = Dn[pW], d11E = Dn[pE], d12W = Dn[nc + pW], d12E = Dn[nc + pE];
    const double d13W = Dn[2 * nc + pW], d13E = Dn[2 * nc + pE], d14W = Dn[3 * nc + pW], d14E = Dn[3 * nc + pE];
    const double d31S = Ds[xyS], d31N = Ds[xyN], d32S = Ds[nvt + xyS], d32N = Ds[nvt + xyN];
    const double d33S = Ds[2 * nvt + xyS], d33N = Ds[2 * nvt + xyN], d34S = Ds[3 * nvt + xyS], d34N = Ds[3 * nvt + xyN];
    const double aW = (cm * d11W + co * d12W) / 3.0, aE = (cm * d11E + co * d12E) / 3.0;
    const double bW = (co * d11W + cm * d12W) / 3.0, bE = (co * d11E + cm * d12E) / 3.0;
    const double sS = (cm * d31S + co * d32S) / 3.0, sN = (cm * d31N + co * d32N) / 3.0;
    const double tS = (co * d31S + cm * d32S) / 3.0, tN = (co * d31N + cm * d32N) / 3.0;
    const double qS = inS * inS * inSv, qN = inN * inN * inNv;
    const double sh = wS * sS - wN * sN;
    // velocity coefficients
    double cW = inW * aW / dx2 + inW * sh / dxz - 0.25 * d13W * (inN * vNWv - inS * vSWv) / dxz;
    double cC = (d33N * qN + d33S * qS) / dz2 + (inE - inW) * sh / dxz - 0.25 * (d13W - d13E) * (qN - qS) / dxz -
                (inE * aE + inW * aW) / dx2;
    double cE = inE * aE / dx2 - inE * sh / dxz + 0.25 * d13E * (inN * vNEv - inS * vSEv) / dxz;
    const double cS = -d33S * qS / dz2 + 0.25 * qS * (d13E - d13W) / dxz + wS * (cSE - cSW) * sS / dxz;
    const double cN = -d33N * qN / dz2 + 0.25 * qN * (d13W - d13E) / dxz - wN * (cNE - cNW) * sN / dxz;
    const double cSWu = (-0.25 * d13W * inS * vSWv + cSW * wS * sS) / dxz;
    const double cSEu = (0.25 * d13E * inS * vSEv - cSE * wS * sS) / dxz;
    const double cNWu = (0.25 * d13W * inN * vNWv - cNW * wN * sN) / dxz;
    const double cNEu = (-0.25 * d13E * inN * vNEv + cNE * wN * sN) / dxz;
    double gSW = -d33S * inS * inSv / dxz - (inW * wN * tN + (cSW - inW) * wS * tS) / dz2 + inW * bW / dxz +
                 0.25 * (d13E * inS * inSv - d13W * (inS * inSv - fSWW * vSWv)) / dx2;
    double gSE = d33S * inS * inSv / dxz + (-inE * wN * tN + (inE - cSE) * wS * tS) / dz2 - inE * bE / dxz -
                 0.25 * (d13E * (inS * inSv - fSEE * vSEv) - d13W * inS * inSv) / dx2;
    double gNW = d33N * inN * inNv / dxz - ((cNW - inW) * wN * tN + inW * wS * tS) / dz2 - inW * bW / dxz -
                 0.25 * (-d13E * inN * inNv + d13W * (inN * inNv - fNWW * vNWv)) / dx2;
    double gNE = -d33N * inN * inNv / dxz + ((inE - cNE) * wN * tN - inE * wS * tS) / dz2 + inE * bE / dxz +
                 0.25 * (-d13E * (inN * inNv - fNEE * vNEv) + d13W * inN * inNv) / dx2;
    const double gSWW = -0.25 * d13W * fSWW * vSWv / dx2, gSEE = -0.25 * d13E * fSEE * vSEv / dx2;
    const double gNWW = -0.25 * d13W * fNWW * vNWv / dx2, gNEE = -0.25 * d13E * fNEE * vNEv / dx2;
    const double gSSW = cSW * wS * tS / dz2, gSSE = cSE * wS * tS / dz2;
    const double gNNW = cNW * wN * tN / dz2, gNNE = cNE * wN * tN / dz2;
    const double psh = (d34S * inS * wS - d34N * inN * wN) / dz;
    const double qW = -inW / dx + inW * (d14W / dx + psh), qE = inE / dx + inE * (-d14E / dx + psh);
    // no-slip ghosts fold into the interior neighbour
    if (t.u[S] == 11) cC -= cS;
    if (t.u[N] == 11) cC -= cN;
    if (t.u[S - 1] == 11) cW -= cSWu;
    if (t.u[S + 1] == 11) cE -= cSEu;
    if (t.u[N - 1] == 11) cW -= cNWu;
    if (t.u[N + 1] == 11) cE -= cNEu;
    if (t.v[vNWi - 1] == 11) gNW -= gNWW;
    if (t.v[vNEi + 1] == 11) gNE -= gNEE;
    if (t.v[vSWi - 1] == 11) gSW -= gSWW;
    if (t.v[vSEi + 1] == 11) gSE -= gSEE;
    double y = cS * u[S] + cW * u[c1 - 1] + cC * u[c1] + cE * u[c1 + 1] + cN * u[N];
    y += cSWu * u[S - 1] + cSEu * u[S + 1] + cNWu * u[N - 1] + cNEu * u[N + 1];
    y += gSW * v[vSWi] + gSE * v[vSEi] + gNW * v[vNWi] + gNE * v[vNEi];
    if (k > 1 || t.v[vSWi - 1] == -1) y += gSWW * v[vSWi - 1];
    if (k < nx - 2 || t.v[vSEi + 1] == -1) y += gSEE * v[vSEi + 1];
    if (k > 1 || t.v[vNWi - 1] == -1) y += gNWW * v[vNWi - 1];
    if (k < nx - 2 || t.v[vNEi + 1] == -1) y += gNEE * v[vNEi + 1];
    if (l > 1) y += gSSW * v[vSWi - nxvz] + gSSE * v[vSEi - nxvz];
    if (l < nz - 1) y += gNNW * v[vNWi + nxvz] + gNNE * v[vNEi + nxvz];
    y += qW * p[pW] + qE * p[pE];
    if (l > 1) y += d34S * inS * wS / dz * (cSW * p[pW - ncx] + cSE * p[pE - ncx]);
    if (l < nz - 1) y -= d34N * inN * wN / dz * (cNW * p[pW + ncx] + cNE * p[pE + ncx]);
    return y * dxz;
}

__device__ double jac_z_row(const sgrid &g, const jac_tags &t, const double *__restrict__ Dn,
                            const double *__restrict__ Ds, double co, const double *__restrict__ u,
                            const double *__restrict__ v, const double *__restrict__ p, int k, int l) {
    const int nx = g.Nx, nz = g.Nz, nxvz = nx + 1, ncx = nx - 1;
    const size_t nc = (size_t)ncx * (nz - 1), nvt = (size_t)nx * nz;
    const double dx = g.dx, dz = g.dz, dx2 = dx * dx, dz2 = dz * dz, dxz = dx * dz, cm = co - 3.0;
    const int c1 = k + l * nx, c2 = (k - 1) + (l - 1) * ncx, c3 = k + l * nxvz;
    const int W = c3 - 1, E = c3 + 1, S = c3 - nxvz, N = c3 + nxvz;     // Vz neighbours
    const int uSWi = c1 - 1, uSEi = c1, uNWi = c1 + nx - 1, uNEi = c1 + nx;
    const int pS = c2, pN = c2 + ncx;
    const int xyW = c1 - 1, xyE = c1;                                      // vertices left / right
    const double inS = fl(t.p[pS] == -1), inN = fl(t.p[pN] == -1);
    const double inW = fl(t.v[W] != 30 && t.v[W] != 13), inE = fl(t.v[E] != 30 && t.v[E] != 13);
    const double inWv = fl(t.g[xyW] != 30), inEv = fl(t.g[xyE] != 30);
    double cSW = 0, cNW = 0, cSE = 0, cNE = 0;
    if (k > 1 || (k == 1 && t.v[W] == -1)) { cSW = fl(t.p[pS - 1] == -1); cNW = fl(t.p[pN - 1] == -1); }
    if (k < nxvz - 2 || (k == nxvz - 2 && t.v[E] == -1)) { cSE = fl(t.p[pS + 1] == -1); cNE = fl(t.p[pN + 1] == -1); }
    const double vSWv = (l > 0) ? fl(t.g[xyW - nx] != 30) : 0.0, vSEv = (l > 0) ? fl(t.g[xyE - nx] != 30) : 0.0;
    const double vNWv = (l < nz - 1) ? fl(t.g[xyW + nx] != 30) : 0.0, vNEv = (l < nz - 1) ? fl(t.g[xyE + nx] != 30) : 0.0;
    const double fSSW = fl(t.u[uSWi - nx] == -1), fSSE = fl(t.u[uSEi - nx] == -1);
    const double fNNW = fl(t.u[uNWi + nx] == -1), fNNE = fl(t.u[uNEi + nx] == -1);
    double wE = inN + inS + cNE + cSE, wW = inN + inS + cNW + cSW;
    if (wW > 1.0) wW = 1.0 / wW;
    if (wE > 1.0) wE = 1.0 / wE;
    const double d21S = Dn[4 * nc + pS], d21N = Dn[4 * nc + pN], d22S = Dn[5 * nc + pS], d22N = Dn[5 * nc + pN];
    const double d23S = Dn[6 * nc + pS], d23N = Dn[6 * nc + pN], d24S = Dn[7 * nc + pS], d24N = Dn[7 * nc + pN];
    const double d31W = Ds[xyW], d31E = Ds[xyE], d32W = Ds[nvt + xyW], d32E = Ds[nvt + xyE];
    const double d33W = Ds[2 * nvt + xyW], d33E = Ds[2 * nvt + xyE], d34W = Ds[3 * nvt + xyW], d34E = Ds[3 * nvt + xyE];
    const double a2S = (co * d21S + cm * d22S) / 3.0, a2N = (co * d21N + cm * d22N) / 3.0;
    const double b2S = (cm * d21S + co * d22S) / 3.0, b2N = (cm * d21N + co * d22N) / 3.0;
    const double sW = (cm * d31W + co * d32W) / 3.0, sE = (cm * d31E + co * d32E) / 3.0;
    const double tW = (co * d31W + cm * d32W) / 3.0, tE = (co * d31E + cm * d32E) / 3.0;
    const double qW = inW * inW * inWv, qE = inE * inE * inEv;
    const double sh = wW * tW - wE * tE;
    double gW = -d33W * qW / dx2 + 0.25 * qW * (d23N * inN - d23S * inS) / dxz + wW * (cNW - cSW) * tW / dxz;
    double gC = -(inN * a2N + inS * a2S) / dz2 - 0.25 * (d23S * inS - d23N * inN) * (qE - qW) / dxz +
                (inN - inS) * sh / dxz + (d33E * qE + d33W * qW) / dx2;
    double gE = -d33E * qE / dx2 + 0.25 * qE * (d23S * inS - d23N * inN) / dxz - wE * (cNE - cSE) * tE / dxz;
    const double gS = inS * (-0.25 * d23S * (inE * vSEv - vSWv * inW) / dxz + a2S / dz2 + sh / dxz);
    const double gN = inN * (0.25 * d23N * (inE * vNEv - vNWv * inW) / dxz + a2N / dz2 - sh / dxz);
    double hSW = 0.25 * (d23N * inN * inW * inWv + d23S * inS * (fSSW * vSWv - inW * inWv)) / dz2 + inS * b2S / dxz -
                 d33W * inW * inWv / dxz - (inS * wE * sE - (inS - cSW) * wW * sW) / dx2;
    double hSE = -0.25 * (-d23N * inE * inEv * inN + d23S * inS * (inE * inEv - vSEv * fSSE)) / dz2 - inS * b2S / dxz +
                 d33E * inE * inEv / dxz + ((inS - cSE) * wE * sE - inS * wW * sW) / dx2;
    double hNW = 0.25 * (d23N * inN * (fNNW * vNWv - inW * inWv) + d23S * inS * inW * inWv) / dz2 - inN * b2N / dxz +
                 d33W * inW * inWv / dxz + (-inN * wE * sE + (inN - cNW) * wW * sW) / dx2;
    double hNE = 0.25 * (-d23N * inN * (inE * inEv - vNEv * fNNE) + d23S * inE * inEv * inS) / dz2 + inN * b2N / dxz -
                 d33E * inE * inEv / dxz + ((inN - cNE) * wE * sE - inN * wW * sW) / dx2;
    const double gSWv = (-0.25 * d23S * inS * vSWv * inW + cSW * wW * tW) / dxz;
    const double gSEv = (0.25 * d23S * inE * inS * vSEv - cSE * wE * tE) / dxz;
    const double gNWv = (0.25 * d23N * inN * vNWv * inW - cNW * wW * tW) / dxz;
    const double gNEv = (-0.25 * d23N * inE * inN * vNEv + cNE * wE * tE) / dxz;
    const double hSWW = cSW * wW * sW / dx2, hSEE = cSE * wE * sE / dx2;
    const double hNWW = cNW * wW * sW / dx2, hNEE = cNE * wE * sE / dx2;
    const double hSSW = -0.25 * d23S * inS * fSSW * vSWv / dz2, hSSE = -0.25 * d23S * inS * vSEv * fSSE / dz2;
    const double hNNW = -0.25 * d23N * inN * fNNW * vNWv / dz2, hNNE = -0.25 * d23N * inN * vNEv * fNNE / dz2;
    const double psh = (d34W * inW * wW - d34E * inE * wE) / dx;
    const double qS = -inS / dz + inS * (d24S / dz + psh), qN = inN / dz + inN * (-d24N / dz + psh);
    if (t.v[W] == 11) gC -= gW;
    if (t.v[E] == 11) gC -= gE;
    if (t.v[S - 1] == 11) gW -= gSWv;
    if (t.v[S + 1] == 11) gE -= gSEv;
    if (t.v[N - 1] == 11) gW -= gNWv;
    if (t.v[N + 1] == 11) gE -= gNEv;
    if (t.u[uNWi + nx] == 11) hNW -= hNNW;
    if (t.u[uNEi + nx] == 11) hNE -= hNNE;
    if (t.u[uSWi - nx] == 11) hSW -= hSSW;
    if (t.u[uSEi - nx] == 11) hSE -= hSSE;
    double y = gS * v[S] + gW * v[W] + gC * v[c3] + gE * v[E] + gN * v[N];
    y += gSWv * v[S - 1] + gSEv * v[S + 1] + gNWv * v[N - 1] + gNEv * v[N + 1];
    y += hSW * u[uSWi] + hSE * u[uSEi] + hNW * u[uNWi] + hNE * u[uNEi];
    y += hSSW * u[uSWi - nx] + hSSE * u[uSEi - nx] + hNNW * u[uNWi + nx] + hNNE * u[uNEi + nx];
    if (k > 1 || t.u[uSWi - 1] == -1) y += hSWW * u[uSWi - 1];
    if (k < nx - 1 || t.u[uSEi + 1] == -1) y += hSEE * u[uSEi + 1];
    if (k > 1 || t.u[uNWi - 1] == -1) y += hNWW * u[uNWi - 1];
    if (k < nx - 1 || t.u[uNEi + 1] == -1) y += hNEE * u[uNEi + 1];
    y += qS * p[pS] + qN * p[pN];
    if (k > 1) y += d34W * inW * wW / dx * (cSW * p[pS - 1] + cNW * p[pN - 1]);
    if (k < nx - 1) y -= d34E * inE * wE / dx * (cSE * p[pS + 1] + cNE * p[pN + 1]);
    return y * dxz;
}

// y = J_gamma x (grid layout): the Jacobian rows above on free velocity
// nodes (plus gamma * C'diag(eta_n/dV)C when gamma != 0), C x on the cells.
__global__ void __launch_bounds__(256)
k_jac_rows(sgrid g, jac_tags t, const double *__restrict__ Dn, const double *__restrict__ Ds, double co, double gamma,
           const double *__restrict__ x, const double *__restrict__ en, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Nz = g.Nz, Ncx = Nx - 1, nu = nvx(g), nv = nvz(g), nc = ncell(g), NxV = Nx + 1;
    const double *u = x, *v = x + nu, *p = x + nu + nv;
    const double dV = g.dx * g.dz;
    if (i < nu) {
        const int k = i % Nx, l = i / Nx;
        const signed char tt = t.u[i];
        double f = 0.0;
        if ((tt == -1 || tt == 2) && l > 0 && l < Nz && k > 0 && k < Nx - 1) {
            f = jac_x_row(g, t, Dn, Ds, co, u, v, p, k, l);
            if (gamma != 0.0) {
                const int cW = k - 1 + (l - 1) * Ncx, cE = cW + 1;
                const double divW = (u[i] - u[i - 1]) / g.dx + (v[k + l * NxV] - v[k + (l - 1) * NxV]) / g.dz;
                const double divE = (u[i + 1] - u[i]) / g.dx + (v[k + 1 + l * NxV] - v[k + 1 + (l - 1) * NxV]) / g.dz;
                f += gamma * (en[cW] * divW - en[cE] * divE) / g.dx * dV;
            }
        }
        y[i] = f;
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % NxV, l = j / NxV;
        const signed char tt = t.v[j];
        double f = 0.0;
        if ((tt == -1 || tt == 2) && k > 0 && k < NxV - 1 && l > 0 && l < Nz - 1) {
            f = jac_z_row(g, t, Dn, Ds, co, u, v, p, k, l);
            if (gamma != 0.0) {
                const int cS = k - 1 + (l - 1) * Ncx, cN = cS + Ncx;
                const double divS = (u[k + l * Nx] - u[k - 1 + l * Nx]) / g.dx + (v[j] - v[j - NxV]) / g.dz;
                const double divN = (u[k + (l + 1) * Nx] - u[k - 1 + (l + 1) * Nx]) / g.dx + (v[j + NxV] - v[j]) / g.dz;
                f += gamma * (en[cS] * divS - en[cN] * divN) / g.dz * dV;
            }
        }
        y[i] = f;
    } else if (i < nu + nv + nc) {
        const int c = i - nu - nv, k = c % Ncx, l = c / Ncx;
        y[i] = dV * ((u[k + 1 + (l + 1) * Nx] - u[k + (l + 1) * Nx]) / g.dx +
                     (v[k + 1 + (l + 1) * NxV] - v[k + 1 + l * NxV]) / g.dz);
    }
}

}  // namespace

// Device state of the Newton operator (coefficients and work arrays).
struct newton_state {
    dbuf<double> Dn, Ds;                 // 8 x cells, 4 x vertices
    dbuf<mdg_phase_prop> ph;
    bool active = false;
};

newton_state *newton_of(mdg_ctx *c) {
    if (!c->newton) c->newton = new newton_state;
    return c->newton;
}

void mdg_newton_release(mdg_ctx *c) {
    newton_state *s = c->newton;
    if (!s) return;
    s->Dn.release(); s->Ds.release(); s->ph.release();
    delete s;
    c->newton = nullptr;
}

// y = J_gamma x on grid-layout device vectors, with the coefficients staged by
// mdg_newton_set; gamma = 0 gives J itself.
int mdg_newton_apply_grid(mdg_ctx *c, double gamma, const double *x, double *y) {
    newton_state *s = c->newton;
    const sgrid g = c->sg;
    const int n = nvx(g) + nvz(g) + ncell(g);
    const jac_tags t{c->BCu_t.p, c->BCv_t.p, c->BCp_t.p, c->BCg_t.p};
    // comp = 1: BuildJacobianOperatorDecoupled assembles with the deviatoric
    // normal rates whatever model.compressible is (StokesAssemblyDecoupled.c:1803)
    hipLaunchKernelGGL(k_jac_rows, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, g, t, (const double *)s->Dn.p,
                       (const double *)s->Ds.p, 1.0, gamma, x, (const double *)c->eta_n.p, y);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

bool mdg_newton_active(const mdg_ctx *c) { return c->newton && c->newton->active; }

extern "C" int mdg_tangent_grid(mdg_ctx *c, const mdg_phase_prop *phases, int nph,
                                const mdg_rheo_params *prm, int centroid, const mdg_rheo_nodes *f,
                                double *D) {
    if (!c || !phases || !prm || !f || !D) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (nph < 1 || nph > MDG_MAX_PHASES) return mdg_fail(c, MDG_ERR_ARG, "nph = %d outside [1, %d]", nph, MDG_MAX_PHASES);
    if (!f->perc || !f->G || !f->T || !f->P || !f->exx || !f->ezz || !f->exz)
        return mdg_fail(c, MDG_ERR_ARG, "perc, G, T, P, exx, ezz and exz are required");
    for (int p = 0; p < nph; p++)
        if (phases[p].plast != 0 && phases[p].plast != 1)
            return mdg_fail(c, MDG_ERR_UNSUPPORTED, "phase %d: plast = %d (only 0 and 1 are implemented)", p, phases[p].plast);
    const int n = centroid ? (c->Nx - 1) * (c->Nz - 1) : c->Nx * c->Nz;
    const int nrow = centroid ? 8 : 4;
    newton_state *s = newton_of(c);
    // stage inputs through the rheology scratch of the grid update
    std::vector<dbuf<double>> tmp(16);
    mdg_rheo_nodes d = {};
    const double *const *ins[16] = {&f->perc, &f->G, &f->T, &f->P, &f->exx, &f->ezz, &f->exz, &f->sxx0,
                                    &f->szz0, &f->sxz0, &f->fric, &f->C, &f->dil, &f->pref, &f->d, &f->phi};
    const double **din[16] = {&d.perc, &d.G, &d.T, &d.P, &d.exx, &d.ezz, &d.exz, &d.sxx0,
                              &d.szz0, &d.sxz0, &d.fric, &d.C, &d.dil, &d.pref, &d.d, &d.phi};
    int rc = MDG_OK;
    for (int q = 0; q < 16 && !rc; q++)
        rc = mdg_stage_in(c, *ins[q], (size_t)n * (q == 0 ? nph : 1), tmp[q], din[q]);
    dbuf<double> dout;
    double *Dd = D;
    if (!rc && !mdg_is_device_ptr(D)) {
        if (dout.reserve((size_t)nrow * n) != hipSuccess) rc = mdg_fail(c, MDG_ERR_HIP, "allocation failed");
        Dd = dout.p;
    }
    if (!rc && s->ph.reserve(nph) != hipSuccess) rc = mdg_fail(c, MDG_ERR_HIP, "allocation failed");
    int fail = 0;
    if (!rc) {
        const rheo_ctl md = rheo::make_ctl(prm, nph);
        if (hipMemcpyAsync(s->ph.p, phases, nph * sizeof(mdg_phase_prop), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            c->flag_dev.reserve(4) != hipSuccess ||
            hipMemsetAsync(c->flag_dev.p, 0, sizeof(int), c->stream) != hipSuccess)
            rc = mdg_fail(c, MDG_ERR_HIP, "staging failed");
        else {
            hipLaunchKernelGGL(rheo::needs_ext(phases, nph) ? k_tangent_grid<true> : k_tangent_grid<false>,
                               dim3(mdg_grid1d(n, 256)), dim3(256), nph * sizeof(mdg_phase_prop),
                               c->stream, md,
                               (const mdg_phase_prop *)s->ph.p, n, centroid ? c->BCp_t.p : c->BCg_t.p,
                               centroid ? 1 : 0, d, Dd, c->flag_dev.p);
            if (hipGetLastError() != hipSuccess) rc = mdg_fail(c, MDG_ERR_HIP, "tangent kernel launch failed");
        }
        if (!rc && Dd != D &&
            hipMemcpyAsync(D, Dd, (size_t)nrow * n * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = mdg_fail(c, MDG_ERR_HIP, "copy back failed");
        if (!rc && (hipMemcpyAsync(&fail, c->flag_dev.p, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                    hipStreamSynchronize(c->stream) != hipSuccess))
            rc = mdg_fail(c, MDG_ERR_HIP, "synchronisation failed");
    }
    (void)hipStreamSynchronize(c->stream);
    for (auto &b : tmp) b.release();
    dout.release();
    if (rc) return rc;
    return rheo::fail_status(c, fail);
}

// Stage the Newton coefficients for mdg_jacobian_apply / the Newton solve.
static int newton_set(mdg_ctx *c, const double *Dn, const double *Ds) {
    newton_state *s = newton_of(c);
    const sgrid g = c->sg;
    const size_t nc = ncell(g), nvt = (size_t)g.Nx * g.Nz;
    HIPCHK(c, s->Dn.reserve(8 * nc));
    HIPCHK(c, s->Ds.reserve(4 * nvt));
    HIPCHK(c, hipMemcpyAsync(s->Dn.p, Dn, 8 * nc * 8, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->Ds.p, Ds, 4 * nvt * 8, hipMemcpyDefault, c->stream));
    return MDG_OK;
}

extern "C" int mdg_jacobian_apply(mdg_ctx *c, const double *Dn, const double *Ds, const double *x, double *y) {
    if (!c || !Dn || !Ds || !x || !y) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    int rc;
    if ((rc = newton_set(c, Dn, Ds)) || (rc = mdg_reserve_grid_vectors(c))) return rc;
    HIPCHK(c, c->eta_n.reserve(ncell(c->sg)));
    const sgrid g = c->sg;
    const double *xd;
    if ((rc = mdg_stage_in(c, x, (size_t)c->neq, c->tmp_c, &xd))) return rc;
    double *gxv = c->gx.p, *gyv = c->gy.p;
    if ((rc = mdg_eqn_to_grid(c, xd, gxv, gxv + nvx(g), gxv + nvx(g) + nvz(g), 1.0))) return rc;
    if ((rc = mdg_newton_apply_grid(c, 0.0, gxv, gyv))) return rc;
    double *yd = mdg_is_device_ptr(y) ? y : c->eqbuf.p;
    if ((rc = mdg_grid_to_eqn(c, gyv, gyv + nvx(g), gyv + nvx(g) + nvz(g), yd))) return rc;
    if (yd != y) HIPCHK(c, hipMemcpyAsync(y, yd, (size_t)c->neq * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_newton_solver(mdg_ctx *c, mdg_solver_opts *o, const double *Dn, const double *Ds,
                                 const double *D11_n, const double *D33_s, const double *rhs_mom,
                                 const double *rhs_cont, double *sol) {
    if (!c || !Dn || !Ds) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    int rc = newton_set(c, Dn, Ds);
    if (rc) return rc;
    c->newton->active = true;
    rc = mdg_killer_solver(c, o, D11_n, D11_n, D33_s, rhs_mom, rhs_cont, sol);
    c->newton->active = false;
    return rc;
}
