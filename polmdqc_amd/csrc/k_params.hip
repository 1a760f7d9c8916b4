// k_params.hip -- derivatives of the energy with respect to the parameters
// (FP64): local multipole components, polarizabilities, Thole damping.
//
// Runs after a GRADIENT evaluation and reuses its state: rotated multipoles,
// induced dipoles mu_d / mu_p, the tile masks, the reciprocal potentials.
//
// Multipoles.  Every pair energy is bilinear in the two multipoles, so the
// derivative of a pair's energy with respect to component c of M_i is the pair
// energy with M_i replaced by the unit multipole e_c; the same holds for the
// fields that build E^d and E^p.  With A = 1/alpha - T symmetric, mu_d =
// A^-1 E^d and mu_p = A^-1 E^p, the variational energy E_pol = -(f/2) mu_d.E^p
// has
//     dE_pol = -(f/2) [ mu_p . dE^d + mu_d . dE^p + mu_p . dT mu_d - mu_p . d(1/alpha) mu_d ],
// so mu_p serves as the adjoint of mu_d and no further solve is needed.  The
// reciprocal and self parts come from the potentials the evaluation already
// has (permanent: phi, grad phi, grad grad phi at each site; induced: the same
// of the two dipole grids).  The global derivatives are rotated back to each
// site's local frame (the chirality flip of chkpole included).
//
// Damping.  The Thole factors depend on D = a_ik (r / (pd_i pd_k))^3 with
// a_ik = min(a_i, a_k) (the larger when the smaller is zero); dlambda/dD is
// analytic, dR_n = (dlambda_n/dD) rr_n in the direct fields and in T, and dD
// is distributed to pd_i, pd_k (-3 D / pd) and to the Thole parameter of the
// smaller (half to each when equal, as a central difference sees it).
#include "pair_mpole.h"
#include "param_pair.h"
#include "pmq_internal.h"
#include "pme_common.h"

namespace pmq {

struct ParamArgs {
    int n, npad;
    const int2 *tiles;
    const int *ntiles;
    const unsigned long long *mask, *mmask;
    const double *sx, *sy, *sz, *smp, *spol;
    const int *perm, *iperm;
    DevBox box;
    double f, aewald;
    int do_mpole, do_polar;
    const double *u;          // 2 x 3 x npad sorted dipoles (d, p)
    const ExclPair *excl;
    int nexcl;
    double off2;
    double *dm, *dpd, *dth;   // sorted outputs (10 npad, npad, npad)
};

__device__ __forceinline__ Mp<double> mp_sorted(const double *smp, int np, int s)
{
    Mp<double> m;
    for (int c = 0; c < 10; c++) mp_set(m, c, smp[(size_t)c * np + s]);
    return m;
}

__global__ __launch_bounds__(64) void k_param_tiles(ParamArgs a)
{
    const int lane = threadIdx.x;
    const int np = a.npad;
    const int ntiles = *a.ntiles;
    __shared__ double jx[64], jy[64], jz[64], jm[10][64], ju[6][64], jpd[64], jth[64];
    __shared__ double acc[12][64]; // 10 multipole derivatives, dE/dpd, dE/dthole
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 IJ = a.tiles[t];
        const int si = IJ.x * 64 + lane, sj = IJ.y * 64 + lane;
        __syncthreads();
        jx[lane] = a.sx[sj]; jy[lane] = a.sy[sj]; jz[lane] = a.sz[sj];
        for (int c = 0; c < 10; c++) jm[c][lane] = a.smp[(size_t)c * np + sj];
        for (int c = 0; c < 6; c++) ju[c][lane] = a.do_polar ? a.u[(size_t)c * np + sj] : 0.0;
        jth[lane] = a.spol[np + sj];
        jpd[lane] = a.spol[2 * np + sj];
        for (int c = 0; c < 12; c++) acc[c][lane] = 0.0;
        const double xi = a.sx[si], yi = a.sy[si], zi = a.sz[si];
        const Mp<double> mi = mp_sorted(a.smp, np, si);
        double udi[3], upi[3];
        for (int c = 0; c < 3; c++) {
            udi[c] = a.do_polar ? a.u[(size_t)c * np + si] : 0.0;
            upi[c] = a.do_polar ? a.u[(size_t)(3 + c) * np + si] : 0.0;
        }
        const double thi = a.spol[np + si], pdi = a.spol[2 * np + si];
        __syncthreads();
        const unsigned long long mp_mask = a.mmask[(size_t)t * 64 + lane];
        unsigned long long mask = a.mask[(size_t)t * 128 + lane] | mp_mask;
        double gi[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, dpdi = 0.0, dthi = 0.0;
        while (mask) {
            const int jj = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            double xr = jx[jj] - xi, yr = jy[jj] - yi, zr = jz[jj] - zi;
            image<double>(a.box, xr, yr, zr);
            const double r2 = xr * xr + yr * yr + zr * zr;
            double Rm[6], R[6], rr[6], lam[6], dl[6], bn[6];
            rr_series<double>(r2, 1.0, rr);
            if (a.aewald > 0) ewald_series<double>(r2, a.aewald, bn);
            for (int q = 0; q < 6; q++) Rm[q] = a.f * (a.aewald > 0 ? bn[q] : rr[q]);
            double D = 0.0, pg = 0.0;
            if (a.do_polar) {
                thole_lambda<double>(sqrt(r2), pdi, jpd[jj], thi, jth[jj], lam);
                if (a.aewald > 0)
                    for (int q = 0; q < 6; q++) R[q] = bn[q] - (1.0 - lam[q]) * rr[q];
                else
                    for (int q = 0; q < 6; q++) R[q] = lam[q] * rr[q];
                thole_dlambda(sqrt(r2), pdi, jpd[jj], thi, jth[jj], dl, &D, &pg);
            } else {
                for (int q = 0; q < 6; q++) { R[q] = 0.0; dl[q] = 0.0; }
            }
            Mp<double> mk;
            for (int c = 0; c < 10; c++) mp_set(mk, c, jm[c][jj]);
            const double udk[3] = {ju[0][jj], ju[1][jj], ju[2][jj]}, upk[3] = {ju[3][jj], ju[4][jj], ju[5][jj]};
            double gk[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, dEdD = 0.0;
            const double sm = ((mp_mask >> jj) & 1ull) ? 1.0 : 0.0;
            const bool pol_pair = (a.mask[(size_t)t * 128 + lane] >> jj) & 1ull;
            param_pair<double, true>(a.do_mpole, a.do_polar, a.f, xr, yr, zr, mi, mk, Rm, R, rr, dl, sm,
                                     pol_pair ? 1.0 : 0.0, pol_pair ? 1.0 : 0.0, pol_pair ? 1.0 : 0.0, udi, upi, udk,
                                     upk, gi, gk, &dEdD);
            for (int c = 0; c < 10; c++) atomicAdd(&acc[c][jj], gk[c]);
            if (D > 0.0 && pol_pair) {
                dpdi += dEdD * (-3.0 * D / pdi);
                atomicAdd(&acc[10][jj], dEdD * (-3.0 * D / jpd[jj]));
                const double dDda = pg > 0.0 ? D / pg : 0.0;
                dthi += dEdD * dDda * thole_share(thi, jth[jj]);
                atomicAdd(&acc[11][jj], dEdD * dDda * thole_share(jth[jj], thi));
            }
        }
        __syncthreads();
        for (int c = 0; c < 10; c++) {
            atomicAdd(&a.dm[(size_t)c * np + si], gi[c]);
            atomicAdd(&a.dm[(size_t)c * np + sj], acc[c][lane]);
        }
        atomicAdd(&a.dpd[si], dpdi);
        atomicAdd(&a.dpd[sj], acc[10][lane]);
        atomicAdd(&a.dth[si], dthi);
        atomicAdd(&a.dth[sj], acc[11][lane]);
    }
}

// scaled pairs: (s - 1) times the undamped series for the multipoles, (s - 1)
// lambda rr for the polarization pieces
__global__ __launch_bounds__(64) void k_param_excl(ParamArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nexcl) return;
    const ExclPair q = a.excl[t];
    const int np = a.npad;
    const int si = a.iperm[q.i], sk = a.iperm[q.k];
    double xr = a.sx[sk] - a.sx[si], yr = a.sy[sk] - a.sy[si], zr = a.sz[sk] - a.sz[si];
    image<double>(a.box, xr, yr, zr);
    const double r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > a.off2) return;
    double Rm[6], R[6], rr[6], lam[6], dl[6];
    rr_series<double>(r2, 1.0, rr);
    for (int c = 0; c < 6; c++) Rm[c] = a.f * rr[c];
    double D = 0.0, pg = 0.0;
    const double pdi = a.spol[2 * np + si], pdk = a.spol[2 * np + sk], thi = a.spol[np + si], thk = a.spol[np + sk];
    if (a.do_polar) {
        thole_lambda<double>(sqrt(r2), pdi, pdk, thi, thk, lam);
        for (int c = 0; c < 6; c++) R[c] = lam[c] * rr[c];
        thole_dlambda(sqrt(r2), pdi, pdk, thi, thk, dl, &D, &pg);
    } else {
        for (int c = 0; c < 6; c++) { R[c] = 0.0; dl[c] = 0.0; }
    }
    const Mp<double> mi = mp_sorted(a.smp, np, si), mk = mp_sorted(a.smp, np, sk);
    double udi[3], upi[3], udk[3], upk[3];
    for (int c = 0; c < 3; c++) {
        udi[c] = a.do_polar ? a.u[(size_t)c * np + si] : 0.0; upi[c] = a.do_polar ? a.u[(size_t)(3 + c) * np + si] : 0.0;
        udk[c] = a.do_polar ? a.u[(size_t)c * np + sk] : 0.0; upk[c] = a.do_polar ? a.u[(size_t)(3 + c) * np + sk] : 0.0;
    }
    double gi[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, gk[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, dEdD = 0.0;
    param_pair<double, true>(a.do_mpole, a.do_polar, a.f, xr, yr, zr, mi, mk, Rm, R, rr, dl, q.m - 1.0, q.d - 1.0,
                             q.p - 1.0, q.u - 1.0, udi, upi, udk, upk, gi, gk, &dEdD);
    for (int c = 0; c < 10; c++) {
        atomicAdd(&a.dm[(size_t)c * np + si], gi[c]);
        atomicAdd(&a.dm[(size_t)c * np + sk], gk[c]);
    }
    if (D > 0.0) {
        atomicAdd(&a.dpd[si], dEdD * (-3.0 * D / pdi));
        atomicAdd(&a.dpd[sk], dEdD * (-3.0 * D / pdk));
        const double dDda = pg > 0.0 ? D / pg : 0.0;
        atomicAdd(&a.dth[si], dEdD * dDda * thole_share(thi, thk));
        atomicAdd(&a.dth[sk], dEdD * dDda * thole_share(thk, thi));
    }
}

struct SiteArgs {
    int n, npad;
    const int *perm;
    const double *smp, *spol;
    const double *cphi;       // 10 npad: reciprocal permanent potential (phi, grad, hessian xx yy zz xy xz yz)
    const double *frac;       // n x 2 x 20 fractional potentials of the dipole grids (d, p), or null
    const double *u;          // 2 x 3 x npad sorted dipoles
    double A[3][3];           // fractional = A r (grid units)
    double f, aewald;
    int ewald, do_mpole, do_polar;
    const double *dm, *dpd, *dth;
    const double *rotm;       // 9n caller order
    const int *flip;
    const double *polarity;   // caller order
    double *dpole, *dpolarity, *dpdamp, *dthole; // caller order outputs
};

// derivative of U = c phi + d.grad phi + q:H with respect to the stored
// multipole components (c, d, q xx xy xz yy yz zz; off-diagonals twice)
__device__ __forceinline__ void dU_dM(double phi, const double *g, const double *H, double s, double *out)
{
    out[0] += s * phi;
    for (int q = 0; q < 3; q++) out[1 + q] += s * g[q];
    out[4] += s * H[0]; out[5] += s * 2.0 * H[3]; out[6] += s * 2.0 * H[4];
    out[7] += s * H[1]; out[8] += s * 2.0 * H[5]; out[9] += s * H[2];
}

__global__ __launch_bounds__(128) void k_param_site(SiteArgs a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n) return;
    const int np = a.npad;
    const int i = a.perm[s];
    double g[10];
    for (int c = 0; c < 10; c++) g[c] = a.dm[(size_t)c * np + s];
    double m[10];
    for (int c = 0; c < 10; c++) m[c] = a.smp[(size_t)c * np + s];
    const double hf = 0.5 * a.f;
    const double ud[3] = {a.do_polar ? a.u[s] : 0.0, a.do_polar ? a.u[np + s] : 0.0, a.do_polar ? a.u[2 * np + s] : 0.0};
    const double up[3] = {a.do_polar ? a.u[3 * np + s] : 0.0, a.do_polar ? a.u[4 * np + s] : 0.0,
                          a.do_polar ? a.u[5 * np + s] : 0.0};
    if (a.ewald) {
        const double al = a.aewald, a2 = al * al, rpi = 1.77245385090551602730;
        if (a.do_mpole) {
            double H[6], gr[3];
            for (int q = 0; q < 3; q++) gr[q] = a.cphi[(size_t)(1 + q) * np + s];
            for (int q = 0; q < 6; q++) H[q] = a.cphi[(size_t)(4 + q) * np + s];
            dU_dM(a.cphi[s], gr, H, a.f, g);
            // self energy -f a/sqrt(pi) [c^2 + 2a^2/3 d.d + 8a^4/5 q:q]
            const double sf = -a.f * al / rpi;
            g[0] += sf * 2.0 * m[0];
            for (int q = 0; q < 3; q++) g[1 + q] += sf * 4.0 * a2 / 3.0 * m[1 + q];
            const double c4 = sf * 16.0 * a2 * a2 / 5.0;
            g[4] += c4 * m[4]; g[5] += c4 * 2.0 * m[5]; g[6] += c4 * 2.0 * m[6];
            g[7] += c4 * m[7]; g[8] += c4 * 2.0 * m[8]; g[9] += c4 * m[9];
        }
        if (a.do_polar && a.frac) {
            // reciprocal potentials of the two dipole grids, each + (f/2)
            for (int set = 0; set < 2; set++) {
                const double *fp = a.frac + ((size_t)s * 2 + set) * 20;
                double phi = fp[0], gr[3], H[6];
                for (int x = 0; x < 3; x++) gr[x] = a.A[0][x] * fp[1] + a.A[1][x] * fp[2] + a.A[2][x] * fp[3];
                const double F2[3][3] = {{fp[4], fp[7], fp[8]}, {fp[7], fp[5], fp[9]}, {fp[8], fp[9], fp[6]}};
                const int hx[6] = {0, 1, 2, 0, 0, 1}, hy[6] = {0, 1, 2, 1, 2, 2};
                for (int q = 0; q < 6; q++) {
                    double v = 0;
                    for (int p = 0; p < 3; p++)
                        for (int r = 0; r < 3; r++) v += a.A[p][hx[q]] * a.A[r][hy[q]] * F2[p][r];
                    H[q] = v;
                }
                dU_dM(phi, gr, H, hf, g);
            }
            // self field (4/3) a^3/sqrt(pi) d in E^d and E^p
            const double self = 4.0 / 3.0 * a2 * al / rpi;
            for (int q = 0; q < 3; q++) g[1 + q] -= hf * self * (ud[q] + up[q]);
        }
    }
    // back to the local frame: d_glob = R^T d_loc (rows x, y, z of R), q_glob = R^T q_loc R
    const double *R = a.rotm + 9 * (size_t)i;
    double out[13];
    out[0] = g[0];
    for (int j = 0; j < 3; j++) out[1 + j] = g[1] * R[3 * j + 0] + g[2] * R[3 * j + 1] + g[3] * R[3 * j + 2];
    // stored global components (h <= k) and their derivatives
    const int sh[6] = {0, 0, 0, 1, 1, 2}, sk[6] = {0, 1, 2, 1, 2, 2};
    const int slot[6] = {4, 5, 6, 7, 8, 9};
    for (int l = 0; l < 3; l++)
        for (int m2 = 0; m2 < 3; m2++) {
            double v = 0.0;
            for (int q = 0; q < 6; q++) v += g[slot[q]] * R[3 * l + sh[q]] * R[3 * m2 + sk[q]];
            out[4 + 3 * l + m2] = v;
        }
    if (a.flip && a.flip[i]) {
        out[2] = -out[2]; out[5] = -out[5]; out[7] = -out[7]; out[9] = -out[9]; out[11] = -out[11];
    }
    for (int c = 0; c < 13; c++) a.dpole[13 * (size_t)i + c] = out[c];
    if (a.dpolarity) {
        const double al = a.polarity[i];
        a.dpolarity[i] = (a.do_polar && al != 0.0)
                             ? -hf / (al * al) * (up[0] * ud[0] + up[1] * ud[1] + up[2] * ud[2])
                             : 0.0;
        a.dpdamp[i] = a.dpd[s];
        a.dthole[i] = a.dth[s];
    }
}

// the per-site step: reciprocal and self parts (Ewald), rotation of the
// global derivatives to the local frames, dE/dpolarity; outputs in caller
// order (13n, n, n, n)
int launch_param_site(pmq_ctx *c, bool mpole, bool polar, double *dpole, double *dpolarity, double *dpdamp,
                      double *dthole)
{
    const int np = c->npad, n = c->n;
    SiteArgs s;
    s.n = n; s.npad = np; s.perm = c->perm.p; s.smp = c->smp.p; s.spol = c->spol.p;
    s.cphi = c->cphi.p;
    s.frac = (polar && c->prm.use_ewald) ? c->pme_frac.p : nullptr;
    s.u = c->uind.p;
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) s.A[p][q] = c->pmeA[p][q];
    s.f = c->prm.electric / c->prm.dielec; s.aewald = c->prm.aewald; s.ewald = c->prm.use_ewald;
    s.do_mpole = mpole; s.do_polar = polar;
    s.dm = c->pg_dm.p; s.dpd = c->pg_dpd.p; s.dth = c->pg_dth.p;
    s.rotm = c->rotm.p; s.flip = c->flip.p;
    s.polarity = c->polarity.p;
    s.dpole = dpole;
    const bool pol_out = polar && !c->h_polarity.empty();
    s.dpolarity = pol_out ? dpolarity : nullptr; s.dpdamp = dpdamp; s.dthole = dthole;
    hipLaunchKernelGGL(k_param_site, dim3((n + 127) / 128), dim3(128), 0, c->stream, s);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int param_gradients(pmq_ctx *c, bool mpole, bool polar, double *dpole, double *dpolarity, double *dpdamp,
                    double *dthole)
{
    const int np = c->npad, n = c->n;
    PMQ_HIP(c->pg_dm.alloc(10 * (size_t)np));
    PMQ_HIP(c->pg_dpd.alloc(np));
    PMQ_HIP(c->pg_dth.alloc(np));
    int rc;
    if ((rc = launch_zero(c, c->pg_dm.p, 10 * (size_t)np))) return rc;
    if ((rc = launch_zero(c, c->pg_dpd.p, np))) return rc;
    if ((rc = launch_zero(c, c->pg_dth.p, np))) return rc;
    double cut = c->prm.use_ewald ? c->prm.ewaldcut : c->prm.mpolecut;
    ParamArgs a;
    a.n = n; a.npad = np;
    a.tiles = c->tiles.p; a.ntiles = c->ntiles_eff.p;
    a.mask = c->tmask.p; a.mmask = c->mmask.p;
    a.sx = c->sx.p; a.sy = c->sy.p; a.sz = c->sz.p; a.smp = c->smp.p; a.spol = c->spol.p;
    a.perm = c->perm.p; a.iperm = c->iperm.p;
    a.box = c->box;
    a.f = c->prm.electric / c->prm.dielec;
    a.aewald = c->prm.use_ewald ? c->prm.aewald : 0.0;
    a.do_mpole = mpole; a.do_polar = polar;
    a.u = c->uind.p;
    a.excl = c->excl.p;
    a.nexcl = (int)c->h_excl.size();
    a.off2 = cut * cut;
    a.dm = c->pg_dm.p; a.dpd = c->pg_dpd.p; a.dth = c->pg_dth.p;
    hipLaunchKernelGGL(k_param_tiles, dim3(c->tgrid), dim3(64), 0, c->stream, a);
    if (a.nexcl) hipLaunchKernelGGL(k_param_excl, dim3((a.nexcl + 63) / 64), dim3(64), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());

    DBuf<double> d_pole, d_pol, d_pd, d_th;
    PMQ_HIP(d_pole.alloc(13 * (size_t)n));
    PMQ_HIP(d_pol.alloc(n)); PMQ_HIP(d_pd.alloc(n)); PMQ_HIP(d_th.alloc(n));
    if ((rc = launch_param_site(c, mpole, polar, d_pole.p, d_pol.p, d_pd.p, d_th.p))) return rc;
    const bool pol_out = polar && !c->h_polarity.empty();
    PMQ_HIP(hipStreamSynchronize(c->stream));
    if (dpole) PMQ_HIP(hipMemcpy(dpole, d_pole.p, 13 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (pol_out) {
        if (dpolarity) PMQ_HIP(hipMemcpy(dpolarity, d_pol.p, n * sizeof(double), hipMemcpyDeviceToHost));
        if (dpdamp) PMQ_HIP(hipMemcpy(dpdamp, d_pd.p, n * sizeof(double), hipMemcpyDeviceToHost));
        if (dthole) PMQ_HIP(hipMemcpy(dthole, d_th.p, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        for (int i = 0; i < n; i++) {
            if (dpolarity) dpolarity[i] = 0.0;
            if (dpdamp) dpdamp[i] = 0.0;
            if (dthole) dthole[i] = 0.0;
        }
    }
    return PMQ_OK;
}

} // namespace pmq
