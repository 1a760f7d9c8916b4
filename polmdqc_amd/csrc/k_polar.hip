// k_polar.hip -- mutual induced-dipole polarization (AMOEBA with Thole
// damping; HIPPO with charge-penetration damping).
//
// The reference declares the state of this term but computes nothing
// (epolar is commented out at energy.cpp:267; induce/PCG defaults at
// kpolar.cpp:80-103, polar.h:35-53, polpot.h:47-66, polpcg.h:21-25).  The
// physics follows Ren and Ponder, J. Phys. Chem. B 107, 5933 (2003):
//
//   mu_d = alpha (E^d + T mu_d),   mu_p = alpha (E^p + T mu_p)
//   E_pol = -(f/2) sum_i mu_d,i . E^p_i
//
// E^d / E^p are the permanent fields with the d (polarization group) and p
// (bond connectivity) exclusion scales, T the dipole field operator with the
// u scales; all Thole damped and, in a periodic box, Ewald summed.  The
// gradient uses the variational form: at convergence it is the position
// derivative at fixed dipoles of
//   (1/2)[U_p(mu_d; M) + U_d(mu_p; M) + U_u(mu_d; mu_p) + U_u(mu_p; mu_d)]
// so no second solve is needed.
//
// HIPPO (charge penetration on; kpolar.cpp:577 turns Thole off then): the
// induced dipoles are points; the permanent field at i is that of the core
// charge of k (undamped) plus its valence multipole damped by the one-site
// factors of k's exponent; T is damped by the overlap factors of the two
// valence densities and scaled by w (induce-1x) instead of u.
//
// Kernels (sorted atom order, 2 dipole sets: 0 = d, 1 = p):
//   k_pol_field_tiles / _excl   permanent field (scale 1 in the tiles,
//                               (s-1) lambda rr corrections for scaled pairs)
//   (T mu for both sets is the pair-matrix product of k_pairs.hip)
//   k_pcg_*                     preconditioned CG: the start, and the step E2
//                               with the diagonal preconditioner (the sparse
//                               one and E1 run in k_pairs.hip / k_pme.hip;
//                               pcg_steps.h)
//   k_pol_grad_tiles / _excl    real-space gradient and torques
#include "pair_mpole.h"
#include "pcg_steps.h"
#include "pmq_internal.h"
#include "tile_pairs.h"

namespace pmq {

__device__ __forceinline__ double wave_sum_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct PolArgs {
    int n, npad;
    const int2 *tiles;
    const int *ntiles;
    TileItems items; // pieces of the tiles (tile_pairs.h)
    const double *sx, *sy, *sz, *smp, *spol;
    const double *scp;   // sorted core charge, valence exponent (charge penetration)
    int cpd;             // HIPPO damping (charge penetration) instead of Thole
    const int *sflag, *perm, *iperm;
    DevBox box;
    double off2, aewald, f;
    const ExclPair *excl;
    int nexcl;
    const unsigned long long *mask; // per tile: 64 row masks of in-cutoff pairs, then 64 column masks
    const double *src;   // dipoles (2 x 3 x npad) for the mutual field / gradient
    double *dst;         // fields (2 x 3 x npad)
    double *dep, *tep;   // caller-order gradient and torque
    long long *dstq;     // mixed precision: fixed-point fields (else null)
    void *part;          // per tile partial sums (2 sides x PMQ_TILE_NV x 64 of the pair arithmetic type)
    long long *depq, *tepq; // mixed precision: fixed-point gradient and torque (else null)
    double *potp;        // charge flux (gradient kernels): dE_pol/dq per atom, caller order (else null)
    long long *potpq;    // mixed precision: its fixed-point accumulator for the scaled pairs (else null)
    int vir;             // gradient kernels: also the pair virial (tile scalars 1..9; totals for scaled pairs)
    double *tscal;
    size_t tstride;
    Totals *tot;
};

__device__ __forceinline__ void field_add(double *d, long long *q, size_t i, double v)
{
    if (q) fx_add(q, i, v, PMQ_FX_FIELD);
    else atomicAdd(d + i, v);
}

// field at i from multipole k and at k from multipole i
template <class T>
__device__ __forceinline__ void mp_fields(const T *R, const Mp<T> &a, const Mp<T> &b, T xr, T yr, T zr, T *Ei, T *Ek)
{
    const T qix = a.qxx * xr + a.qxy * yr + a.qxz * zr, qiy = a.qxy * xr + a.qyy * yr + a.qyz * zr,
                 qiz = a.qxz * xr + a.qyz * yr + a.qzz * zr;
    const T qkx = b.qxx * xr + b.qxy * yr + b.qxz * zr, qky = b.qxy * xr + b.qyy * yr + b.qyz * zr,
            qkz = b.qxz * xr + b.qyz * yr + b.qzz * zr;
    const T dir = a.dx * xr + a.dy * yr + a.dz * zr, dkr = b.dx * xr + b.dy * yr + b.dz * zr;
    const T qir = qix * xr + qiy * yr + qiz * zr, qkr = qkx * xr + qky * yr + qkz * zr;
    const T si = b.c * R[1] - dkr * R[2] + qkr * R[3];
    const T sk = a.c * R[1] + dir * R[2] + qir * R[3];
    Ei[0] = -si * xr - R[1] * b.dx + T(2) * R[2] * qkx;
    Ei[1] = -si * yr - R[1] * b.dy + T(2) * R[2] * qky;
    Ei[2] = -si * zr - R[1] * b.dz + T(2) * R[2] * qkz;
    Ek[0] = sk * xr - R[1] * a.dx - T(2) * R[2] * qix;
    Ek[1] = sk * yr - R[1] * a.dy - T(2) * R[2] * qiy;
    Ek[2] = sk * zr - R[1] * a.dz - T(2) * R[2] * qiz;
}

// one side of mp_fields: the field at i of multipole b at k, and at k of a at i
template <class T>
__device__ __forceinline__ void field_at_i(const T *R, const Mp<T> &b, T xr, T yr, T zr, T *Ei)
{
    const T qkx = b.qxx * xr + b.qxy * yr + b.qxz * zr, qky = b.qxy * xr + b.qyy * yr + b.qyz * zr,
            qkz = b.qxz * xr + b.qyz * yr + b.qzz * zr;
    const T dkr = b.dx * xr + b.dy * yr + b.dz * zr, qkr = qkx * xr + qky * yr + qkz * zr;
    const T si = b.c * R[1] - dkr * R[2] + qkr * R[3];
    Ei[0] += -si * xr - R[1] * b.dx + T(2) * R[2] * qkx;
    Ei[1] += -si * yr - R[1] * b.dy + T(2) * R[2] * qky;
    Ei[2] += -si * zr - R[1] * b.dz + T(2) * R[2] * qkz;
}

template <class T>
__device__ __forceinline__ void field_at_k(const T *R, const Mp<T> &a, T xr, T yr, T zr, T *Ek)
{
    const T qix = a.qxx * xr + a.qxy * yr + a.qxz * zr, qiy = a.qxy * xr + a.qyy * yr + a.qyz * zr,
            qiz = a.qxz * xr + a.qyz * yr + a.qzz * zr;
    const T dir = a.dx * xr + a.dy * yr + a.dz * zr, qir = qix * xr + qiy * yr + qiz * zr;
    const T sk = a.c * R[1] + dir * R[2] + qir * R[3];
    Ek[0] += sk * xr - R[1] * a.dx - T(2) * R[2] * qix;
    Ek[1] += sk * yr - R[1] * a.dy - T(2) * R[2] * qiy;
    Ek[2] += sk * zr - R[1] * a.dz - T(2) * R[2] * qiz;
}

// the HIPPO series of one pair at scale s: the permanent field of site k at
// i is that of its core undamped plus its valence multipole damped by k's
// one-site factors, summed as [total multipole, Rt] + [valence multipole,
// Xk = -s (1 - dmpk) rr] so the core/valence cancellation never forms (Xi
// the same for i); Rm the dipole-dipole series (overlap factors).  With
// Ewald (ew) Rt = B - (1 - s) rr and Rm = B - (1 - s dmpik) rr, else s rr
// and s dmpik rr.
template <class T>
__device__ __forceinline__ void hippo_pol_series(T r2, T ai, T ak, bool ew, T aewald, T s, T *Rt, T *Xi, T *Xk, T *Rm)
{
    T rr[6], bn[6], xi[5], xk[5], xik[6];
    rr_series<T>(r2, T(1), rr);
    if (ew) ewald_series<T>(r2, aewald, bn);
    damp_gordon_comp<T, 1>(sqrt(r2), ai, ak, xi, xk, xik);
#pragma unroll
    for (int j = 0; j < 6; j++) {
        Rt[j] = ew ? bn[j] - (T(1) - s) * rr[j] : s * rr[j];
        Xi[j] = j < 5 ? -s * xi[j] * rr[j] : T(0);
        Xk[j] = j < 5 ? -s * xk[j] * rr[j] : T(0);
        Rm[j] = ew ? bn[j] - (T(1) - s + s * xik[j]) * rr[j] : s * (T(1) - xik[j]) * rr[j];
    }
}

template <class T = double>
__device__ __forceinline__ Mp<T> load_mp(const double *smp, int np, int s)
{
    Mp<T> m;
    m.c = smp[s]; m.dx = smp[np + s]; m.dy = smp[2 * np + s]; m.dz = smp[3 * np + s];
    m.qxx = smp[4 * np + s]; m.qxy = smp[5 * np + s]; m.qxz = smp[6 * np + s];
    m.qyy = smp[7 * np + s]; m.qyz = smp[8 * np + s]; m.qzz = smp[9 * np + s];
    return m;
}

template <class T = double>
__device__ __forceinline__ Mp<T> dip(T x, T y, T z)
{
    Mp<T> m{};
    m.dx = x; m.dy = y; m.dz = z;
    return m;
}

// ---------------------------------------------------------------------------
// tile kernels share this frame: lane = i atom of block I, j atoms of block J
// in LDS, per-lane in-cutoff masks, j accumulators in LDS
#define TILE_PROLOGUE                                                                                          \
    const int lane = threadIdx.x;                                                                              \
    const int np = a.npad;                                                                                     \
    const int ntiles = *a.ntiles;

// T: arithmetic of the pair terms (double, or float in the mixed form with
// FP64 displacements)
template <bool EW, class T, bool CPD = false>
__global__ __launch_bounds__(64) void k_pol_field_tiles(PolArgs a)
{
    TILE_PROLOGUE
    // side 0: block I, side 1: block J; compacted pair schedule (tile_pairs.h)
    __shared__ double bx[2][64], by[2][64], bz[2][64];
    __shared__ T bm[2][10][64], bpd[2][64], bth[2][64];
    __shared__ double acc[2][3][64];
    __shared__ unsigned short plist[64 * 64];
    for (int v = blockIdx.x, nitems = ntiles + *a.items.xcnt; v < nitems; v += gridDim.x) {
        TileItem it;
        if (!tile_item(v, ntiles, a.items, it)) continue;
        const int u = it.u;
        const int t = it.t;
        const int2 IJ = a.tiles[t];
        __syncthreads();
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int s = (side ? IJ.y : IJ.x) * 64 + lane;
            bx[side][lane] = a.sx[s]; by[side][lane] = a.sy[s]; bz[side][lane] = a.sz[s];
            for (int c = 0; c < 10; c++) bm[side][c][lane] = a.smp[c * np + s];
            // Thole: thole and pdamp; HIPPO: total charge and valence exponent
            bth[side][lane] = CPD ? a.scp[2 * np + s] : a.spol[np + s];
            bpd[side][lane] = CPD ? a.scp[np + s] : a.spol[2 * np + s];
            acc[side][0][lane] = acc[side][1][lane] = acc[side][2][lane] = 0.0;
        }
        const int npair = min(tile_pair_list(a.mask[(size_t)t * 128 + lane], plist, lane), it.b1 * 64);
        __syncthreads();
        for (int base = it.b0 * 64; base < npair; base += 64) {
            if (base + lane >= npair) break;
            const int ent = plist[base + lane];
            const int il = ent >> 8, jl = ent & 63;
            double xd = bx[1][jl] - bx[0][il], yd = by[1][jl] - by[0][il], zd = bz[1][jl] - bz[0][il];
            image<double>(a.box, xd, yd, zd);
            const T xr = T(xd), yr = T(yd), zr = T(zd);
            const T r2 = xr * xr + yr * yr + zr * zr;
            Mp<T> mi, mk;
            mi.c = bm[0][0][il]; mi.dx = bm[0][1][il]; mi.dy = bm[0][2][il]; mi.dz = bm[0][3][il];
            mi.qxx = bm[0][4][il]; mi.qxy = bm[0][5][il]; mi.qxz = bm[0][6][il]; mi.qyy = bm[0][7][il];
            mi.qyz = bm[0][8][il]; mi.qzz = bm[0][9][il];
            mk.c = bm[1][0][jl]; mk.dx = bm[1][1][jl]; mk.dy = bm[1][2][jl]; mk.dz = bm[1][3][jl];
            mk.qxx = bm[1][4][jl]; mk.qxy = bm[1][5][jl]; mk.qxz = bm[1][6][jl]; mk.qyy = bm[1][7][jl];
            mk.qyz = bm[1][8][jl]; mk.qzz = bm[1][9][jl];
            T Ei[3] = {0, 0, 0}, Ek[3] = {0, 0, 0};
            if (CPD) {
                T Rt[6], Xi[6], Xk[6], Rm[6];
                hippo_pol_series<T>(r2, bpd[0][il], bpd[1][jl], EW, T(a.aewald), T(1), Rt, Xi, Xk, Rm);
                Mp<T> tk = mk, ti = mi;
                tk.c = bth[1][jl];
                ti.c = bth[0][il];
                field_at_i<T>(Rt, tk, xr, yr, zr, Ei);
                field_at_i<T>(Xk, mk, xr, yr, zr, Ei);
                field_at_k<T>(Rt, ti, xr, yr, zr, Ek);
                field_at_k<T>(Xi, mi, xr, yr, zr, Ek);
            } else {
                T lam[6], R[6];
                thole_lambda<T>(sqrt(r2), bpd[0][il], bpd[1][jl], bth[0][il], bth[1][jl], lam);
                pol_series<EW, T>(r2, T(a.aewald), lam, R);
                mp_fields<T>(R, mi, mk, xr, yr, zr, Ei, Ek);
            }
            for (int c = 0; c < 3; c++) {
                atomicAdd(&acc[0][c][il], (double)Ei[c]);
                atomicAdd(&acc[1][c][jl], (double)Ek[c]);
            }
        }
        __syncthreads();
        // partial sums: side 0 the i atoms, side 1 the j atoms
        T *P = (T *)a.part + (size_t)u * 2 * PMQ_TILE_NV * 64;
        T *Q = P + PMQ_TILE_NV * 64;
        for (int c = 0; c < 3; c++) {
            P[c * 64 + lane] = T(acc[0][c][lane]);
            Q[c * 64 + lane] = T(acc[1][c][lane]);
        }
    }
}

__global__ __launch_bounds__(64) void k_pol_field_excl(PolArgs a)
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
    double lam[6], R[6], Ei[3], Ek[3];
    if (!a.cpd)
        thole_lambda<double>(sqrt(r2), a.spol[2 * np + si], a.spol[2 * np + sk], a.spol[np + si], a.spol[np + sk],
                             lam);
    const Mp<double> mi = load_mp<double>(a.smp, np, si), mk = load_mp<double>(a.smp, np, sk);
    const double sc[2] = {q.d, q.p};
    const bool ew = a.aewald > 0.0;
    double rr[6], bn[6];
    rr_series<double>(r2, 1.0, rr);
    if (ew) ewald_series<double>(r2, a.aewald, bn);
    for (int set = 0; set < 2; set++) {
        // the whole pair at its scale (the tiles leave the scaled pairs out)
        if (sc[set] == 0.0 && !ew) continue;
        if (a.cpd) {
            double Rt[6], Xi[6], Xk[6], Rm[6];
            hippo_pol_series<double>(r2, a.scp[np + si], a.scp[np + sk], ew, a.aewald, sc[set], Rt, Xi, Xk, Rm);
            Mp<double> tk = mk, ti = mi;
            tk.c = a.scp[2 * np + sk];
            ti.c = a.scp[2 * np + si];
            Ei[0] = Ei[1] = Ei[2] = Ek[0] = Ek[1] = Ek[2] = 0.0;
            field_at_i<double>(Rt, tk, xr, yr, zr, Ei);
            field_at_i<double>(Xk, mk, xr, yr, zr, Ei);
            field_at_k<double>(Rt, ti, xr, yr, zr, Ek);
            field_at_k<double>(Xi, mi, xr, yr, zr, Ek);
        } else {
            for (int j = 0; j < 6; j++) R[j] = ew ? bn[j] - (1.0 - sc[set] * lam[j]) * rr[j] : sc[set] * lam[j] * rr[j];
            mp_fields<double>(R, mi, mk, xr, yr, zr, Ei, Ek);
        }
        const size_t o = (size_t)set * 3 * np;
        for (int c = 0; c < 3; c++) {
            field_add(a.dst, a.dstq, o + c * np + si, Ei[c]);
            field_add(a.dst, a.dstq, o + c * np + sk, Ek[c]);
        }
    }
}

// ---------------------------------------------------------------------------
// real-space polarization gradient
template <class T>
__device__ __forceinline__ void pol_pair_grad(const T *R, const Mp<T> &mi, const Mp<T> &mk, const T *udi, const T *upi,
                                              const T *udk, const T *upk, T xr, T yr, T zr, T sdp, T sd, T sp, T su,
                                              T *F, T *Ti, T *Tk)
{
    // sdp weights the (mu_d + mu_p) permanent terms when d and p scales are equal
    // (tiles); sd/sp separately otherwise; su the dipole-dipole term
    PairOut<T> o;
    if (sdp != T(0)) {
        pair_series<T, true>(R, dip<T>(udi[0] + upi[0], udi[1] + upi[1], udi[2] + upi[2]), mk, xr, yr, zr, o);
        F[0] += sdp * o.fx; F[1] += sdp * o.fy; F[2] += sdp * o.fz;
        Tk[0] += sdp * o.tkx; Tk[1] += sdp * o.tky; Tk[2] += sdp * o.tkz;
        pair_series<T, true>(R, mi, dip<T>(udk[0] + upk[0], udk[1] + upk[1], udk[2] + upk[2]), xr, yr, zr, o);
        F[0] += sdp * o.fx; F[1] += sdp * o.fy; F[2] += sdp * o.fz;
        Ti[0] += sdp * o.tix; Ti[1] += sdp * o.tiy; Ti[2] += sdp * o.tiz;
    }
    if (sp != T(0)) { // mu_d with the p-scaled permanent field
        pair_series<T, true>(R, dip<T>(udi[0], udi[1], udi[2]), mk, xr, yr, zr, o);
        F[0] += sp * o.fx; F[1] += sp * o.fy; F[2] += sp * o.fz;
        Tk[0] += sp * o.tkx; Tk[1] += sp * o.tky; Tk[2] += sp * o.tkz;
        pair_series<T, true>(R, mi, dip<T>(udk[0], udk[1], udk[2]), xr, yr, zr, o);
        F[0] += sp * o.fx; F[1] += sp * o.fy; F[2] += sp * o.fz;
        Ti[0] += sp * o.tix; Ti[1] += sp * o.tiy; Ti[2] += sp * o.tiz;
    }
    if (sd != T(0)) { // mu_p with the d-scaled permanent field
        pair_series<T, true>(R, dip<T>(upi[0], upi[1], upi[2]), mk, xr, yr, zr, o);
        F[0] += sd * o.fx; F[1] += sd * o.fy; F[2] += sd * o.fz;
        Tk[0] += sd * o.tkx; Tk[1] += sd * o.tky; Tk[2] += sd * o.tkz;
        pair_series<T, true>(R, mi, dip<T>(upk[0], upk[1], upk[2]), xr, yr, zr, o);
        F[0] += sd * o.fx; F[1] += sd * o.fy; F[2] += sd * o.fz;
        Ti[0] += sd * o.tix; Ti[1] += sd * o.tiy; Ti[2] += sd * o.tiz;
    }
    if (su != T(0)) {
        pair_series<T, true>(R, dip<T>(udi[0], udi[1], udi[2]), dip<T>(upk[0], upk[1], upk[2]), xr, yr, zr, o);
        F[0] += su * o.fx; F[1] += su * o.fy; F[2] += su * o.fz;
        pair_series<T, true>(R, dip<T>(upi[0], upi[1], upi[2]), dip<T>(udk[0], udk[1], udk[2]), xr, yr, zr, o);
        F[0] += su * o.fx; F[1] += su * o.fy; F[2] += su * o.fz;
    }
}

// HIPPO form of pol_pair_grad: the induced dipole at i sees the total
// multipole of k through Rt plus its valence multipole through Xk (and the
// reverse through Xi); the dipole-dipole term uses Rm.  qi/qk: total charges
// (mi/mk carry the valence charges)
// With charge flux (PI/PK not null) also the potential of the induced
// dipoles at each site, the derivative in its valence charge (which enters
// both the total and the valence multipole), in *PI, *PK.
template <class T>
__device__ __forceinline__ void pol_pair_grad_cp(const T *Rt, const T *Xi, const T *Xk, const T *Rm, T qi, T qk,
                                                 const Mp<T> &mi, const Mp<T> &mk, const T *udi, const T *upi,
                                                 const T *udk, const T *upk, T xr, T yr, T zr, T sdp, T sd, T sp,
                                                 T su, T *F, T *Ti, T *Tk, T *PI = nullptr, T *PK = nullptr)
{
    PairOut<T> o;
    Mp<T> ti = mi, tk = mk;
    ti.c = qi;
    tk.c = qk;
    auto perm = [&](T w, T dxi, T dyi, T dzi, T dxk, T dyk, T dzk) {
        if (w == T(0)) return;
        const Mp<T> ui = dip<T>(dxi, dyi, dzi), uk = dip<T>(dxk, dyk, dzk);
        pair_series<T, true>(Rt, ui, tk, xr, yr, zr, o);
        F[0] += w * o.fx; F[1] += w * o.fy; F[2] += w * o.fz;
        Tk[0] += w * o.tkx; Tk[1] += w * o.tky; Tk[2] += w * o.tkz;
        if (PK) *PK += w * o.phik;
        pair_series<T, true>(Xk, ui, mk, xr, yr, zr, o);
        F[0] += w * o.fx; F[1] += w * o.fy; F[2] += w * o.fz;
        Tk[0] += w * o.tkx; Tk[1] += w * o.tky; Tk[2] += w * o.tkz;
        if (PK) *PK += w * o.phik;
        pair_series<T, true>(Rt, ti, uk, xr, yr, zr, o);
        F[0] += w * o.fx; F[1] += w * o.fy; F[2] += w * o.fz;
        Ti[0] += w * o.tix; Ti[1] += w * o.tiy; Ti[2] += w * o.tiz;
        if (PI) *PI += w * o.phii;
        pair_series<T, true>(Xi, mi, uk, xr, yr, zr, o);
        F[0] += w * o.fx; F[1] += w * o.fy; F[2] += w * o.fz;
        Ti[0] += w * o.tix; Ti[1] += w * o.tiy; Ti[2] += w * o.tiz;
        if (PI) *PI += w * o.phii;
    };
    perm(sdp, udi[0] + upi[0], udi[1] + upi[1], udi[2] + upi[2], udk[0] + upk[0], udk[1] + upk[1], udk[2] + upk[2]);
    perm(sp, udi[0], udi[1], udi[2], udk[0], udk[1], udk[2]); // mu_d with the p-scaled permanent field
    perm(sd, upi[0], upi[1], upi[2], upk[0], upk[1], upk[2]); // mu_p with the d-scaled permanent field
    if (su != T(0)) {
        pair_series<T, true>(Rm, dip<T>(udi[0], udi[1], udi[2]), dip<T>(upk[0], upk[1], upk[2]), xr, yr, zr, o);
        F[0] += su * o.fx; F[1] += su * o.fy; F[2] += su * o.fz;
        pair_series<T, true>(Rm, dip<T>(upi[0], upi[1], upi[2]), dip<T>(udk[0], udk[1], udk[2]), xr, yr, zr, o);
        F[0] += su * o.fx; F[1] += su * o.fy; F[2] += su * o.fz;
    }
}

template <bool EW, class T, bool CPD = false, bool CF = false>
__global__ __launch_bounds__(64) void k_pol_grad_tiles(PolArgs a)
{
    TILE_PROLOGUE
    // side 0: block I, side 1: block J; compacted pair schedule (tile_pairs.h)
    constexpr int NA = CF ? 7 : 6; // gradient, torque (and dE/dq with charge flux)
    __shared__ double bx[2][64], by[2][64], bz[2][64];
    __shared__ T bm[2][10][64], bu[2][6][64], bpd[2][64], bth[2][64];
    __shared__ double acc[2][NA][64];
    __shared__ unsigned short ring[TILE_RING];
    const double hf = 0.5 * a.f;
    for (int v = blockIdx.x, nitems = ntiles + *a.items.xcnt; v < nitems; v += gridDim.x) {
        TileItem it;
        if (!tile_item(v, ntiles, a.items, it)) continue;
        const int u = it.u;
        const int t = it.t;
        const int2 IJ = a.tiles[t];
        Vir6 vir;
        __syncthreads();
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int s = (side ? IJ.y : IJ.x) * 64 + lane;
            bx[side][lane] = a.sx[s]; by[side][lane] = a.sy[s]; bz[side][lane] = a.sz[s];
            for (int c = 0; c < 10; c++) bm[side][c][lane] = a.smp[c * np + s];
            for (int c = 0; c < 6; c++) bu[side][c][lane] = a.src[c * np + s];
            // Thole: thole and pdamp; HIPPO: total charge and valence exponent
            bth[side][lane] = CPD ? a.scp[2 * np + s] : a.spol[np + s];
            bpd[side][lane] = CPD ? a.scp[np + s] : a.spol[2 * np + s];
            for (int c = 0; c < NA; c++) acc[side][c][lane] = 0.0;
        }
        TilePairStream ps(a.mask[(size_t)t * 128 + lane]);
        __syncthreads();
        for (int cnt, ent, b = 0; b < it.b1 && (cnt = ps.next(ring, lane, ent)) > 0; b++) {
            if (b < it.b0) continue;
            if (lane >= cnt) break;
            const int il = ent >> 8, jl = ent & 63;
            double xd = bx[1][jl] - bx[0][il], yd = by[1][jl] - by[0][il], zd = bz[1][jl] - bz[0][il];
            image<double>(a.box, xd, yd, zd);
            const T xr = T(xd), yr = T(yd), zr = T(zd);
            const T r2 = xr * xr + yr * yr + zr * zr;
            Mp<T> mi, mk;
            mi.c = bm[0][0][il]; mi.dx = bm[0][1][il]; mi.dy = bm[0][2][il]; mi.dz = bm[0][3][il];
            mi.qxx = bm[0][4][il]; mi.qxy = bm[0][5][il]; mi.qxz = bm[0][6][il]; mi.qyy = bm[0][7][il];
            mi.qyz = bm[0][8][il]; mi.qzz = bm[0][9][il];
            mk.c = bm[1][0][jl]; mk.dx = bm[1][1][jl]; mk.dy = bm[1][2][jl]; mk.dz = bm[1][3][jl];
            mk.qxx = bm[1][4][jl]; mk.qxy = bm[1][5][jl]; mk.qxz = bm[1][6][jl]; mk.qyy = bm[1][7][jl];
            mk.qyz = bm[1][8][jl]; mk.qzz = bm[1][9][jl];
            const T ui[6] = {bu[0][0][il], bu[0][1][il], bu[0][2][il], bu[0][3][il], bu[0][4][il], bu[0][5][il]};
            const T udk[3] = {bu[1][0][jl], bu[1][1][jl], bu[1][2][jl]}, upk[3] = {bu[1][3][jl], bu[1][4][jl], bu[1][5][jl]};
            T f3[3] = {0, 0, 0}, ti[3] = {0, 0, 0}, tk[3] = {0, 0, 0}, pi = T(0), pk = T(0);
            if (CPD) {
                T Rt[6], Xi[6], Xk[6], Rm[6];
                hippo_pol_series<T>(r2, bpd[0][il], bpd[1][jl], EW, T(a.aewald), T(1), Rt, Xi, Xk, Rm);
                pol_pair_grad_cp<T>(Rt, Xi, Xk, Rm, bth[0][il], bth[1][jl], mi, mk, ui, ui + 3, udk, upk, xr, yr, zr,
                                    T(1), T(0), T(0), T(1), f3, ti, tk, CF ? &pi : nullptr, CF ? &pk : nullptr);
            } else {
                T lam[6], R[6];
                thole_lambda<T>(sqrt(r2), bpd[0][il], bpd[1][jl], bth[0][il], bth[1][jl], lam);
                pol_series<EW, T>(r2, T(a.aewald), lam, R);
                pol_pair_grad<T>(R, mi, mk, ui, ui + 3, udk, upk, xr, yr, zr, T(1), T(0), T(0), T(1), f3, ti, tk);
            }
            for (int c = 0; c < 3; c++) {
                atomicAdd(&acc[0][c][il], (double)f3[c]);
                atomicAdd(&acc[0][3 + c][il], (double)ti[c]);
                atomicAdd(&acc[1][c][jl], -(double)f3[c]);
                atomicAdd(&acc[1][3 + c][jl], (double)tk[c]);
            }
            if (CF) {
                atomicAdd(&acc[0][NA - 1][il], (double)pi);
                atomicAdd(&acc[1][NA - 1][jl], (double)pk);
            }
            if (a.vir) vir.pair(xd, yd, zd, hf * (double)f3[0], hf * (double)f3[1], hf * (double)f3[2]);
        }
        __syncthreads();
        // partial sums: side 0 the i atoms, side 1 the j atoms
        T *P = (T *)a.part + (size_t)u * 2 * PMQ_TILE_NV * 64;
        T *Q = P + PMQ_TILE_NV * 64;
        for (int c = 0; c < NA; c++) {
            P[c * 64 + lane] = T(hf * acc[0][c][lane]);
            Q[c * 64 + lane] = T(hf * acc[1][c][lane]);
        }
        if (a.vir) {
            vir.wave_sum();
            if (lane == 0) vir.store_tile(a.tscal, a.tstride, u);
        }
    }
}

__global__ __launch_bounds__(64) void k_pol_grad_excl(PolArgs a)
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
    const Mp<double> mi = load_mp<double>(a.smp, np, si), mk = load_mp<double>(a.smp, np, sk);
    double udi[3], upi[3], udk[3], upk[3];
    for (int c = 0; c < 3; c++) {
        udi[c] = a.src[c * np + si]; upi[c] = a.src[(3 + c) * np + si];
        udk[c] = a.src[c * np + sk]; upk[c] = a.src[(3 + c) * np + sk];
    }
    double F[3] = {0, 0, 0}, Ti[3] = {0, 0, 0}, Tk[3] = {0, 0, 0}, pi = 0.0, pk = 0.0;
    double *PI = a.potp ? &pi : nullptr, *PK = a.potp ? &pk : nullptr;
    // the whole pair (the tiles leave the scaled pairs out): the series at
    // scale s is linear in s, B - rr + s dmp rr, so the damped part takes the
    // scales as weights and the Ewald part B - rr weight 1
    double rr[6];
    rr_series<double>(r2, 1.0, rr);
    if (a.cpd) {
        double Rt[6], Xi[6], Xk[6], Rm[6];
        hippo_pol_series<double>(r2, a.scp[np + si], a.scp[np + sk], false, 0.0, 1.0, Rt, Xi, Xk, Rm);
        pol_pair_grad_cp<double>(Rt, Xi, Xk, Rm, a.scp[2 * np + si], a.scp[2 * np + sk], mi, mk, udi, upi, udk, upk,
                                 xr, yr, zr, 0.0, q.d, q.p, q.u, F, Ti, Tk, PI, PK);
    } else {
        double lam[6], R[6];
        thole_lambda<double>(sqrt(r2), a.spol[2 * np + si], a.spol[2 * np + sk], a.spol[np + si], a.spol[np + sk],
                             lam);
        for (int c = 0; c < 6; c++) R[c] = lam[c] * rr[c];
        pol_pair_grad<double>(R, mi, mk, udi, upi, udk, upk, xr, yr, zr, 0.0, q.d, q.p, q.u, F, Ti, Tk);
    }
    if (a.aewald > 0.0) {
        double bn[6], R0[6];
        ewald_series<double>(r2, a.aewald, bn);
        for (int c = 0; c < 6; c++) R0[c] = bn[c] - rr[c];
        const double Z[6] = {0, 0, 0, 0, 0, 0};
        if (a.cpd)
            pol_pair_grad_cp<double>(R0, Z, Z, R0, a.scp[2 * np + si], a.scp[2 * np + sk], mi, mk, udi, upi, udk, upk,
                                     xr, yr, zr, 0.0, 1.0, 1.0, 1.0, F, Ti, Tk, PI, PK);
        else
            pol_pair_grad<double>(R0, mi, mk, udi, upi, udk, upk, xr, yr, zr, 0.0, 1.0, 1.0, 1.0, F, Ti, Tk);
    }
    const double hf = 0.5 * a.f;
    for (int c = 0; c < 3; c++) {
        grad_add(a.dep, a.depq, 3 * q.i + c, hf * F[c]);
        grad_add(a.dep, a.depq, 3 * q.k + c, -hf * F[c]);
        grad_add(a.tep, a.tepq, 3 * q.i + c, hf * Ti[c]);
        grad_add(a.tep, a.tepq, 3 * q.k + c, hf * Tk[c]);
    }
    if (a.potp) {
        if (a.potpq) {
            fx_add(a.potpq, q.i, hf * pi, PMQ_FX_GRAD);
            fx_add(a.potpq, q.k, hf * pk, PMQ_FX_GRAD);
        } else {
            atomicAdd(&a.potp[q.i], hf * pi);
            atomicAdd(&a.potp[q.k], hf * pk);
        }
    }
    if (a.vir) {
        Vir6 v;
        v.pair(xr, yr, zr, hf * F[0], hf * F[1], hf * F[2]);
        a.tot->add_vir(v);
    }
}


// ---------------------------------------------------------------------------
// preconditioned conjugate gradient on (1/alpha - T) mu = E, both sets at
// once (pcg_steps.h: the iteration, its scalars and reductions).  Vectors
// are 2 x 3 x npad, component t = c npad + s.  The convergence test runs on
// the device, so iterations are enqueued in batches without host round
// trips; once PCG_DONE is set the remaining launches of a batch return at
// once.

// the first guess u = a E, and the positions of both packed records (fixed
// during the solve); u also into the row product's packed input, and with
// the sparse preconditioner into its packed input, whose launch then adds
// the near field of these dipoles to the guess (k_pcg_prec with E)
__global__ __launch_bounds__(256) void k_pcg_guess(PcgVec a, const double *E, const double *sx, const double *sy,
                                                   const double *sz)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * a.npad) return;
    const int s = t % a.npad, c = t / a.npad;
    const double u = a.spol[s] * E[t];
    a.u[t] = u;
    a.pks[(size_t)PMQ_PK * s + 3 + c] = u;
    if (a.prec) a.pkw[(size_t)PMQ_PK * s + 3 + c] = u;
    if (c < 3) {
        const double x = c == 0 ? sx[s] : c == 1 ? sy[s] : sz[s];
        a.pks[(size_t)PMQ_PK * s + c] = x;
        if (a.prec) a.pkw[(size_t)PMQ_PK * s + c] = x;
    }
}

// the initial residual r = E + T u - u/a (Tu = T u, the whole operator's
// field), p = v = 0, the scalars cleared; with the sparse preconditioner
// a r into its packed input (it forms z and the partials), else z = a r and
// the partials here
__global__ __launch_bounds__(PCG_TB) void k_pcg_init(PcgVec a, const double *E, const double *Tu)
{
    if (blockIdx.x == 0 && threadIdx.x < 16) a.sc[threadIdx.x] = 0.0;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double rz[2] = {0, 0}, rr[2] = {0, 0};
    const double none[2] = {0.0, 0.0};
    if (t < 6 * a.npad) {
        const int s = t % a.npad, c = t / a.npad;
        const double pol = a.spol[s];
        const double r = pol != 0.0 ? E[t] + Tu[t] - a.u[t] / pol : 0.0;
        a.r[t] = r;
        a.p[t] = 0.0;
        a.v[t] = 0.0;
        if (a.prec) a.pkw[(size_t)PMQ_PK * s + 3 + c] = pol * r;
        else pcg_e2(a, s, c, pol * r, none, rz[c / 3], rr[c / 3]);
    }
    if (!a.prec) pcg_e2_end(a, rz[0], rz[1], rr[0], rr[1]);
}

// the step E2 with the diagonal preconditioner (P v = a v)
__global__ __launch_bounds__(PCG_TB) void k_pcg_step(PcgVec a)
{
    if (a.sc[PCG_DONE] != 0.0) return;
    double step[2];
    pcg_e2_begin(a, step);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double rz[2] = {0, 0}, rr[2] = {0, 0};
    if (t < 6 * a.npad) {
        const int s = t % a.npad, c = t / a.npad;
        pcg_e2(a, s, c, a.spol[s] * a.v[t], step, rz[c / 3], rr[c / 3]);
    }
    pcg_e2_end(a, rz[0], rz[1], rr[0], rr[1]);
}

// the convergence test of iteration a.iter, the last of a batch (one block)
__global__ __launch_bounds__(PCG_TB) void k_pcg_check(PcgVec a)
{
    if (a.sc[PCG_DONE] != 0.0) return;
    pcg_converged(a, a.iter + 1);
}

// block sum of one value, one atomic per block (blockDim 256)
__device__ __forceinline__ void block_add1(double v, double *dst)
{
    __shared__ double red[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(dst, ((red[0] + red[1]) + red[2]) + red[3]);
}

// E_pol = -(f/2) sum mu_d . E^p
// the induced dipoles to the caller's order (AoS), and the polarization
// energy -(f/2) mu_d . E^p summed per atom, then per block
// (analysis: the per-atom energies into aep and the nonzero ones counted
// in nep)
__global__ __launch_bounds__(256) void k_pol_finish(int n, int npad, const int *perm, const double *u, double *ud_out,
                                                   double *up_out, double f, const double *Ep, Totals *tot,
                                                   double *aep)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (s < n) {
        const int i = perm[s];
        for (int c = 0; c < 3; c++) {
            const double ud = u[c * npad + s];
            ud_out[3 * i + c] = ud;
            up_out[3 * i + c] = u[(3 + c) * npad + s];
            v += ud * Ep[c * npad + s];
        }
        v *= -0.5 * f;
        if (aep) {
            aep[i] = v;
            if (v != 0.0) atomicAdd(&tot->nep, 1ull);
        }
    }
    block_add1(v, &tot->ep);
}

// ---------------------------------------------------------------------------
// launchers
int pme_mutual_field(pmq_ctx *c, const double *dip, double *field, const double *skip, PcgVec *pcg);
int pme_polar_grad(pmq_ctx *c, bool vir);
int launch_zero(pmq_ctx *c, double *p, size_t n);

static PolArgs pol_args(pmq_ctx *c, double off2)
{
    PolArgs a;
    a.n = c->n; a.npad = c->npad;
    a.tiles = c->tiles.p; a.ntiles = c->ntiles_eff.p;
    a.sx = c->sx.p; a.sy = c->sy.p; a.sz = c->sz.p; a.smp = c->smp.p; a.spol = c->spol.p;
    a.scp = c->scp.p;
    a.cpd = c->prm.pentype != PMQ_PEN_NONE ? 1 : 0;
    a.sflag = c->sflag.p; a.perm = c->perm.p; a.iperm = c->iperm.p;
    a.box = c->box;
    a.off2 = off2;
    a.aewald = c->prm.use_ewald ? c->prm.aewald : 0.0;
    a.f = c->prm.electric / c->prm.dielec;
    a.excl = c->excl.p;
    a.nexcl = (int)c->h_excl.size();
    a.mask = c->xmask.p; // the scaled pairs are done whole by the exclusion kernels
    a.src = nullptr; a.dst = nullptr; a.dep = c->dep.p; a.tep = c->tep.p;
    a.dstq = nullptr; a.depq = nullptr; a.tepq = nullptr;
    a.potp = nullptr; a.potpq = nullptr;
    a.part = c->tpart.p;
    a.vir = 0;
    a.tscal = c->tscal.p;
    a.tstride = item_stride(c);
    a.items = tile_items(c);
    a.tot = c->totals.p;
    return a;
}

static int tile_grid(pmq_ctx *c) { return c->igrid; }

// dst = T src (real space, from the pair matrix) + reciprocal and self
// field; skip: the solve's converged flag for the iterations of a
// speculative batch (their launches return at once once it is set), or null;
// pcg: the PCG step E1 on T src (src = z) instead of storing it, or null
static int mutual_field(pmq_ctx *c, const double *src, double *dst, const double *skip = nullptr,
                        PcgVec *pcg = nullptr)
{
    // the product reads src from its packed copy (written with src by the
    // PCG kernels); the reciprocal part reads src itself
    if (c->prm.use_ewald) return pme_mutual_field(c, src, dst, skip, pcg); // product and spreading in one launch
    return pair_matrix_apply(c, dst, skip, pcg, pcg ? PCG_APPLY_E1 : PCG_APPLY_NONE);
}

int polar_prepare(pmq_ctx *c, double off2, bool mixed, bool safe)
{
    // the real-space permanent field (the sorted parameters and the zeroed
    // field come with the multipole rotation: launch_rotpole)
    const int np = c->npad;
    int rc;
    if ((rc = pair_matrix_build(c, mixed, safe))) return rc;
    PolArgs a = pol_args(c, off2);
    a.dst = c->pfield.p;
    a.dstq = mixed ? c->fx_field.p : nullptr;
    {
        ProfScope ps_(c, "pol_field_tiles");
        const dim3 g(tile_grid(c));
#define PMQ_FIELD(EW, T, CPD) hipLaunchKernelGGL((k_pol_field_tiles<EW, T, CPD>), g, dim3(64), 0, c->stream, a)
        const int sel = (mixed ? 4 : 0) + (a.aewald > 0 ? 2 : 0) + (a.cpd ? 1 : 0);
        switch (sel) {
        case 0: PMQ_FIELD(false, double, false); break;
        case 1: PMQ_FIELD(false, double, true); break;
        case 2: PMQ_FIELD(true, double, false); break;
        case 3: PMQ_FIELD(true, double, true); break;
        case 4: PMQ_FIELD(false, float, false); break;
        case 5: PMQ_FIELD(false, float, true); break;
        case 6: PMQ_FIELD(true, float, false); break;
        default: PMQ_FIELD(true, float, true); break;
        }
#undef PMQ_FIELD
    }
    PMQ_HIP(hipGetLastError());
    // the scaled pairs first: the reduction takes their fixed-point share of
    // each set (mixed form) into its sums
    if (a.nexcl) hipLaunchKernelGGL(k_pol_field_excl, dim3((a.nexcl + 63) / 64), dim3(64), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    TileReduceArgs r{};
    for (int q = 0; q < 3; q++) {
        r.out[q] = c->pfield.p + (size_t)q * np;
        r.out2[q] = c->pfield.p + (size_t)(3 + q) * np;
        r.stride[q] = 1;
        r.fxscale[q] = PMQ_FX_FIELD;
        r.fxq[q] = mixed ? c->fx_field.p + (size_t)q * np : nullptr;
        r.fxq2[q] = mixed ? c->fx_field.p + (size_t)(3 + q) * np : nullptr;
    }
    return tile_reduce(c, r, 3, mixed, mixed);
}

// safe: iterations go out one at a time, each followed by a host check of
// convergence; otherwise one batch of (last count + 1) iterations is
// enqueued with no round trip, iterations past convergence return at once,
// and the count is read back with the evaluation's status (it is redone
// safely if the batch fell short)
int polar_solve(pmq_ctx *c, bool safe, int *iters_out)
{
    const int np = c->npad;
    const size_t nv = 6 * (size_t)np;
    const int gb = (int)((nv + PCG_TB - 1) / PCG_TB);
    const bool prec = c->prec_on;
    // E1 partials (the field gather, at most 64 lanes per atom, or the row
    // product), then the E2 ones (the preconditioner or the vector kernels)
    const int n2 = prec ? pair_matrix_prec_blocks(c) : gb;
    c->pcg_part1_cap = 2 * ((size_t)np / 4 + 64);
    PMQ_HIP(c->pcg_part.alloc(c->pcg_part1_cap + 4 * (size_t)n2));
    PMQ_HIP(c->pk_src.alloc((size_t)PMQ_PK * np));
    if (prec) PMQ_HIP(c->pk_w.alloc((size_t)PMQ_PK * np));
    PcgVec a;
    a.npad = np; a.iter = 0; a.prec = prec ? 1 : 0;
    a.spol = c->spol.p; a.udiag = c->prm.udiag;
    a.u = c->uind.p; a.r = c->pres.p; a.z = c->pz.p; a.p = c->pdir.p; a.v = c->pav.p;
    a.pks = c->pk_src.p;
    a.pkw = prec ? c->pk_w.p : nullptr;
    a.part1 = c->pcg_part.p;
    a.n1 = 0;
    a.part2 = c->pcg_part.p + c->pcg_part1_cap;
    a.n2 = n2;
    a.sc = c->pcg.p;
    int npolar = 0;
    for (double p : c->h_polarity) npolar += p != 0.0;
    a.npolar = std::max(npolar, 1);
    a.poleps = c->prm.poleps;
    int rc;
    hipLaunchKernelGGL(k_pcg_guess, dim3(gb), dim3(256), 0, c->stream, a, c->pfield.p, c->sx.p, c->sy.p, c->sz.p);
    // with the sparse preconditioner the guess takes the near field of the
    // direct dipoles once more: a (E + T_s a E), one Jacobi step on the near
    // part (8 -> 7 iterations to 1e-6 D on the 12,288-atom water box)
    if (prec && (rc = pair_matrix_precondition(c, a, c->pfield.p))) return rc;
    if ((rc = mutual_field(c, c->uind.p, c->pvec.p))) return rc;
    hipLaunchKernelGGL(k_pcg_init, dim3(gb), dim3(PCG_TB), 0, c->stream, a, c->pfield.p, c->pvec.p);
    // with PME and the sparse preconditioner, the preconditioner's launches
    // also carry the reciprocal potential grids of p and u by linearity, so
    // the gradient's reciprocal part needs no spreading or transform of u
    void *qg_, *gp_, *gu_;
    size_t pts_;
    bool g32_;
    const bool carry = prec && pme_carry_grids(c, &qg_, &gp_, &gu_, &pts_, &g32_);
    c->phi_carried = false;
    if (prec && (rc = pair_matrix_precondition(c, a, nullptr, carry))) return rc;
    PMQ_HIP(hipGetLastError());
    const bool spec = !safe && c->pcg_guess_iters > 0;
    int batch = spec ? std::min(c->pcg_guess_iters + 1, c->prm.politer) : 1;
    int it = 0;
    double sc[11];
    bool done = false;
    while (!done && it < c->prm.politer) {
        const int nb = std::min(batch, c->prm.politer - it);
        for (int k = 0; k < nb; k++) {
            // before the speculative last iteration of a batch, the test of
            // the one before it, so that all of that iteration's launches
            // return at once when it is not needed (otherwise only the
            // launches after its first one would)
            if (spec && k == nb - 1 && k > 0) hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(PCG_TB), 0, c->stream, a);
            a.iter = ++it;
            // the convergence test of the last iteration, the operator on z
            // ending in the step E1 (which records its block count), then E2
            if ((rc = mutual_field(c, a.z, c->pvec.p, a.sc + PCG_DONE, &a))) return rc;
            if (prec) {
                if ((rc = pair_matrix_precondition(c, a, nullptr, carry))) return rc;
            } else {
                hipLaunchKernelGGL(k_pcg_step, dim3(gb), dim3(PCG_TB), 0, c->stream, a);
            }
            PMQ_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(PCG_TB), 0, c->stream, a);
        PMQ_HIP(hipGetLastError());
        if (spec) {
            *iters_out = 0; // known after the status check
            c->phi_carried = carry;
            return PMQ_OK;
        }
        PMQ_HIP(hipMemcpyAsync(sc, c->pcg.p, sizeof sc, hipMemcpyDeviceToHost, c->stream));
        PMQ_HIP(hipStreamSynchronize(c->stream));
        done = sc[PCG_DONE] != 0.0;
    }
    const int iters = done ? (int)sc[PCG_ITERS] : it;
    if (done) c->pcg_guess_iters = iters;
    *iters_out = done ? iters : -iters;
    c->phi_carried = carry;
    return PMQ_OK;
}

int polar_gradient(pmq_ctx *c, double off2, bool mixed, bool vir)
{
    PolArgs a = pol_args(c, off2);
    a.src = c->uind.p;
    a.vir = vir ? 1 : 0;
    if (mixed) {
        a.depq = c->fx_grad.p;
        a.tepq = c->fx_grad.p + 3 * (size_t)c->n;
    }
    // charge flux (with the charge-penetration damping): the induced dipoles'
    // potential at every site as well
    const bool cf = a.cpd && c->prm.use_chgflx;
    if (cf) {
        a.potp = c->potp.p;
        a.potpq = mixed ? c->fx_pot.p + c->n : nullptr;
    }
    {
        ProfScope ps_(c, "pol_grad_tiles");
        const dim3 g(tile_grid(c));
#define PMQ_GRAD(EW, T, CPD) hipLaunchKernelGGL((k_pol_grad_tiles<EW, T, CPD>), g, dim3(64), 0, c->stream, a)
#define PMQ_GRADCF(EW, T) hipLaunchKernelGGL((k_pol_grad_tiles<EW, T, true, true>), g, dim3(64), 0, c->stream, a)
        const int sel = (mixed ? 4 : 0) + (a.aewald > 0 ? 2 : 0) + (a.cpd ? 1 : 0);
        switch (sel) {
        case 0: PMQ_GRAD(false, double, false); break;
        case 1: if (cf) PMQ_GRADCF(false, double); else PMQ_GRAD(false, double, true); break;
        case 2: PMQ_GRAD(true, double, false); break;
        case 3: if (cf) PMQ_GRADCF(true, double); else PMQ_GRAD(true, double, true); break;
        case 4: PMQ_GRAD(false, float, false); break;
        case 5: if (cf) PMQ_GRADCF(false, float); else PMQ_GRAD(false, float, true); break;
        case 6: PMQ_GRAD(true, float, false); break;
        default: if (cf) PMQ_GRADCF(true, float); else PMQ_GRAD(true, float, true); break;
        }
#undef PMQ_GRAD
#undef PMQ_GRADCF
    }
    PMQ_HIP(hipGetLastError());
    TileReduceArgs r{};
    for (int q = 0; q < 3; q++) {
        r.out[q] = c->dep.p; r.out[3 + q] = c->tep.p;
        r.caller_order[q] = r.caller_order[3 + q] = 1;
        r.stride[q] = r.stride[3 + q] = 3;
        r.off[q] = r.off[3 + q] = q;
        r.fxscale[q] = r.fxscale[3 + q] = PMQ_FX_GRAD;
    }
    if (cf) {
        r.out[6] = c->potp.p;
        r.caller_order[6] = 1; r.stride[6] = 1; r.off[6] = 0; r.fxscale[6] = PMQ_FX_GRAD;
    }
    for (int q = 0; q < 6; q++) r.fxq[q] = mixed ? c->fx_grad.p + (q < 3 ? 0 : 3 * (size_t)c->n) : nullptr;
    // the scaled pairs first: the reduction takes their fixed-point share of
    // the gradient and torque (mixed form) into its sums, and the virial's
    // tile scalars in the same launch
    if (a.nexcl) hipLaunchKernelGGL(k_pol_grad_excl, dim3((a.nexcl + 63) / 64), dim3(64), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    int rc;
    if ((rc = tile_reduce(c, r, cf ? 8 : 6, mixed, mixed, 0, vir ? 1 : 0, vir ? 10 : 0))) return rc;
    if (a.aewald > 0) return pme_polar_grad(c, vir);
    return PMQ_OK;
}

int polar_finish(pmq_ctx *c, double *ud_out, double *up_out, double *aep)
{
    const int np = c->npad;
    hipLaunchKernelGGL(k_pol_finish, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, np, c->perm.p,
                       c->uind.p, ud_out, up_out, c->prm.electric / c->prm.dielec, c->pfield.p + 3 * (size_t)np,
                       c->totals.p, aep);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

} // namespace pmq
