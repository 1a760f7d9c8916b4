// k_batch.hip -- many independent non-periodic clusters (no cutoff) in one
// persistent launch.
//
// Each workgroup takes clusters from a work queue (largest first) and solves
// one cluster entirely in LDS: direct (permanent) fields, the mutual-dipole
// PCG to the cluster's own convergence test, the multipole and polarization
// energies, and (GRADIENT) forces and torques.  Pair work is row-owned: lane
// group `row` owns atom `row` of the cluster and sums over every other atom
// of the cluster (TPR lanes per row split the partner loop and combine with
// shuffles), so the all-pairs sums need no atomics; the scaled (excluded)
// pairs of the cluster are corrected afterwards through LDS accumulators (64-
// bit fixed point in the mixed form).  Frames (k_rotpole) run before and the
// torque-to-force conversion (k_torque) after this kernel, in the same graph.
//
// The physics per pair is the one of the periodic path with no Ewald
// screening and every pair inside the cutoff: permanent fields and the
// dipole tensor with Thole damping at scale 1 plus (s - 1) lambda rr
// corrections, PCG with the sparse preconditioner of the periodic path (the
// pairs within usolvcut; with pcgprec off the polarizability alone) and its
// first guess a (E + T_s a E), the test debye * sqrt(max_set r.r / npolar)
// < poleps, E_pol = -(f/2) mu_d . E^p, and the variational gradient.
#include <algorithm>
#include <cstddef>
#include <mutex>

#include "pair_mpole.h"
#include "param_pair.h"
#include "pmq_internal.h"

namespace pmq {

constexpr int BATCH_MAX = 256; // atoms per cluster

struct BatchArgs {
    int nsys;
    const int *order;     // clusters, largest first
    const int *off;       // nsys + 1 atom offsets
    const int *eoff;      // nsys + 1 offsets into the excluded-pair list
    const ExclPair *excl;
    const double *x, *y, *z;
    const double *smp;    // 10 x npad rotated multipoles, caller order
    int npad;
    const double *pol, *thl, *pdm;
    const int *useflag;   // per atom: 1 active (use), else 0
    double f, poleps;
    int politer;
    // the sparse preconditioner (pcgprec): P x = udiag a x + max(a, 1e-8)
    // T_s (a x), T_s the dipole field of the pairs within usolvcut; off:
    // P x = a x
    int prec;
    double us2, udiag;
    int grad, do_mpole, do_polar;
    double *dem, *te, *dep, *tep;      // caller order (written)
    double *em_c, *ep_c, *it_c;        // per cluster
    int *counter;
    Totals *tot;
    // parameter gradients (pgrad): per atom, caller order, global frame
    int pgrad;
    double *pg_dm;                     // 10 x npad: dE/d(rotated multipole components)
    double *pg_dpd, *pg_dth;           // dE/dpdamp, dE/dthole
    double *uout;                      // 6 x npad: the dipoles (d, p), for dE/dpolarity
};

template <class T, int ROWS>
__device__ __forceinline__ Mp<T> mp_at(const T (*m)[ROWS], int k)
{
    Mp<T> r;
    r.c = m[0][k]; r.dx = m[1][k]; r.dy = m[2][k]; r.dz = m[3][k];
    r.qxx = m[4][k]; r.qxy = m[5][k]; r.qxz = m[6][k]; r.qyy = m[7][k]; r.qyz = m[8][k]; r.qzz = m[9][k];
    return r;
}

// field at i from multipole k and at k from multipole i (Thole-damped series R)
template <class T>
__device__ __forceinline__ void pair_fields(const T *R, const Mp<T> &a, const Mp<T> &b, T xr, T yr, T zr, T *Ei, T *Ek)
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

template <class T>
__device__ __forceinline__ Mp<T> dipole(T x, T y, T z)
{
    Mp<T> m{};
    m.dx = x; m.dy = y; m.dz = z;
    return m;
}

// gradient and torques of the variational polarization energy for one pair
// (same terms as the periodic path's pair gradient)
template <class T>
__device__ __forceinline__ void polgrad_pair(const T *R, const Mp<T> &mi, const Mp<T> &mk, const T *udi, const T *upi,
                                             const T *udk, const T *upk, T xr, T yr, T zr, T sdp, T sd, T sp, T su,
                                             T *F, T *Ti, T *Tk)
{
    PairOut<T> o;
    if (sdp != T(0)) {
        pair_series<T, true>(R, dipole<T>(udi[0] + upi[0], udi[1] + upi[1], udi[2] + upi[2]), mk, xr, yr, zr, o);
        F[0] += sdp * o.fx; F[1] += sdp * o.fy; F[2] += sdp * o.fz;
        Tk[0] += sdp * o.tkx; Tk[1] += sdp * o.tky; Tk[2] += sdp * o.tkz;
        pair_series<T, true>(R, mi, dipole<T>(udk[0] + upk[0], udk[1] + upk[1], udk[2] + upk[2]), xr, yr, zr, o);
        F[0] += sdp * o.fx; F[1] += sdp * o.fy; F[2] += sdp * o.fz;
        Ti[0] += sdp * o.tix; Ti[1] += sdp * o.tiy; Ti[2] += sdp * o.tiz;
    }
    if (sp != T(0)) {
        pair_series<T, true>(R, dipole<T>(udi[0], udi[1], udi[2]), mk, xr, yr, zr, o);
        F[0] += sp * o.fx; F[1] += sp * o.fy; F[2] += sp * o.fz;
        Tk[0] += sp * o.tkx; Tk[1] += sp * o.tky; Tk[2] += sp * o.tkz;
        pair_series<T, true>(R, mi, dipole<T>(udk[0], udk[1], udk[2]), xr, yr, zr, o);
        F[0] += sp * o.fx; F[1] += sp * o.fy; F[2] += sp * o.fz;
        Ti[0] += sp * o.tix; Ti[1] += sp * o.tiy; Ti[2] += sp * o.tiz;
    }
    if (sd != T(0)) {
        pair_series<T, true>(R, dipole<T>(upi[0], upi[1], upi[2]), mk, xr, yr, zr, o);
        F[0] += sd * o.fx; F[1] += sd * o.fy; F[2] += sd * o.fz;
        Tk[0] += sd * o.tkx; Tk[1] += sd * o.tky; Tk[2] += sd * o.tkz;
        pair_series<T, true>(R, mi, dipole<T>(upk[0], upk[1], upk[2]), xr, yr, zr, o);
        F[0] += sd * o.fx; F[1] += sd * o.fy; F[2] += sd * o.fz;
        Ti[0] += sd * o.tix; Ti[1] += sd * o.tiy; Ti[2] += sd * o.tiz;
    }
    if (su != T(0)) {
        pair_series<T, true>(R, dipole<T>(udi[0], udi[1], udi[2]), dipole<T>(upk[0], upk[1], upk[2]), xr, yr, zr, o);
        F[0] += su * o.fx; F[1] += su * o.fy; F[2] += su * o.fz;
        pair_series<T, true>(R, dipole<T>(upi[0], upi[1], upi[2]), dipole<T>(udk[0], udk[1], udk[2]), xr, yr, zr, o);
        F[0] += su * o.fx; F[1] += su * o.fy; F[2] += su * o.fz;
    }
}

// LDS accumulator: FP64 atomics, or 64-bit fixed point (FX) so that the sum
// does not depend on the order of the corrections
template <bool FX, int ROWS>
struct LdsAcc {
    double d[6][ROWS];
    long long q[6][ROWS];
    __device__ void clear(int tid, int nthr)
    {
        for (int k = tid; k < 6 * ROWS; k += nthr) {
            (&d[0][0])[k] = 0.0;
            (&q[0][0])[k] = 0;
        }
    }
    __device__ void add(int c, int i, double v, double scale)
    {
        if (FX) atomicAdd((unsigned long long *)&q[c][i], (unsigned long long)__double2ll_rn(v * scale));
        else atomicAdd(&d[c][i], v);
    }
    __device__ double get(int c, int i, double scale) const { return FX ? (double)q[c][i] / scale : d[c][i]; }
};

// sum over the TPR lanes of one row (adjacent lanes)
template <int TPR>
__device__ __forceinline__ double row_sum(double v)
{
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// deterministic block sum of two values (one per row owner; fixed tree)
template <int NT>
__device__ __forceinline__ void block_sum2(double a, double b, double *red, double &sa, double &sb)
{
    __syncthreads();
    red[threadIdx.x] = a;
    red[NT + threadIdx.x] = b;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[threadIdx.x] += red[threadIdx.x + w];
            red[NT + threadIdx.x] += red[NT + threadIdx.x + w];
        }
        __syncthreads();
    }
    sa = red[0];
    sb = red[NT];
    __syncthreads();
}

template <class T, int ROWS, int TPR, bool FX>
__global__ __launch_bounds__(ROWS * TPR) void k_batch_clusters(BatchArgs a)
{
    constexpr int NT = ROWS * TPR;
    __shared__ double px[ROWS], py[ROWS], pz[ROWS];
    __shared__ T mp[10][ROWS];
    __shared__ T spd[ROWS], sth[ROWS];
    __shared__ double vec[6][ROWS]; // shared vector: PCG direction, then the dipoles
    __shared__ LdsAcc<FX, ROWS> acc;
    __shared__ double red[2 * NT];
    __shared__ int job;
    const int tid = threadIdx.x;
    const int row = tid / TPR, part = tid % TPR;
    const double debye = 4.80321;
    const double fx_field = PMQ_FX_FIELD, fx_grad = PMQ_FX_GRAD;

    for (;;) {
        __syncthreads();
        if (tid == 0) job = atomicAdd(a.counter, 1);
        __syncthreads();
        if (job >= a.nsys) break;
        const int cl = a.order[job];
        const int o0 = a.off[cl], nc = a.off[cl + 1] - o0;
        const int e0 = a.eoff[cl], e1 = a.eoff[cl + 1];
        // ---- load the cluster
        for (int k = tid; k < ROWS; k += NT) {
            const bool on = k < nc;
            const int g = o0 + k;
            px[k] = on ? a.x[g] : 0.0; py[k] = on ? a.y[g] : 0.0; pz[k] = on ? a.z[g] : 0.0;
            for (int c = 0; c < 10; c++) mp[c][k] = on ? T(a.smp[(size_t)c * a.npad + g]) : T(0);
            spd[k] = on ? T(a.pdm[g]) : T(0);
            sth[k] = on ? T(a.thl[g]) : T(0);
        }
        acc.clear(tid, NT);
        __syncthreads();
        const bool own = row < nc && part == 0; // the lane that keeps the row's state
        const bool rowon = row < nc;
        const int gi = o0 + (rowon ? row : 0);
        const double alpha = rowon ? a.pol[gi] : 0.0;
        const double xi = px[rowon ? row : 0], yi = py[rowon ? row : 0], zi = pz[rowon ? row : 0];
        const Mp<T> mi = mp_at<T, ROWS>(mp, rowon ? row : 0);
        const T pdi = spd[rowon ? row : 0], thi = sth[rowon ? row : 0];
        double emc = 0.0, epc = 0.0;
        int iters = 0;

        // ---- permanent fields at scale 1 (row-owned), then the d / p corrections
        double Ed[3] = {0, 0, 0}, Ep[3] = {0, 0, 0};
        if (a.do_polar) {
            double e3[3] = {0, 0, 0};
            if (rowon)
                for (int k = part; k < nc; k += TPR) {
                    if (k == row) continue;
                    double xd = px[k] - xi, yd = py[k] - yi, zd = pz[k] - zi;
                    const T xr = T(xd), yr = T(yd), zr = T(zd);
                    const T r2 = xr * xr + yr * yr + zr * zr;
                    T lam[6], R[6], Ei[3], Ek[3];
                    thole_lambda<T>(sqrt(r2), pdi, spd[k], thi, sth[k], lam);
                    pol_series<false, T>(r2, T(0), lam, R);
                    pair_fields<T>(R, mi, mp_at<T, ROWS>(mp, k), xr, yr, zr, Ei, Ek);
                    for (int c = 0; c < 3; c++) e3[c] += Ei[c];
                }
            for (int c = 0; c < 3; c++) e3[c] = row_sum<TPR>(e3[c]);
            for (int q = e0 + tid; q < e1; q += NT) {
                const ExclPair ex = a.excl[q];
                if (ex.d == 1.0 && ex.p == 1.0) continue;
                const int i = ex.i - o0, k = ex.k - o0;
                const double xr = px[k] - px[i], yr = py[k] - py[i], zr = pz[k] - pz[i];
                const double r2 = xr * xr + yr * yr + zr * zr;
                double lam[6], R[6], Ei[3], Ek[3];
                thole_lambda<double>(sqrt(r2), a.pdm[ex.i], a.pdm[ex.k], a.thl[ex.i], a.thl[ex.k], lam);
                Mp<double> m1, m2;
                m1 = Mp<double>{};
                m2 = Mp<double>{};
                m1.c = a.smp[ex.i]; m1.dx = a.smp[a.npad + ex.i]; m1.dy = a.smp[2 * (size_t)a.npad + ex.i];
                m1.dz = a.smp[3 * (size_t)a.npad + ex.i]; m1.qxx = a.smp[4 * (size_t)a.npad + ex.i];
                m1.qxy = a.smp[5 * (size_t)a.npad + ex.i]; m1.qxz = a.smp[6 * (size_t)a.npad + ex.i];
                m1.qyy = a.smp[7 * (size_t)a.npad + ex.i]; m1.qyz = a.smp[8 * (size_t)a.npad + ex.i];
                m1.qzz = a.smp[9 * (size_t)a.npad + ex.i];
                m2.c = a.smp[ex.k]; m2.dx = a.smp[a.npad + ex.k]; m2.dy = a.smp[2 * (size_t)a.npad + ex.k];
                m2.dz = a.smp[3 * (size_t)a.npad + ex.k]; m2.qxx = a.smp[4 * (size_t)a.npad + ex.k];
                m2.qxy = a.smp[5 * (size_t)a.npad + ex.k]; m2.qxz = a.smp[6 * (size_t)a.npad + ex.k];
                m2.qyy = a.smp[7 * (size_t)a.npad + ex.k]; m2.qyz = a.smp[8 * (size_t)a.npad + ex.k];
                m2.qzz = a.smp[9 * (size_t)a.npad + ex.k];
                const double sc[2] = {ex.d, ex.p};
                for (int set = 0; set < 2; set++) {
                    if (sc[set] == 1.0) continue;
                    corr_series<double>(r2, sc[set], lam, R);
                    pair_fields<double>(R, m1, m2, xr, yr, zr, Ei, Ek);
                    for (int c = 0; c < 3; c++) {
                        acc.add(3 * set + c, i, Ei[c], fx_field);
                        acc.add(3 * set + c, k, Ek[c], fx_field);
                    }
                }
            }
            __syncthreads();
            if (rowon)
                for (int c = 0; c < 3; c++) {
                    Ed[c] = e3[c] + acc.get(c, row, fx_field);
                    Ep[c] = e3[c] + acc.get(3 + c, row, fx_field);
                }
        }

        // ---- PCG, both sets at once; u, r, z live with the row, p in LDS
        double u[6] = {0, 0, 0, 0, 0, 0};
        if (a.do_polar) {
            int npolar = 0;
            {
                double s1, s2;
                block_sum2<NT>(own && alpha != 0.0 ? 1.0 : 0.0, 0.0, red, s1, s2);
                npolar = (int)s1;
            }
            const double npol = npolar > 0 ? (double)npolar : 1.0;
            const double E6[6] = {Ed[0], Ed[1], Ed[2], Ep[0], Ep[1], Ep[2]};
            for (int c = 0; c < 6; c++) u[c] = alpha * E6[c];
            // T v into tv (row-owned + u-scale corrections); v taken from vec;
            // near: only the pairs within usolvcut (the preconditioner's T_s)
            auto apply_T = [&](double *tv, bool near) {
                double s[6] = {0, 0, 0, 0, 0, 0};
                if (rowon)
                    for (int k = part; k < nc; k += TPR) {
                        if (k == row) continue;
                        const double xd = px[k] - xi, yd = py[k] - yi, zd = pz[k] - zi;
                        if (near && xd * xd + yd * yd + zd * zd > a.us2) continue;
                        const T xr = T(xd), yr = T(yd), zr = T(zd);
                        const T r2 = xr * xr + yr * yr + zr * zr;
                        T lam[6], R[6];
                        thole_lambda<T>(sqrt(r2), pdi, spd[k], thi, sth[k], lam);
                        pol_series<false, T>(r2, T(0), lam, R);
                        for (int set = 0; set < 2; set++) {
                            const T vx = T(vec[3 * set][k]), vy = T(vec[3 * set + 1][k]), vz = T(vec[3 * set + 2][k]);
                            const T vr = vx * xr + vy * yr + vz * zr;
                            s[3 * set + 0] += double(vr * R[2] * xr - R[1] * vx);
                            s[3 * set + 1] += double(vr * R[2] * yr - R[1] * vy);
                            s[3 * set + 2] += double(vr * R[2] * zr - R[1] * vz);
                        }
                    }
                for (int c = 0; c < 6; c++) s[c] = row_sum<TPR>(s[c]);
                acc.clear(tid, NT);
                __syncthreads();
                for (int q = e0 + tid; q < e1; q += NT) {
                    const ExclPair ex = a.excl[q];
                    if (ex.u == 1.0) continue;
                    const int i = ex.i - o0, k = ex.k - o0;
                    const double xr = px[k] - px[i], yr = py[k] - py[i], zr = pz[k] - pz[i];
                    const double r2 = xr * xr + yr * yr + zr * zr;
                    if (near && r2 > a.us2) continue;
                    double lam[6], R[6];
                    thole_lambda<double>(sqrt(r2), a.pdm[ex.i], a.pdm[ex.k], a.thl[ex.i], a.thl[ex.k], lam);
                    corr_series<double>(r2, ex.u, lam, R);
                    const double rv[3] = {xr, yr, zr};
                    for (int set = 0; set < 2; set++) {
                        const double *vi = nullptr;
                        const double ui[3] = {vec[3 * set][i], vec[3 * set + 1][i], vec[3 * set + 2][i]};
                        const double uk[3] = {vec[3 * set][k], vec[3 * set + 1][k], vec[3 * set + 2][k]};
                        (void)vi;
                        const double uir = ui[0] * xr + ui[1] * yr + ui[2] * zr, ukr = uk[0] * xr + uk[1] * yr + uk[2] * zr;
                        for (int c = 0; c < 3; c++) {
                            acc.add(3 * set + c, i, ukr * R[2] * rv[c] - R[1] * uk[c], fx_field);
                            acc.add(3 * set + c, k, uir * R[2] * rv[c] - R[1] * ui[c], fx_field);
                        }
                    }
                }
                __syncthreads();
                for (int c = 0; c < 6; c++) tv[c] = rowon ? s[c] + acc.get(c, row, fx_field) : 0.0;
            };
            // z = P x for this row, x = the row's residual (both sets)
            double tv[6];
            auto precond = [&](const double *x, double *z) {
                if (!a.prec) {
                    for (int c = 0; c < 6; c++) z[c] = alpha * x[c];
                    return;
                }
                __syncthreads();
                if (own)
                    for (int c = 0; c < 6; c++) vec[c][row] = alpha * x[c];
                __syncthreads();
                apply_T(tv, true);
                for (int c = 0; c < 6; c++) z[c] = a.udiag * (alpha * x[c]) + fmax(1e-8, alpha) * tv[c];
            };
            // with the sparse preconditioner the first guess takes the near
            // field of the direct dipoles once: u = a E + a T_s (a E)
            if (a.prec) {
                __syncthreads();
                if (own)
                    for (int c = 0; c < 6; c++) vec[c][row] = u[c];
                __syncthreads();
                apply_T(tv, true);
                for (int c = 0; c < 6; c++) u[c] = alpha != 0.0 ? u[c] + fmax(1e-8, alpha) * tv[c] : 0.0;
            }
            // r = E + T u - u / alpha, z = P r, p = z
            __syncthreads();
            if (own)
                for (int c = 0; c < 6; c++) vec[c][row] = u[c];
            __syncthreads();
            double r[6], z[6], p[6];
            apply_T(tv, false);
            for (int c = 0; c < 6; c++) r[c] = alpha != 0.0 ? E6[c] + tv[c] - u[c] / alpha : 0.0;
            precond(r, z);
            for (int c = 0; c < 6; c++) p[c] = z[c];
            double rz[2];
            block_sum2<NT>(own ? r[0] * z[0] + r[1] * z[1] + r[2] * z[2] : 0.0,
                           own ? r[3] * z[3] + r[4] * z[4] + r[5] * z[5] : 0.0, red, rz[0], rz[1]);
            bool done = false;
            int it = 0;
            while (!done && it < a.politer) {
                it++;
                __syncthreads();
                if (own)
                    for (int c = 0; c < 6; c++) vec[c][row] = p[c];
                __syncthreads();
                apply_T(tv, false);
                double v[6];
                for (int c = 0; c < 6; c++) v[c] = alpha != 0.0 ? p[c] / alpha - tv[c] : 0.0;
                double pv[2];
                block_sum2<NT>(own ? p[0] * v[0] + p[1] * v[1] + p[2] * v[2] : 0.0,
                               own ? p[3] * v[3] + p[4] * v[4] + p[5] * v[5] : 0.0, red, pv[0], pv[1]);
                const double al[2] = {pv[0] != 0.0 ? rz[0] / pv[0] : 0.0, pv[1] != 0.0 ? rz[1] / pv[1] : 0.0};
                for (int c = 0; c < 6; c++) {
                    const double s = al[c / 3];
                    u[c] += s * p[c];
                    r[c] -= s * v[c];
                }
                precond(r, z);
                double rzn[2], rr[2];
                block_sum2<NT>(own ? r[0] * z[0] + r[1] * z[1] + r[2] * z[2] : 0.0,
                               own ? r[3] * z[3] + r[4] * z[4] + r[5] * z[5] : 0.0, red, rzn[0], rzn[1]);
                block_sum2<NT>(own ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : 0.0,
                               own ? r[3] * r[3] + r[4] * r[4] + r[5] * r[5] : 0.0, red, rr[0], rr[1]);
                const double eps = debye * sqrt(fmax(rr[0], rr[1]) / npol);
                if (eps < a.poleps) {
                    done = true;
                    break;
                }
                const double be[2] = {rz[0] != 0.0 ? rzn[0] / rz[0] : 0.0, rz[1] != 0.0 ? rzn[1] / rz[1] : 0.0};
                for (int c = 0; c < 6; c++) p[c] = z[c] + be[c / 3] * p[c];
                rz[0] = rzn[0];
                rz[1] = rzn[1];
            }
            iters = done ? it : -it;
            // E_pol = -(f/2) sum mu_d . E^p
            double s1, s2;
            block_sum2<NT>(own ? -0.5 * a.f * (u[0] * Ep[0] + u[1] * Ep[1] + u[2] * Ep[2]) : 0.0, 0.0, red, s1, s2);
            epc = s1;
            // the dipoles go to LDS for the gradient
            __syncthreads();
            if (own)
                for (int c = 0; c < 6; c++) vec[c][row] = u[c];
            __syncthreads();
        }

        // ---- multipole energy and gradient, polarization gradient
        double fm[3] = {0, 0, 0}, tm[3] = {0, 0, 0}, fp[3] = {0, 0, 0}, tp[3] = {0, 0, 0}, er = 0.0;
        if (rowon) {
            T udi[3], upi[3];
            for (int c = 0; c < 3; c++) { udi[c] = T(vec[c][row]); upi[c] = T(vec[3 + c][row]); }
            for (int k = part; k < nc; k += TPR) {
                if (k == row) continue;
                const double xd = px[k] - xi, yd = py[k] - yi, zd = pz[k] - zi;
                const T xr = T(xd), yr = T(yd), zr = T(zd);
                const T r2 = xr * xr + yr * yr + zr * zr;
                const Mp<T> mk = mp_at<T, ROWS>(mp, k);
                if (a.do_mpole) {
                    T Rs[6];
                    rr_series<T>(r2, T(a.f), Rs);
                    PairOut<T> o;
                    if (a.grad) pair_series<T, true>(Rs, mi, mk, xr, yr, zr, o);
                    else pair_series<T, false>(Rs, mi, mk, xr, yr, zr, o);
                    er += 0.5 * (double)o.e;
                    if (a.grad) {
                        fm[0] += o.fx; fm[1] += o.fy; fm[2] += o.fz;
                        tm[0] += o.tix; tm[1] += o.tiy; tm[2] += o.tiz;
                    }
                }
                if (a.do_polar && a.grad) {
                    T lam[6], R[6];
                    thole_lambda<T>(sqrt(r2), pdi, spd[k], thi, sth[k], lam);
                    pol_series<false, T>(r2, T(0), lam, R);
                    const T udk[3] = {T(vec[0][k]), T(vec[1][k]), T(vec[2][k])};
                    const T upk[3] = {T(vec[3][k]), T(vec[4][k]), T(vec[5][k])};
                    T F[3] = {0, 0, 0}, Ti[3] = {0, 0, 0}, Tk[3] = {0, 0, 0};
                    polgrad_pair<T>(R, mi, mk, udi, upi, udk, upk, xr, yr, zr, T(1), T(0), T(0), T(1), F, Ti, Tk);
                    for (int c = 0; c < 3; c++) { fp[c] += F[c]; tp[c] += Ti[c]; }
                }
            }
        }
        er = row_sum<TPR>(er);
        for (int c = 0; c < 3; c++) {
            fm[c] = row_sum<TPR>(fm[c]); tm[c] = row_sum<TPR>(tm[c]);
            fp[c] = row_sum<TPR>(fp[c]); tp[c] = row_sum<TPR>(tp[c]);
        }
        // scaled-pair corrections: multipole (m - 1) and polarization (d, p, u - 1)
        acc.clear(tid, NT);
        __syncthreads();
        double ecorr = 0.0;
        for (int q = e0 + tid; q < e1; q += NT) {
            const ExclPair ex = a.excl[q];
            const int i = ex.i - o0, k = ex.k - o0;
            const double xr = px[k] - px[i], yr = py[k] - py[i], zr = pz[k] - pz[i];
            const double r2 = xr * xr + yr * yr + zr * zr;
            Mp<double> m1, m2;
            m1.c = mp[0][i]; m1.dx = mp[1][i]; m1.dy = mp[2][i]; m1.dz = mp[3][i]; m1.qxx = mp[4][i];
            m1.qxy = mp[5][i]; m1.qxz = mp[6][i]; m1.qyy = mp[7][i]; m1.qyz = mp[8][i]; m1.qzz = mp[9][i];
            m2.c = mp[0][k]; m2.dx = mp[1][k]; m2.dy = mp[2][k]; m2.dz = mp[3][k]; m2.qxx = mp[4][k];
            m2.qxy = mp[5][k]; m2.qxz = mp[6][k]; m2.qyy = mp[7][k]; m2.qyz = mp[8][k]; m2.qzz = mp[9][k];
            if (FX) { // FP64 multipoles for the corrections
                const size_t np = a.npad;
                m1.c = a.smp[ex.i]; m1.dx = a.smp[np + ex.i]; m1.dy = a.smp[2 * np + ex.i]; m1.dz = a.smp[3 * np + ex.i];
                m1.qxx = a.smp[4 * np + ex.i]; m1.qxy = a.smp[5 * np + ex.i]; m1.qxz = a.smp[6 * np + ex.i];
                m1.qyy = a.smp[7 * np + ex.i]; m1.qyz = a.smp[8 * np + ex.i]; m1.qzz = a.smp[9 * np + ex.i];
                m2.c = a.smp[ex.k]; m2.dx = a.smp[np + ex.k]; m2.dy = a.smp[2 * np + ex.k]; m2.dz = a.smp[3 * np + ex.k];
                m2.qxx = a.smp[4 * np + ex.k]; m2.qxy = a.smp[5 * np + ex.k]; m2.qxz = a.smp[6 * np + ex.k];
                m2.qyy = a.smp[7 * np + ex.k]; m2.qyz = a.smp[8 * np + ex.k]; m2.qzz = a.smp[9 * np + ex.k];
            }
            if (a.do_mpole && ex.m != 1.0) {
                double Rs[6];
                rr_series<double>(r2, (ex.m - 1.0) * a.f, Rs);
                PairOut<double> o;
                pair_series<double, true>(Rs, m1, m2, xr, yr, zr, o);
                ecorr += o.e;
                if (a.grad) {
                    acc.add(0, i, o.fx, fx_grad); acc.add(1, i, o.fy, fx_grad); acc.add(2, i, o.fz, fx_grad);
                    acc.add(0, k, -o.fx, fx_grad); acc.add(1, k, -o.fy, fx_grad); acc.add(2, k, -o.fz, fx_grad);
                    acc.add(3, i, o.tix, fx_grad); acc.add(4, i, o.tiy, fx_grad); acc.add(5, i, o.tiz, fx_grad);
                    acc.add(3, k, o.tkx, fx_grad); acc.add(4, k, o.tky, fx_grad); acc.add(5, k, o.tkz, fx_grad);
                }
            }
        }
        {
            double s1, s2;
            block_sum2<NT>((own ? er : 0.0) + ecorr, 0.0, red, s1, s2);
            emc = s1;
        }
        if (a.grad && own) {
            const int g = o0 + row;
            for (int c = 0; c < 3; c++) {
                a.dem[3 * (size_t)g + c] = fm[c] + acc.get(c, row, fx_grad);
                a.te[3 * (size_t)g + c] = tm[c] + acc.get(3 + c, row, fx_grad);
            }
        }
        if (a.do_polar && a.grad) {
            __syncthreads(); // every lane has read the multipole corrections
            acc.clear(tid, NT);
            __syncthreads();
            for (int q = e0 + tid; q < e1; q += NT) {
                const ExclPair ex = a.excl[q];
                if (ex.d == 1.0 && ex.p == 1.0 && ex.u == 1.0) continue;
                const int i = ex.i - o0, k = ex.k - o0;
                const double xr = px[k] - px[i], yr = py[k] - py[i], zr = pz[k] - pz[i];
                const double r2 = xr * xr + yr * yr + zr * zr;
                double lam[6], rr[6], R[6];
                thole_lambda<double>(sqrt(r2), a.pdm[ex.i], a.pdm[ex.k], a.thl[ex.i], a.thl[ex.k], lam);
                rr_series<double>(r2, 1.0, rr);
                for (int c = 0; c < 6; c++) R[c] = lam[c] * rr[c];
                const size_t np = a.npad;
                Mp<double> m1, m2;
                m1.c = a.smp[ex.i]; m1.dx = a.smp[np + ex.i]; m1.dy = a.smp[2 * np + ex.i]; m1.dz = a.smp[3 * np + ex.i];
                m1.qxx = a.smp[4 * np + ex.i]; m1.qxy = a.smp[5 * np + ex.i]; m1.qxz = a.smp[6 * np + ex.i];
                m1.qyy = a.smp[7 * np + ex.i]; m1.qyz = a.smp[8 * np + ex.i]; m1.qzz = a.smp[9 * np + ex.i];
                m2.c = a.smp[ex.k]; m2.dx = a.smp[np + ex.k]; m2.dy = a.smp[2 * np + ex.k]; m2.dz = a.smp[3 * np + ex.k];
                m2.qxx = a.smp[4 * np + ex.k]; m2.qxy = a.smp[5 * np + ex.k]; m2.qxz = a.smp[6 * np + ex.k];
                m2.qyy = a.smp[7 * np + ex.k]; m2.qyz = a.smp[8 * np + ex.k]; m2.qzz = a.smp[9 * np + ex.k];
                const double udi[3] = {vec[0][i], vec[1][i], vec[2][i]}, upi[3] = {vec[3][i], vec[4][i], vec[5][i]};
                const double udk[3] = {vec[0][k], vec[1][k], vec[2][k]}, upk[3] = {vec[3][k], vec[4][k], vec[5][k]};
                double F[3] = {0, 0, 0}, Ti[3] = {0, 0, 0}, Tk[3] = {0, 0, 0};
                polgrad_pair<double>(R, m1, m2, udi, upi, udk, upk, xr, yr, zr, 0.0, ex.d - 1.0, ex.p - 1.0,
                                     ex.u - 1.0, F, Ti, Tk);
                for (int c = 0; c < 3; c++) {
                    acc.add(c, i, F[c], fx_grad);
                    acc.add(c, k, -F[c], fx_grad);
                    acc.add(3 + c, i, Ti[c], fx_grad);
                    acc.add(3 + c, k, Tk[c], fx_grad);
                }
            }
            __syncthreads();
            if (own) {
                const int g = o0 + row;
                const double hf = 0.5 * a.f;
                for (int c = 0; c < 3; c++) {
                    a.dep[3 * (size_t)g + c] = hf * (fp[c] + acc.get(c, row, fx_grad));
                    a.tep[3 * (size_t)g + c] = hf * (tp[c] + acc.get(3 + c, row, fx_grad));
                }
            }
        }
        if (a.pgrad) {
            // dE/d(parameters) of every atom of the cluster (param_pair.h):
            // row-owned sums over all partners at scale 1, then the scaled
            // pairs' corrections through the LDS accumulator in two passes
            // (components 0-5; components 6-9, pdamp, Thole)
            double g[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, gpd = 0.0, gth = 0.0;
            if (rowon) {
                T udi[3], upi[3];
                for (int c = 0; c < 3; c++) {
                    udi[c] = a.do_polar ? T(vec[c][row]) : T(0);
                    upi[c] = a.do_polar ? T(vec[3 + c][row]) : T(0);
                }
                for (int k = part; k < nc; k += TPR) {
                    if (k == row) continue;
                    const double xd = px[k] - xi, yd = py[k] - yi, zd = pz[k] - zi;
                    const T xr = T(xd), yr = T(yd), zr = T(zd);
                    const T r2 = xr * xr + yr * yr + zr * zr;
                    T rr[6], Rm[6], R[6], dl[6], D = T(0), pg = T(0);
                    rr_series<T>(r2, T(1), rr);
                    for (int q = 0; q < 6; q++) Rm[q] = T(a.f) * rr[q];
                    if (a.do_polar) {
                        T lam[6];
                        thole_lambda<T>(sqrt(r2), pdi, spd[k], thi, sth[k], lam);
                        for (int q = 0; q < 6; q++) R[q] = lam[q] * rr[q];
                        thole_dlambda<T>(sqrt(r2), pdi, spd[k], thi, sth[k], dl, &D, &pg);
                    } else {
                        for (int q = 0; q < 6; q++) R[q] = dl[q] = T(0);
                    }
                    const Mp<T> mk = mp_at<T, ROWS>(mp, k);
                    T udk[3], upk[3];
                    for (int c = 0; c < 3; c++) {
                        udk[c] = a.do_polar ? T(vec[c][k]) : T(0);
                        upk[c] = a.do_polar ? T(vec[3 + c][k]) : T(0);
                    }
                    double dEdD = 0.0;
                    param_pair<T, false>(a.do_mpole, a.do_polar, a.f, xr, yr, zr, mi, mk, Rm, R, rr, dl, T(1), T(1),
                                         T(1), T(1), udi, upi, udk, upk, g, nullptr, &dEdD);
                    if (D > T(0)) {
                        gpd += dEdD * (-3.0 * (double)D / (double)pdi);
                        gth += dEdD * ((double)D / (double)pg) * (double)thole_share<T>(thi, sth[k]);
                    }
                }
            }
            for (int c = 0; c < 10; c++) g[c] = row_sum<TPR>(g[c]);
            gpd = row_sum<TPR>(gpd);
            gth = row_sum<TPR>(gth);
            const size_t np = a.npad;
            for (int pass = 0; pass < 2; pass++) {
                __syncthreads();
                acc.clear(tid, NT);
                __syncthreads();
                for (int q = e0 + tid; q < e1; q += NT) {
                    const ExclPair ex = a.excl[q];
                    if (ex.m == 1.0 && ex.d == 1.0 && ex.p == 1.0 && ex.u == 1.0) continue;
                    const int i = ex.i - o0, k = ex.k - o0;
                    const double xr = px[k] - px[i], yr = py[k] - py[i], zr = pz[k] - pz[i];
                    const double r2 = xr * xr + yr * yr + zr * zr;
                    double rr[6], Rm[6], R[6], dl[6], D = 0.0, pg = 0.0;
                    rr_series<double>(r2, 1.0, rr);
                    for (int c = 0; c < 6; c++) Rm[c] = a.f * rr[c];
                    const double pdi2 = a.pdm[ex.i], pdk2 = a.pdm[ex.k], thi2 = a.thl[ex.i], thk2 = a.thl[ex.k];
                    if (a.do_polar) {
                        double lam[6];
                        thole_lambda<double>(sqrt(r2), pdi2, pdk2, thi2, thk2, lam);
                        for (int c = 0; c < 6; c++) R[c] = lam[c] * rr[c];
                        thole_dlambda<double>(sqrt(r2), pdi2, pdk2, thi2, thk2, dl, &D, &pg);
                    } else {
                        for (int c = 0; c < 6; c++) R[c] = dl[c] = 0.0;
                    }
                    Mp<double> m1, m2;
                    for (int c = 0; c < 10; c++) {
                        mp_set(m1, c, a.smp[c * np + ex.i]);
                        mp_set(m2, c, a.smp[c * np + ex.k]);
                    }
                    double udi[3], upi[3], udk[3], upk[3];
                    for (int c = 0; c < 3; c++) {
                        udi[c] = a.do_polar ? vec[c][i] : 0.0; upi[c] = a.do_polar ? vec[3 + c][i] : 0.0;
                        udk[c] = a.do_polar ? vec[c][k] : 0.0; upk[c] = a.do_polar ? vec[3 + c][k] : 0.0;
                    }
                    double gi[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, gk[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                    double dEdD = 0.0;
                    param_pair<double, true>(a.do_mpole, a.do_polar, a.f, xr, yr, zr, m1, m2, Rm, R, rr, dl, ex.m - 1.0,
                                             ex.d - 1.0, ex.p - 1.0, ex.u - 1.0, udi, upi, udk, upk, gi, gk, &dEdD);
                    if (pass == 0) {
                        for (int c = 0; c < 6; c++) {
                            acc.add(c, i, gi[c], fx_grad);
                            acc.add(c, k, gk[c], fx_grad);
                        }
                    } else {
                        for (int c = 0; c < 4; c++) {
                            acc.add(c, i, gi[6 + c], fx_grad);
                            acc.add(c, k, gk[6 + c], fx_grad);
                        }
                        if (D > 0.0) {
                            acc.add(4, i, dEdD * (-3.0 * D / pdi2), fx_grad);
                            acc.add(4, k, dEdD * (-3.0 * D / pdk2), fx_grad);
                            acc.add(5, i, dEdD * (D / pg) * thole_share<double>(thi2, thk2), fx_grad);
                            acc.add(5, k, dEdD * (D / pg) * thole_share<double>(thk2, thi2), fx_grad);
                        }
                    }
                }
                __syncthreads();
                if (own) {
                    if (pass == 0) {
                        for (int c = 0; c < 6; c++) g[c] += acc.get(c, row, fx_grad);
                    } else {
                        for (int c = 0; c < 4; c++) g[6 + c] += acc.get(c, row, fx_grad);
                        gpd += acc.get(4, row, fx_grad);
                        gth += acc.get(5, row, fx_grad);
                    }
                }
            }
            if (own) {
                const size_t gidx = (size_t)(o0 + row);
                for (int c = 0; c < 10; c++) a.pg_dm[c * np + gidx] = g[c];
                a.pg_dpd[gidx] = gpd;
                a.pg_dth[gidx] = gth;
                for (int c = 0; c < 6; c++) a.uout[c * np + gidx] = a.do_polar ? vec[c][row] : 0.0;
            }
        }
        if (tid == 0) {
            a.em_c[cl] = a.do_mpole ? emc : 0.0;
            a.ep_c[cl] = epc;
            a.it_c[cl] = (double)iters;
        }
    }
}

// per-cluster energies into the totals, in cluster order
__global__ void k_batch_totals(int nsys, const double *em_c, const double *ep_c, Totals *tot)
{
    __shared__ double red[2][256];
    double a = 0.0, b = 0.0;
    for (int c = threadIdx.x; c < nsys; c += 256) { a += em_c[c]; b += ep_c[c]; }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tot->em = red[0][0];
        tot->ep = red[1][0];
    }
}

__global__ void k_batch_reset(int *counter)
{
    if (threadIdx.x < PMQ_BATCH_BUCKETS) counter[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// host
//
// Clusters are grouped by size into buckets of up to 256, 128 and 64 atoms
// (the workgroup's row count), each bucket one persistent launch over its
// own slice of the largest-first order.  A shape is (rows, lanes per row,
// workgroups per CU); the autotuner times every shape a bucket admits on
// that bucket and keeps the fastest, per precision.

static const int kBucketRows[PMQ_BATCH_BUCKETS] = {256, 128, 64};
static const int kLanes[PMQ_BATCH_BUCKETS][4] = {{1, 2, 0, 0}, {1, 2, 4, 0}, {1, 2, 4, 8}};
static const int kWgPerCu[3] = {1, 2, 4};

int batch_bucket_rows(int b) { return kBucketRows[b]; }

static int lane_options(int b)
{
    int n = 0;
    while (n < 4 && kLanes[b][n]) n++;
    return n;
}

int batch_variants(int b) { return lane_options(b) * 3; }

void batch_shape(int b, int v, int *rows, int *lanes, int *wpc)
{
    const int nl = lane_options(b);
    *rows = kBucketRows[b];
    *lanes = kLanes[b][v % nl];
    *wpc = kWgPerCu[v / nl];
}

template <class T, bool FX>
static void launch_shape(int rows, int tpr, int wg, hipStream_t st, const BatchArgs &a)
{
    const dim3 g(wg), t(rows * tpr);
    switch (rows * 16 + tpr) {
    case 256 * 16 + 1: hipLaunchKernelGGL((k_batch_clusters<T, 256, 1, FX>), g, t, 0, st, a); break;
    case 256 * 16 + 2: hipLaunchKernelGGL((k_batch_clusters<T, 256, 2, FX>), g, t, 0, st, a); break;
    case 128 * 16 + 1: hipLaunchKernelGGL((k_batch_clusters<T, 128, 1, FX>), g, t, 0, st, a); break;
    case 128 * 16 + 2: hipLaunchKernelGGL((k_batch_clusters<T, 128, 2, FX>), g, t, 0, st, a); break;
    case 128 * 16 + 4: hipLaunchKernelGGL((k_batch_clusters<T, 128, 4, FX>), g, t, 0, st, a); break;
    case 64 * 16 + 1: hipLaunchKernelGGL((k_batch_clusters<T, 64, 1, FX>), g, t, 0, st, a); break;
    case 64 * 16 + 2: hipLaunchKernelGGL((k_batch_clusters<T, 64, 2, FX>), g, t, 0, st, a); break;
    case 64 * 16 + 4: hipLaunchKernelGGL((k_batch_clusters<T, 64, 4, FX>), g, t, 0, st, a); break;
    default: hipLaunchKernelGGL((k_batch_clusters<T, 64, 8, FX>), g, t, 0, st, a); break;
    }
}

static BatchArgs batch_args(pmq_ctx *c, int mode, bool mpole, bool polar, const BatchScratch *bs, bool pgrad)
{
    BatchArgs a;
    a.nsys = c->nsys;
    a.order = c->b_order.p; a.off = c->b_off.p; a.eoff = c->b_eoff.p;
    a.excl = c->excl.p;
    a.x = c->x.p; a.y = c->y.p; a.z = c->z.p;
    a.smp = c->smp.p; a.npad = c->npad;
    a.pol = c->polarity.p; a.thl = c->thole.p; a.pdm = c->pdamp.p;
    a.useflag = nullptr;
    a.f = c->prm.electric / c->prm.dielec;
    a.poleps = c->prm.poleps;
    a.politer = c->prm.politer;
    a.prec = c->prm.pcgprec && c->prm.usolvcut > 0.0 ? 1 : 0;
    a.us2 = c->prm.usolvcut * c->prm.usolvcut;
    a.udiag = c->prm.udiag;
    a.grad = mode == PMQ_GRADIENT;
    a.do_mpole = mpole; a.do_polar = polar;
    a.dem = c->dem.p; a.te = c->te.p; a.dep = c->dep.p; a.tep = c->tep.p;
    a.em_c = bs ? bs->em_c : c->b_em.p;
    a.ep_c = bs ? bs->ep_c : c->b_ep.p;
    a.it_c = bs ? bs->it_c : c->b_it.p;
    a.counter = bs ? bs->counter : c->b_counter.p;
    a.tot = bs ? bs->tot : c->totals.p;
    a.pgrad = pgrad && mode == PMQ_GRADIENT;
    a.pg_dm = c->pg_dm.p; a.pg_dpd = c->pg_dpd.p; a.pg_dth = c->pg_dth.p;
    a.uout = c->uind.p;
    return a;
}

// one bucket's persistent launch with shape v on stream st (its counter already zero)
static void launch_bucket(pmq_ctx *c, BatchArgs a, int b, int v, bool mixed, hipStream_t st, bool merged)
{
    const int n = merged ? c->nsys : c->b_beg[b + 1] - c->b_beg[b];
    int rows, tpr, wpc;
    batch_shape(b, v, &rows, &tpr, &wpc);
    a.order = c->b_order.p + (merged ? 0 : c->b_beg[b]);
    a.nsys = n;
    a.counter += b;
    const int wg = std::max(1, std::min(n, c->num_cus * wpc));
    if (mixed) launch_shape<float, true>(rows, tpr, wg, st, a);
    else launch_shape<double, false>(rows, tpr, wg, st, a);
}

// every bucket's launch, the buckets side by side: the first on the
// context's stream, the others on their own streams forked from it and
// joined back (the same in a captured graph).  merged: all clusters in one
// launch with the first bucket's workgroups (256 rows).
static int launch_buckets(pmq_ctx *c, const BatchArgs &a, const int *variants, bool mixed, bool merged)
{
    int act[PMQ_BATCH_BUCKETS], na = 0;
    for (int b = 0; b < PMQ_BATCH_BUCKETS; b++)
        if (merged ? b == 0 : c->b_beg[b + 1] > c->b_beg[b]) act[na++] = b;
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, c->stream, a.counter);
    if (na > 1) PMQ_HIP(hipEventRecord(c->b_fork, c->stream));
    for (int k = 1; k < na; k++) {
        PMQ_HIP(hipStreamWaitEvent(c->b_side[k - 1], c->b_fork, 0));
        launch_bucket(c, a, act[k], variants[act[k]], mixed, c->b_side[k - 1], false);
        PMQ_HIP(hipEventRecord(c->b_join[k - 1], c->b_side[k - 1]));
    }
    if (na > 0) launch_bucket(c, a, act[0], variants[act[0]], mixed, c->stream, merged);
    for (int k = 1; k < na; k++) PMQ_HIP(hipStreamWaitEvent(c->stream, c->b_join[k - 1], 0));
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int launch_batch(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed, const int *variants, const BatchScratch *bs,
                 bool pgrad, bool merged)
{
    const BatchArgs a = batch_args(c, mode, mpole, polar, bs, pgrad);
    int rc;
    {
        ProfScope ps_(c, "batch_clusters");
        if ((rc = launch_buckets(c, a, variants, mixed, merged))) return rc;
    }
    hipLaunchKernelGGL(k_batch_totals, dim3(1), dim3(256), 0, c->stream, c->nsys, a.em_c, a.ep_c, a.tot);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// The autotuner: coordinate descent over the buckets (largest clusters
// first), each bucket's shapes timed on the whole batch with the other
// buckets at their current shapes (the buckets run side by side, so one
// bucket's best shape depends on the others'); then every shape of one
// merged launch of all clusters in 256-row workgroups, kept if faster.  The
// second of two launches is timed.  Records the choice in c->b_var, c->b_ms
// and c->b_merged for this precision; shapes fixed by pmq_set_batch_shape are
// kept (and rule out the merged launch).
int batch_tune(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed, bool pgrad)
{
    const int p = mixed ? 1 : 0;
    int rc;
    if ((rc = launch_rotpole(c, false))) return rc;
    const BatchArgs a = batch_args(c, mode, mpole, polar, nullptr, pgrad);
    hipEvent_t e0, e1;
    PMQ_HIP(hipEventCreate(&e0));
    PMQ_HIP(hipEventCreate(&e1));
    int var[PMQ_BATCH_BUCKETS];
    bool pinned = false;
    for (int b = 0; b < PMQ_BATCH_BUCKETS; b++) {
        var[b] = c->b_pin[p][b] ? c->b_var[p][b] : 0;
        pinned = pinned || c->b_pin[p][b];
    }
    auto timed = [&](bool merged, float *ms) -> int {
        for (int rep = 0; rep < 2; rep++) {
            if (rep) PMQ_HIP(hipEventRecord(e0, c->stream));
            int r = launch_buckets(c, a, var, mixed, merged);
            if (r) return r;
            if (rep) PMQ_HIP(hipEventRecord(e1, c->stream));
        }
        PMQ_HIP(hipEventSynchronize(e1));
        PMQ_HIP(hipEventElapsedTime(ms, e0, e1));
        return PMQ_OK;
    };
    float split = 1e30f;
    for (int b = 0; b < PMQ_BATCH_BUCKETS; b++) {
        if (c->b_pin[p][b]) continue;
        c->b_ms[p][b] = -1.0f;
        if (c->b_beg[b + 1] == c->b_beg[b]) continue;
        float best = 1e30f;
        int bestv = 0;
        for (int v = 0; v < batch_variants(b); v++) {
            var[b] = v;
            float ms = 0.0f;
            if ((rc = timed(false, &ms))) return rc;
            if (ms < best) {
                best = ms;
                bestv = v;
            }
        }
        var[b] = bestv;
        c->b_ms[p][b] = best;
        split = best;
    }
    bool merged = false;
    if (!pinned && c->nsys > 0) {
        const int keep0 = var[0];
        float best = split;
        int bestv = -1;
        for (int v = 0; v < batch_variants(0); v++) {
            var[0] = v;
            float ms = 0.0f;
            if ((rc = timed(true, &ms))) return rc;
            if (ms < best) {
                best = ms;
                bestv = v;
            }
        }
        var[0] = keep0;
        if (bestv >= 0) {
            merged = true;
            var[0] = bestv;
            for (int b = 0; b < PMQ_BATCH_BUCKETS; b++) c->b_ms[p][b] = b == 0 ? best : -1.0f;
        }
    }
    for (int b = 0; b < PMQ_BATCH_BUCKETS; b++) c->b_var[p][b] = var[b];
    c->b_merged[p] = merged;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return PMQ_OK;
}

// ---------------------------------------------------------------------------
// The allocation-free batched evaluation (pmq_compute_batch_noalloc) keeps
// its scratch -- the work-queue counter, the per-cluster energies and
// iteration counts, the energy totals -- in this fixed device array instead
// of context buffers.  Each GPU holds its own copy for the life of the
// process; its device address is taken once per device when a context is
// created (static_scratch_attach), and kernels are only ever given that
// address.  Every call on one GPU shares the array: calls that overlap on
// one GPU race.
struct StaticScratchLayout {
    Totals tot;
    int counter[PMQ_BATCH_BUCKETS];
    int pad[64 - PMQ_BATCH_BUCKETS];
    double em_c[PMQ_NOALLOC_MAX_CLUSTERS];
    double ep_c[PMQ_NOALLOC_MAX_CLUSTERS];
    double it_c[PMQ_NOALLOC_MAX_CLUSTERS];
};
static __device__ __align__(256) unsigned char scratch[sizeof(StaticScratchLayout)];

static std::mutex scratch_mu;
static void *scratch_dev[PMQ_NOALLOC_MAX_DEVICES] = {};

size_t static_scratch_bytes() { return sizeof(StaticScratchLayout); }

// the calling thread's current device must be `device`
int static_scratch_attach(pmq_ctx *c, int device)
{
    if (device < 0 || device >= PMQ_NOALLOC_MAX_DEVICES) {
        c->err = "device ordinal beyond the static scratch table";
        return PMQ_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[device]) {
        void *p = nullptr;
        PMQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            c->err = "hipGetSymbolAddress returned no address for the static scratch";
            return PMQ_ERR_HIP;
        }
        scratch_dev[device] = p;
    }
    c->static_scratch = scratch_dev[device];
    return PMQ_OK;
}

// results of the allocation-free evaluation into the caller's device arrays
__global__ void k_batch_static_out(int nsys, int n3, const double *em_c, const double *ep_c, const double *it_c,
                                   const Totals *tot, double *em, double *ep, int *iters, double *etot,
                                   const double *dem, const double *dep, int mp, int pol, double *grad)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int k = t; k < nsys; k += stride) {
        if (em) em[k] = em_c[k];
        if (ep) ep[k] = ep_c[k];
        if (iters) iters[k] = (int)it_c[k];
    }
    if (grad)
        for (int j = t; j < n3; j += stride) grad[j] = (mp ? dem[j] : 0.0) + (pol ? dep[j] : 0.0);
    if (etot && t == 0) {
        etot[0] = tot->em;
        etot[1] = tot->ep;
    }
}

int batch_noalloc_enqueue(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed, const int *variants, bool merged,
                          double *em, double *ep, int *iters, double *etot, double *grad)
{
    // device addresses inside this device's array (from static_scratch_attach)
    unsigned char *base = (unsigned char *)c->static_scratch;
    BatchScratch bs;
    bs.tot = (Totals *)(base + offsetof(StaticScratchLayout, tot));
    bs.counter = (int *)(base + offsetof(StaticScratchLayout, counter));
    bs.em_c = (double *)(base + offsetof(StaticScratchLayout, em_c));
    bs.ep_c = (double *)(base + offsetof(StaticScratchLayout, ep_c));
    bs.it_c = (double *)(base + offsetof(StaticScratchLayout, it_c));
    const bool grad_mode = mode == PMQ_GRADIENT;
    int rc;
    // the totals start at zero (a fill kernel: nothing here records a memset)
    if ((rc = launch_zero(c, (double *)bs.tot, sizeof(Totals) / sizeof(double)))) return rc;
    if ((rc = launch_rotpole(c, false))) return rc;
    if ((rc = launch_batch(c, mode, mpole, polar, mixed, variants, &bs, false, merged))) return rc;
    if (grad_mode) {
        if (mpole && (rc = launch_torque(c, PMQ_GRADIENT, c->te.p, c->dem.p, mixed))) return rc;
        if (polar && (rc = launch_torque(c, PMQ_GRADIENT, c->tep.p, c->dep.p, mixed))) return rc;
    }
    const int n3 = 3 * c->n;
    const int work = std::max(c->nsys, grad_mode && grad ? n3 : 0);
    const int blocks = std::max(1, std::min((work + 255) / 256, 1024));
    hipLaunchKernelGGL(k_batch_static_out, dim3(blocks), dim3(256), 0, c->stream, c->nsys, n3, bs.em_c, bs.ep_c,
                       bs.it_c, bs.tot, em, ep, iters, etot, c->dem.p, c->dep.p, mpole ? 1 : 0, polar ? 1 : 0,
                       grad_mode ? grad : nullptr);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

} // namespace pmq
