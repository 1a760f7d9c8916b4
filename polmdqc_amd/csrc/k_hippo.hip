// k_hippo.hip -- HIPPO damped dispersion and Pauli repulsion.
//
// The reference carries the parameters of both terms and no energy:
// dispersion is absent (energy.cpp:280; dspsix/dspdmp per class at
// kdsp.h:18-19, dsp2..5scale at dsppot.h:21-25, DISPERSION at
// readprm.cpp:1047); erepel_a is an empty pair loop after chkrepole,
// rotpole(Repel) and cutoffSwitch(Repuls) (erepel.cpp:42-153; sizpr, dmppr,
// elepr, r2..5scale at repel.h:25-32, reppot.h:20-23, krepel.cpp:92-153).
//
// Dispersion (Rackers et al., JCP 149, 084115 (2018)):
//   E_ik = -s C_i C_k f^2 / r^6,  f = 1.5 d5 - 0.5 d3,
// d3, d5 the overlap damping factors of the two adisp exponents (the
// valence-valence factors of the charge-penetration model).  With
// dispersion Ewald (coefficient a) the real-space pair is
// -C_i C_k [s f^2 - 1 + g(a r)] / r^6, g(x) = (1 + x^2 + x^4/2) e^{-x^2}; the
// reciprocal part and the self term a^6/12 C^2 are in k_pme.hip.  Without
// Ewald the pair is switched off between disptaper and dispcut.
//
// Repulsion (Rackers and Ponder, JCP 150, 084104 (2019)):
//   E_ik = K_i K_k s / r * M(B),
// M the multipole pair contraction (monopole = elepr, the valence electron
// count; the rotated dipoles and quadrupoles) over the radial series
// B_n = D^n S^2 (rep_series in pair_mpole.h), switched off between
// reptaper and repcut.  The torques go to the frames through k_torque.
//
// Charge transfer (Rackers and Ponder, JCP 150, 084104 (2019); the reference
// assigns chgct and dmpct, kchgtrn.cpp, and has no energy, energy.cpp:276):
//   E_ik = -s f (c_i e^{-a_k r} + c_k e^{-a_i r}),
// f = electric/dielec, a = dmpct (0 read as 1000), s the pair's m scale,
// switched off between ctrntaper and ctrncut; a pair counts when either
// atom is active.
//
// Kernels follow the multipole tiles: every unscaled pair in k_*_tiles
// (per-tile partial sums reduced per atom in a fixed order), and the scaled
// pairs whole, in FP64, in k_*_excl.
#include "pair_mpole.h"
#include "pmq_internal.h"
#include "tile_pairs.h"

namespace pmq {

__global__ void k_hippo_sort(int n, int npad, const int *perm, const double *csix, const double *adisp,
                             const double *sizpr, const double *dmppr, const double *elepr, const double *chgct,
                             const double *dmpct, const int *use, double *shp)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    const int i = s < n ? perm[s] : -1;
    shp[s] = i >= 0 && csix ? csix[i] : 0.0;
    shp[npad + s] = i >= 0 && adisp ? adisp[i] : 1.0;
    shp[2 * npad + s] = i >= 0 && sizpr ? sizpr[i] : 0.0;
    shp[3 * npad + s] = i >= 0 && dmppr ? dmppr[i] : 1.0;
    shp[4 * npad + s] = i >= 0 && elepr ? elepr[i] : 0.0;
    shp[5 * npad + s] = i >= 0 && chgct ? chgct[i] : 0.0;
    const double a = i >= 0 && dmpct ? dmpct[i] : 0.0;
    shp[6 * npad + s] = a != 0.0 ? a : 1000.0;
    shp[7 * npad + s] = i >= 0 && use[i] ? 1.0 : 0.0;
}

int hippo_sort(pmq_ctx *c)
{
    const int np = c->npad;
    hipLaunchKernelGGL(k_hippo_sort, dim3((np + 255) / 256), dim3(256), 0, c->stream, c->n, np, c->perm.p,
                       c->have_disp ? c->csix.p : nullptr, c->have_disp ? c->adisp.p : nullptr,
                       c->have_rep ? c->sizpr.p : nullptr, c->have_rep ? c->dmppr.p : nullptr,
                       c->have_rep ? c->elepr.p : nullptr, c->have_ct ? c->chgct.p : nullptr,
                       c->have_ct ? c->dmpct.p : nullptr, c->use_raw.p, c->shp.p);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

struct HippoArgs {
    int n, npad;
    const int2 *tiles;
    const int *ntiles;
    TileItems items; // pieces of the tiles (tile_pairs.h)
    const double *sx, *sy, *sz, *smp, *shp;
    const int *perm, *iperm;
    DevBox box;
    double off2, dewald;   // term cutoff^2; dispersion Ewald coefficient (0 = none)
    double felec;          // electric / dielec (charge transfer)
    Taper taper;
    const unsigned long long *mask; // 128 per tile: [0,64) row masks of the in-cutoff pairs, scaled pairs left out
    void *part;            // per tile partial sums
    double *tscal;
    const ExclPair *excl;
    int nexcl;
    double *de, *te;       // caller order gradient (and torque, repulsion)
    long long *deq, *teq;  // mixed form: fixed-point accumulators
    Totals *tot;
    int vir;               // also the pair virial (tile scalars 1..9; totals for scaled pairs)
    size_t tstride;        // scalar stride of tscal
};

template <class T>
__device__ __forceinline__ T wsum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// dispersion pair: energy and dE/dr at scale s (EW: the Ewald real-space form)
template <class T, bool EW>
__device__ __forceinline__ void disp_pair(T r2, T ci, T ck, T ai, T ak, T s, T dewald, const Taper &tp, T &e, T &de)
{
    const T r = sqrt(r2);
    T dmp[6];
    overlap_damp<T>(r, ai, ak, dmp);
    const T f = T(1.5) * dmp[2] - T(0.5) * dmp[1];
    // the factors obey d(d_n r^-n)/dr = -n d_{n+2} r^-(n+1), so
    // d3' = 3 (d3 - d5) / r and d5' = 5 (d5 - d7) / r
    const T fp = (T(7.5) * (dmp[2] - dmp[3]) - T(1.5) * (dmp[1] - dmp[2])) / r;
    const T r6 = r2 * r2 * r2, cc = ci * ck;
    if (EW) {
        const T a2 = dewald * dewald, x2 = a2 * r2, ex = exp(-x2);
        const T g = (T(1) + x2 + T(0.5) * x2 * x2) * ex, gp = -a2 * r * x2 * x2 * ex;
        const T u = s * f * f - T(1) + g;
        e = -cc * u / r6;
        de = -cc * ((T(2) * s * f * fp + gp) / r6 - T(6) * u / (r6 * r));
    } else {
        T tv, tdv;
        taper_eval<T>(tp, r, tv, tdv);
        const T e0 = -cc * s * f * f / r6;
        const T de0 = -cc * s * (T(2) * f * fp / r6 - T(6) * f * f / (r6 * r));
        e = e0 * tv;
        de = de0 * tv + e0 * tdv;
    }
}

template <bool EW, class T>
__global__ __launch_bounds__(64) void k_disp_tiles(HippoArgs a)
{
    // side 0: block I, side 1: block J; compacted pair schedule (tile_pairs.h)
    __shared__ double bx[2][64], by[2][64], bz[2][64];
    __shared__ T bc[2][64], ba[2][64];
    __shared__ double acc[2][3][64];
    __shared__ unsigned short ring[TILE_RING];
    const int lane = threadIdx.x, np = a.npad;
    const int ntiles = *a.ntiles;
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
            bc[side][lane] = a.shp[s];
            ba[side][lane] = a.shp[np + s];
            acc[side][0][lane] = acc[side][1][lane] = acc[side][2][lane] = 0.0;
        }
        TilePairStream ps(a.mask[(size_t)t * 128 + lane]);
        __syncthreads();
        double e = 0.0;
        Vir6 vir;
        for (int cnt, ent, b = 0; b < it.b1 && (cnt = ps.next(ring, lane, ent)) > 0; b++) {
            if (b < it.b0) continue;
            if (lane >= cnt) break;
            const int il = ent >> 8, jl = ent & 63;
            double xd = bx[1][jl] - bx[0][il], yd = by[1][jl] - by[0][il], zd = bz[1][jl] - bz[0][il];
            image<double>(a.box, xd, yd, zd);
            const double r2d = xd * xd + yd * yd + zd * zd;
            const T ci = bc[0][il], ck = bc[1][jl];
            if (r2d > a.off2 || ci == T(0) || ck == T(0)) continue;
            const T xr = T(xd), yr = T(yd), zr = T(zd), r2 = T(r2d);
            T ep, de;
            disp_pair<T, EW>(r2, ci, ck, ba[0][il], ba[1][jl], T(1), T(a.dewald), a.taper, ep, de);
            e += (double)ep;
            const double g = (double)(de / sqrt(r2)); // dE/dx_k = g r_vec
            atomicAdd(&acc[0][0][il], -g * (double)xr);
            atomicAdd(&acc[0][1][il], -g * (double)yr);
            atomicAdd(&acc[0][2][il], -g * (double)zr);
            atomicAdd(&acc[1][0][jl], g * (double)xr);
            atomicAdd(&acc[1][1][jl], g * (double)yr);
            atomicAdd(&acc[1][2][jl], g * (double)zr);
            if (a.vir) vir.pair(xd, yd, zd, -g * (double)xr, -g * (double)yr, -g * (double)zr);
        }
        __syncthreads();
        T *P = (T *)a.part + (size_t)u * 2 * PMQ_TILE_NV * 64;
        T *Q = P + PMQ_TILE_NV * 64;
        for (int c = 0; c < 3; c++) {
            P[c * 64 + lane] = T(acc[0][c][lane]);
            Q[c * 64 + lane] = T(acc[1][c][lane]);
        }
        const double ew = wsum(e);
        if (a.vir) vir.wave_sum();
        if (lane == 0) {
            a.tscal[u] = ew; // scalar 0 of item u
            if (a.vir) vir.store_tile(a.tscal, a.tstride, u);
        }
    }
}

// scaled pairs, whole, at their scale (Ewald form, or switched by the taper)
__global__ __launch_bounds__(64) void k_disp_excl(HippoArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nexcl) return;
    const ExclPair q = a.excl[t];
    const int np = a.npad;
    if (q.dsp == 0.0 && !(a.dewald > 0.0)) return;
    const int si = a.iperm[q.i], sk = a.iperm[q.k];
    double xr = a.sx[sk] - a.sx[si], yr = a.sy[sk] - a.sy[si], zr = a.sz[sk] - a.sz[si];
    image<double>(a.box, xr, yr, zr);
    const double r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > a.off2) return;
    double e, de;
    if (a.dewald > 0.0)
        disp_pair<double, true>(r2, a.shp[si], a.shp[sk], a.shp[np + si], a.shp[np + sk], q.dsp, a.dewald, a.taper, e,
                                de);
    else
        disp_pair<double, false>(r2, a.shp[si], a.shp[sk], a.shp[np + si], a.shp[np + sk], q.dsp, 0.0, a.taper, e, de);
    atomicAdd(&a.tot->edsp, e);
    const double g = de / sqrt(r2);
    grad_add(a.de, a.deq, 3 * q.k + 0, g * xr); grad_add(a.de, a.deq, 3 * q.k + 1, g * yr);
    grad_add(a.de, a.deq, 3 * q.k + 2, g * zr);
    grad_add(a.de, a.deq, 3 * q.i + 0, -g * xr); grad_add(a.de, a.deq, 3 * q.i + 1, -g * yr);
    grad_add(a.de, a.deq, 3 * q.i + 2, -g * zr);
    if (a.vir) {
        Vir6 v;
        v.pair(xr, yr, zr, -g * xr, -g * yr, -g * zr);
        a.tot->add_vir(v);
    }
}

// repulsion pair at scale s: energy, dE/dx_i, torques
template <class T>
__device__ __forceinline__ T rep_pair(T r2, T xr, T yr, T zr, T si_, T sk_, T di, T dk, const Mp<T> &mi,
                                      const Mp<T> &mk, T s, const Taper &tp, T *F, T *Ti, T *Tk)
{
    const T r = sqrt(r2), rinv = T(1) / r;
    T B[6];
    rep_series<T>(r, di, dk, B);
    PairOut<T> o;
    pair_series<T, true>(B, mi, mk, xr, yr, zr, o);
    const T K = si_ * sk_ * s;
    T tv, tdv;
    taper_eval<T>(tp, r, tv, tdv);
    const T e0 = K * o.e * rinv;
    const T w = tv * K * rinv;
    const T c1 = tv * K * o.e * rinv * rinv * rinv - e0 * tdv * rinv;
    F[0] = w * o.fx + c1 * xr;
    F[1] = w * o.fy + c1 * yr;
    F[2] = w * o.fz + c1 * zr;
    Ti[0] = w * o.tix; Ti[1] = w * o.tiy; Ti[2] = w * o.tiz;
    Tk[0] = w * o.tkx; Tk[1] = w * o.tky; Tk[2] = w * o.tkz;
    return e0 * tv;
}

template <class T>
__device__ __forceinline__ Mp<T> rep_mp(const double *smp, const double *shp, int np, int s)
{
    Mp<T> m;
    m.c = shp[4 * np + s];
    m.dx = smp[np + s]; m.dy = smp[2 * np + s]; m.dz = smp[3 * np + s];
    m.qxx = smp[4 * np + s]; m.qxy = smp[5 * np + s]; m.qxz = smp[6 * np + s];
    m.qyy = smp[7 * np + s]; m.qyz = smp[8 * np + s]; m.qzz = smp[9 * np + s];
    return m;
}

template <class T>
__global__ __launch_bounds__(64) void k_rep_tiles(HippoArgs a)
{
    // side 0: block I, side 1: block J; compacted pair schedule (tile_pairs.h)
    __shared__ double bx[2][64], by[2][64], bz[2][64];
    __shared__ T bm[2][10][64], bsz[2][64], bdm[2][64];
    __shared__ double acc[2][6][64];
    __shared__ unsigned short ring[TILE_RING];
    const int lane = threadIdx.x, np = a.npad;
    const int ntiles = *a.ntiles;
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
            const Mp<T> m = rep_mp<T>(a.smp, a.shp, np, s);
            bm[side][0][lane] = m.c; bm[side][1][lane] = m.dx; bm[side][2][lane] = m.dy; bm[side][3][lane] = m.dz;
            bm[side][4][lane] = m.qxx; bm[side][5][lane] = m.qxy; bm[side][6][lane] = m.qxz;
            bm[side][7][lane] = m.qyy; bm[side][8][lane] = m.qyz; bm[side][9][lane] = m.qzz;
            bsz[side][lane] = a.shp[2 * np + s];
            bdm[side][lane] = a.shp[3 * np + s];
            for (int c = 0; c < 6; c++) acc[side][c][lane] = 0.0;
        }
        TilePairStream ps(a.mask[(size_t)t * 128 + lane]);
        __syncthreads();
        double e = 0.0;
        Vir6 vir;
        for (int cnt, ent, b = 0; b < it.b1 && (cnt = ps.next(ring, lane, ent)) > 0; b++) {
            if (b < it.b0) continue;
            if (lane >= cnt) break;
            const int il = ent >> 8, jl = ent & 63;
            double xd = bx[1][jl] - bx[0][il], yd = by[1][jl] - by[0][il], zd = bz[1][jl] - bz[0][il];
            image<double>(a.box, xd, yd, zd);
            const double r2d = xd * xd + yd * yd + zd * zd;
            if (r2d > a.off2 || bsz[0][il] == T(0) || bsz[1][jl] == T(0)) continue;
            const T xr = T(xd), yr = T(yd), zr = T(zd), r2 = T(r2d);
            Mp<T> mi, mk;
            mi.c = bm[0][0][il]; mi.dx = bm[0][1][il]; mi.dy = bm[0][2][il]; mi.dz = bm[0][3][il];
            mi.qxx = bm[0][4][il]; mi.qxy = bm[0][5][il]; mi.qxz = bm[0][6][il]; mi.qyy = bm[0][7][il];
            mi.qyz = bm[0][8][il]; mi.qzz = bm[0][9][il];
            mk.c = bm[1][0][jl]; mk.dx = bm[1][1][jl]; mk.dy = bm[1][2][jl]; mk.dz = bm[1][3][jl];
            mk.qxx = bm[1][4][jl]; mk.qxy = bm[1][5][jl]; mk.qxz = bm[1][6][jl]; mk.qyy = bm[1][7][jl];
            mk.qyz = bm[1][8][jl]; mk.qzz = bm[1][9][jl];
            T f[3], ti[3], tk[3];
            e += (double)rep_pair<T>(r2, xr, yr, zr, bsz[0][il], bsz[1][jl], bdm[0][il], bdm[1][jl], mi, mk, T(1),
                                     a.taper, f, ti, tk);
            for (int c = 0; c < 3; c++) {
                atomicAdd(&acc[0][c][il], (double)f[c]);
                atomicAdd(&acc[0][3 + c][il], (double)ti[c]);
                atomicAdd(&acc[1][c][jl], -(double)f[c]);
                atomicAdd(&acc[1][3 + c][jl], (double)tk[c]);
            }
            if (a.vir) vir.pair(xd, yd, zd, (double)f[0], (double)f[1], (double)f[2]);
        }
        __syncthreads();
        T *P = (T *)a.part + (size_t)u * 2 * PMQ_TILE_NV * 64;
        T *Q = P + PMQ_TILE_NV * 64;
        for (int c = 0; c < 6; c++) {
            P[c * 64 + lane] = T(acc[0][c][lane]);
            Q[c * 64 + lane] = T(acc[1][c][lane]);
        }
        const double ew = wsum(e);
        if (a.vir) vir.wave_sum();
        if (lane == 0) {
            a.tscal[u] = ew; // scalar 0 of item u
            if (a.vir) vir.store_tile(a.tscal, a.tstride, u);
        }
    }
}

__global__ __launch_bounds__(64) void k_rep_excl(HippoArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nexcl) return;
    const ExclPair q = a.excl[t];
    if (q.rep == 0.0) return;
    const int np = a.npad;
    const int si = a.iperm[q.i], sk = a.iperm[q.k];
    const double szi = a.shp[2 * np + si], szk = a.shp[2 * np + sk];
    if (szi == 0.0 || szk == 0.0) return;
    double xr = a.sx[sk] - a.sx[si], yr = a.sy[sk] - a.sy[si], zr = a.sz[sk] - a.sz[si];
    image<double>(a.box, xr, yr, zr);
    const double r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > a.off2) return;
    const Mp<double> mi = rep_mp<double>(a.smp, a.shp, np, si), mk = rep_mp<double>(a.smp, a.shp, np, sk);
    double F[3], Ti[3], Tk[3];
    const double e = rep_pair<double>(r2, xr, yr, zr, szi, szk, a.shp[3 * np + si], a.shp[3 * np + sk], mi, mk,
                                      q.rep, a.taper, F, Ti, Tk);
    atomicAdd(&a.tot->er, e);
    for (int c = 0; c < 3; c++) {
        grad_add(a.de, a.deq, 3 * q.i + c, F[c]);
        grad_add(a.de, a.deq, 3 * q.k + c, -F[c]);
        grad_add(a.te, a.teq, 3 * q.i + c, Ti[c]);
        grad_add(a.te, a.teq, 3 * q.k + c, Tk[c]);
    }
    if (a.vir) {
        Vir6 v;
        v.pair(xr, yr, zr, F[0], F[1], F[2]);
        a.tot->add_vir(v);
    }
}

// charge-transfer pair at scale-times-f sf: energy and dE/dr
template <class T>
__device__ __forceinline__ void ct_pair(T r, T ci, T ck, T ai, T ak, T sf, const Taper &tp, T &e, T &de)
{
    const T ei = exp(-ai * r), ek = exp(-ak * r);
    const T e0 = -sf * (ci * ek + ck * ei);
    const T de0 = sf * (ci * ak * ek + ck * ai * ei);
    T tv, tdv;
    taper_eval<T>(tp, r, tv, tdv);
    e = e0 * tv;
    de = de0 * tv + e0 * tdv;
}

template <class T>
__global__ __launch_bounds__(64) void k_ct_tiles(HippoArgs a)
{
    // side 0: block I, side 1: block J; compacted pair schedule (tile_pairs.h)
    __shared__ double bx[2][64], by[2][64], bz[2][64];
    __shared__ T bc[2][64], ba[2][64];
    __shared__ int bu[2][64];
    __shared__ double acc[2][3][64];
    __shared__ unsigned short ring[TILE_RING];
    const int lane = threadIdx.x, np = a.npad;
    const int ntiles = *a.ntiles;
    const T f = T(a.felec);
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
            bc[side][lane] = a.shp[5 * np + s];
            ba[side][lane] = a.shp[6 * np + s];
            bu[side][lane] = a.shp[7 * np + s] != 0.0;
            acc[side][0][lane] = acc[side][1][lane] = acc[side][2][lane] = 0.0;
        }
        TilePairStream ps(a.mask[(size_t)t * 128 + lane]);
        __syncthreads();
        double e = 0.0;
        Vir6 vir;
        for (int cnt, ent, b = 0; b < it.b1 && (cnt = ps.next(ring, lane, ent)) > 0; b++) {
            if (b < it.b0) continue;
            if (lane >= cnt) break;
            const int il = ent >> 8, jl = ent & 63;
            double xd = bx[1][jl] - bx[0][il], yd = by[1][jl] - by[0][il], zd = bz[1][jl] - bz[0][il];
            image<double>(a.box, xd, yd, zd);
            const double r2d = xd * xd + yd * yd + zd * zd;
            const T ci = bc[0][il], ck = bc[1][jl];
            if (r2d > a.off2 || !(bu[0][il] | bu[1][jl]) || (ci == T(0) && ck == T(0))) continue;
            const T r = T(sqrt(r2d));
            T ep, de;
            ct_pair<T>(r, ci, ck, ba[0][il], ba[1][jl], f, a.taper, ep, de);
            e += (double)ep;
            const double g = (double)(de / r); // dE/dx_k = g r_vec
            atomicAdd(&acc[0][0][il], -g * xd);
            atomicAdd(&acc[0][1][il], -g * yd);
            atomicAdd(&acc[0][2][il], -g * zd);
            atomicAdd(&acc[1][0][jl], g * xd);
            atomicAdd(&acc[1][1][jl], g * yd);
            atomicAdd(&acc[1][2][jl], g * zd);
            if (a.vir) vir.pair(xd, yd, zd, -g * xd, -g * yd, -g * zd);
        }
        __syncthreads();
        T *P = (T *)a.part + (size_t)u * 2 * PMQ_TILE_NV * 64;
        T *Q = P + PMQ_TILE_NV * 64;
        for (int c = 0; c < 3; c++) {
            P[c * 64 + lane] = T(acc[0][c][lane]);
            Q[c * 64 + lane] = T(acc[1][c][lane]);
        }
        const double ew = wsum(e);
        if (a.vir) vir.wave_sum();
        if (lane == 0) {
            a.tscal[u] = ew; // scalar 0 of item u
            if (a.vir) vir.store_tile(a.tscal, a.tstride, u);
        }
    }
}

// scaled pairs, whole, at their m scale
__global__ __launch_bounds__(64) void k_ct_excl(HippoArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nexcl) return;
    const ExclPair q = a.excl[t];
    if (q.m == 0.0) return;
    const int np = a.npad;
    const int si = a.iperm[q.i], sk = a.iperm[q.k];
    const double ci = a.shp[5 * np + si], ck = a.shp[5 * np + sk];
    if ((ci == 0.0 && ck == 0.0) || !(a.shp[7 * np + si] != 0.0 || a.shp[7 * np + sk] != 0.0)) return;
    double xr = a.sx[sk] - a.sx[si], yr = a.sy[sk] - a.sy[si], zr = a.sz[sk] - a.sz[si];
    image<double>(a.box, xr, yr, zr);
    const double r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > a.off2) return;
    const double r = sqrt(r2);
    double e, de;
    ct_pair<double>(r, ci, ck, a.shp[6 * np + si], a.shp[6 * np + sk], q.m * a.felec, a.taper, e, de);
    atomicAdd(&a.tot->ect, e);
    const double g = de / r;
    grad_add(a.de, a.deq, 3 * q.k + 0, g * xr); grad_add(a.de, a.deq, 3 * q.k + 1, g * yr);
    grad_add(a.de, a.deq, 3 * q.k + 2, g * zr);
    grad_add(a.de, a.deq, 3 * q.i + 0, -g * xr); grad_add(a.de, a.deq, 3 * q.i + 1, -g * yr);
    grad_add(a.de, a.deq, 3 * q.i + 2, -g * zr);
    if (a.vir) {
        Vir6 v;
        v.pair(xr, yr, zr, -g * xr, -g * yr, -g * zr);
        a.tot->add_vir(v);
    }
}

// ---------------------------------------------------------------------------
// host side

Taper make_taper(double cut, double off)
{
    Taper tp;
    tp.cut = 1e300;
    tp.w = 0.0;
    for (int q = 0; q < 6; q++) tp.c[q] = 0.0;
    if (!(cut < off)) return tp;
    const double den = pow(off - cut, 5), off2 = off * off, cut2 = cut * cut;
    tp.cut = cut;
    tp.w = 1.0 / (off - cut);
    tp.c[0] = off * off2 * (off2 - 5.0 * off * cut + 10.0 * cut2) / den;
    tp.c[1] = -30.0 * off2 * cut2 / den;
    tp.c[2] = 30.0 * (off2 * cut + off * cut2) / den;
    tp.c[3] = -10.0 * (off2 + 4.0 * off * cut + cut2) / den;
    tp.c[4] = 15.0 * (off + cut) / den;
    tp.c[5] = -6.0 / den;
    return tp;
}

static HippoArgs hippo_args(pmq_ctx *c, double off2)
{
    HippoArgs a;
    a.n = c->n; a.npad = c->npad;
    a.tiles = c->tiles.p; a.ntiles = c->ntiles_eff.p;
    a.sx = c->sx.p; a.sy = c->sy.p; a.sz = c->sz.p; a.smp = c->smp.p; a.shp = c->shp.p;
    a.perm = c->perm.p; a.iperm = c->iperm.p;
    a.box = c->box;
    a.off2 = off2;
    a.dewald = 0.0;
    a.felec = c->prm.electric / c->prm.dielec;
    a.taper = make_taper(1e300, 0.0);
    a.mask = c->xmask.p;
    a.part = c->tpart.p;
    a.tscal = c->tscal.p;
    a.excl = c->excl.p;
    a.nexcl = (int)c->h_excl.size();
    a.de = nullptr; a.te = nullptr; a.deq = nullptr; a.teq = nullptr;
    a.tot = c->totals.p;
    a.vir = 0;
    a.tstride = item_stride(c);
    a.items = tile_items(c);
    return a;
}

int launch_disp(pmq_ctx *c, double off2, bool grad, bool mixed, bool vir)
{
    HippoArgs a = hippo_args(c, off2);
    a.vir = vir ? 1 : 0;
    const bool ew = c->prm.use_dewald && c->box.kind != PMQ_BOX_NONE;
    a.dewald = ew ? c->prm.dewald : 0.0;
    if (!ew) a.taper = make_taper(c->prm.disptaper, sqrt(off2));
    a.de = c->dedsp.p;
    a.deq = mixed ? c->fx_grad.p : nullptr;
    {
        ProfScope ps_(c, "disp_tiles");
        const dim3 g(c->igrid);
        if (mixed) {
            if (ew) hipLaunchKernelGGL((k_disp_tiles<true, float>), g, dim3(64), 0, c->stream, a);
            else hipLaunchKernelGGL((k_disp_tiles<false, float>), g, dim3(64), 0, c->stream, a);
        } else {
            if (ew) hipLaunchKernelGGL((k_disp_tiles<true, double>), g, dim3(64), 0, c->stream, a);
            else hipLaunchKernelGGL((k_disp_tiles<false, double>), g, dim3(64), 0, c->stream, a);
        }
    }
    // the scaled pairs first: the reduction takes their fixed-point share
    // (mixed form) into its sums, and the tile scalars in the same launch
    if (a.nexcl) hipLaunchKernelGGL(k_disp_excl, dim3((a.nexcl + 63) / 64), dim3(64), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    int rc;
    if (grad) {
        TileReduceArgs r{};
        for (int q = 0; q < 3; q++) {
            r.out[q] = c->dedsp.p;
            r.caller_order[q] = 1;
            r.stride[q] = 3;
            r.off[q] = q;
            r.fxscale[q] = PMQ_FX_GRAD;
            r.fxq[q] = mixed ? c->fx_grad.p : nullptr;
        }
        return tile_reduce(c, r, 3, mixed, mixed, 3, 0, vir ? 10 : 1);
    }
    if ((rc = tile_scalars(c, 3, false))) return rc;
    if (vir && (rc = tile_virial(c))) return rc;
    return PMQ_OK;
}

int launch_rep(pmq_ctx *c, double off2, bool grad, bool mixed, bool vir)
{
    HippoArgs a = hippo_args(c, off2);
    a.vir = vir ? 1 : 0;
    a.taper = make_taper(c->prm.reptaper, sqrt(off2));
    a.de = c->der.p;
    a.te = c->ter.p;
    a.deq = mixed ? c->fx_grad.p : nullptr;
    a.teq = mixed ? c->fx_grad.p + 3 * (size_t)c->n : nullptr;
    {
        ProfScope ps_(c, "rep_tiles");
        const dim3 g(c->igrid);
        if (mixed) hipLaunchKernelGGL((k_rep_tiles<float>), g, dim3(64), 0, c->stream, a);
        else hipLaunchKernelGGL((k_rep_tiles<double>), g, dim3(64), 0, c->stream, a);
    }
    // the scaled pairs first: the reduction takes their fixed-point share
    // (mixed form) into its sums, and the tile scalars in the same launch
    if (a.nexcl) hipLaunchKernelGGL(k_rep_excl, dim3((a.nexcl + 63) / 64), dim3(64), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    int rc;
    if (grad) {
        TileReduceArgs r{};
        for (int q = 0; q < 3; q++) {
            r.out[q] = c->der.p; r.out[3 + q] = c->ter.p;
            r.caller_order[q] = r.caller_order[3 + q] = 1;
            r.stride[q] = r.stride[3 + q] = 3;
            r.off[q] = r.off[3 + q] = q;
            r.fxscale[q] = r.fxscale[3 + q] = PMQ_FX_GRAD;
            r.fxq[q] = mixed ? c->fx_grad.p : nullptr;
            r.fxq[3 + q] = mixed ? c->fx_grad.p + 3 * (size_t)c->n : nullptr;
        }
        return tile_reduce(c, r, 6, mixed, mixed, 2, 0, vir ? 10 : 1);
    }
    if ((rc = tile_scalars(c, 2, false))) return rc;
    if (vir && (rc = tile_virial(c))) return rc;
    return PMQ_OK;
}

} // namespace pmq

namespace pmq {

int launch_ct(pmq_ctx *c, double off2, bool grad, bool mixed, bool vir)
{
    HippoArgs a = hippo_args(c, off2);
    a.vir = vir ? 1 : 0;
    a.taper = make_taper(c->prm.ctrntaper, sqrt(off2));
    a.de = c->dect.p;
    a.deq = mixed ? c->fx_grad.p : nullptr;
    {
        ProfScope ps_(c, "ct_tiles");
        if (mixed) hipLaunchKernelGGL((k_ct_tiles<float>), dim3(c->igrid), dim3(64), 0, c->stream, a);
        else hipLaunchKernelGGL((k_ct_tiles<double>), dim3(c->igrid), dim3(64), 0, c->stream, a);
    }
    // the scaled pairs first: the reduction takes their fixed-point share
    // (mixed form) into its sums, and the tile scalars in the same launch
    if (a.nexcl) hipLaunchKernelGGL(k_ct_excl, dim3((a.nexcl + 63) / 64), dim3(64), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    int rc;
    if (grad) {
        TileReduceArgs r{};
        for (int q = 0; q < 3; q++) {
            r.out[q] = c->dect.p;
            r.caller_order[q] = 1;
            r.stride[q] = 3;
            r.off[q] = q;
            r.fxscale[q] = PMQ_FX_GRAD;
            r.fxq[q] = mixed ? c->fx_grad.p : nullptr;
        }
        return tile_reduce(c, r, 3, mixed, mixed, 5, 0, vir ? 10 : 1); // energy component 5: ect
    }
    if ((rc = tile_scalars(c, 5, false))) return rc; // energy component 5: ect
    if (vir && (rc = tile_virial(c))) return rc;
    return PMQ_OK;
}

} // namespace pmq
