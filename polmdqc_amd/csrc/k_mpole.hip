// k_mpole.hip -- permanent-multipole kernels: local->global frames, the
// block-tile pair kernel, excluded-pair corrections, torque->force and charge
// flux.  Plain FP64 form; the reference's arithmetic is FP64 throughout
// (precision.h:16).
//
//   k_rotpole      chkpole.cpp:44-89 + rotmat/rotsite (rotpole.cpp:102-425),
//                  charge-flux monopole shift (alterchg.cpp:50-53), one
//                  thread per atom, writes the sorted SoA multipoles
//   k_mpole_tiles  the i<k loop of empole_a (empole.cpp:198-345) over 64x64
//                  block tiles, every pair but the scaled ones at scale 1
//   k_mpole_excl   the 1-2..1-5 (and polarization-group) pairs whole, at
//                  their scale (empole.cpp:224-235: later classes override),
//                  in FP64
//   k_torque       torque.cpp:42-397
//   k_cf_delta / k_cf_grad   alterchg.cpp:36-169 / dcflux.cpp:31-225
#include "pair_mpole.h"
#include "pmq_internal.h"
#include "tile_pairs.h"

namespace pmq {

__global__ void k_zero(double *__restrict__ p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// d += q / scale, q = 0 (fixed-point accumulators back to FP64)
__global__ void k_fx_flush(size_t n, long long *__restrict__ q, double *__restrict__ d, double inv_scale)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        d[i] += (double)q[i] * inv_scale;
        q[i] = 0;
    }
}

int launch_fx_flush(pmq_ctx *c, long long *q, double *d, size_t n, double scale)
{
    if (!n) return PMQ_OK;
    const int tb = 256;
    const int gb = (int)std::min<size_t>((n + tb - 1) / tb, 4096);
    hipLaunchKernelGGL(k_fx_flush, dim3(gb), dim3(tb), 0, c->stream, n, q, d, 1.0 / scale);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// two accumulators held back to back (q[0, n) and q[n, 2n), e.g. forces
// then torques) flushed into two separate FP64 arrays by one launch
__global__ void k_fx_flush_pair(size_t n, long long *__restrict__ q, double *__restrict__ d0,
                                double *__restrict__ d1, double inv_scale)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < 2 * n; i += (size_t)gridDim.x * blockDim.x) {
        double *d = i < n ? d0 + i : d1 + (i - n);
        *d += (double)q[i] * inv_scale;
        q[i] = 0;
    }
}

int launch_fx_flush_pair(pmq_ctx *c, long long *q, double *d0, double *d1, size_t n, double scale)
{
    if (!n) return PMQ_OK;
    const int tb = 256;
    const int gb = (int)std::min<size_t>((2 * n + tb - 1) / tb, 4096);
    hipLaunchKernelGGL(k_fx_flush_pair, dim3(gb), dim3(tb), 0, c->stream, n, q, d0, d1, 1.0 / scale);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// several buffers cleared by one launch (segment k: blockIdx.y == k)
struct ZeroList {
    double *p[PMQ_ZERO_MAX];
    size_t n[PMQ_ZERO_MAX];
};

__global__ void k_zero_many(ZeroList z)
{
    double *d = z.p[blockIdx.y];
    const size_t n = z.n[blockIdx.y];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = 0.0;
}

// the evaluation's status record (run counts and the PCG scalars) written
// straight into the pinned host mirror, one launch instead of two copies
__global__ void k_status_out(const RunStatus *__restrict__ st, const double *__restrict__ sc, HostStatus *h)
{
    const int t = threadIdx.x;
    if (t == 0) h->st = *st;
    if (t < 11) h->sc[t] = sc ? sc[t] : 0.0;
}

int launch_status_out(pmq_ctx *c)
{
    if (!c->h_status_dev) {
        PMQ_HIP(hipHostGetDevicePointer((void **)&c->h_status_dev, c->h_status, 0));
    }
    hipLaunchKernelGGL(k_status_out, dim3(1), dim3(64), 0, c->stream, c->status.p, c->pcg.p, c->h_status_dev);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int launch_zero_many(pmq_ctx *c, int k, double *const *p, const size_t *n)
{
    if (k <= 0) return PMQ_OK;
    if (k > PMQ_ZERO_MAX) {
        c->err = "launch_zero_many: too many buffers";
        return PMQ_ERR_ARG;
    }
    ZeroList z{};
    size_t mx = 0;
    for (int i = 0; i < k; i++) {
        z.p[i] = p[i];
        z.n[i] = n[i];
        mx = std::max(mx, n[i]);
    }
    const int tb = 256;
    const int gb = (int)std::max<size_t>(1, std::min<size_t>((mx + tb - 1) / tb, 1024));
    hipLaunchKernelGGL(k_zero_many, dim3(gb, k), dim3(tb), 0, c->stream, z);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int launch_zero(pmq_ctx *c, double *p, size_t n)
{
    if (!n) return PMQ_OK;
    const int tb = 256;
    const int gb = (int)std::min<size_t>((n + tb - 1) / tb, 4096);
    hipLaunchKernelGGL(k_zero, dim3(gb), dim3(tb), 0, c->stream, p, n);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// ---------------------------------------------------------------------------
// local -> global frame
struct RotArgs {
    int n, npad;
    const int *perm;
    const double *x, *y, *z;
    const double *pole;               // 13n local
    const int *polaxe, *zaxis, *xaxis, *yaxis;
    int cp, cf;
    const double *pcore, *pval0, *palpha, *pdelta;
    double *smp;                      // 10 * npad sorted
    double *scp;                      // 3 * npad sorted: core, alpha, charge (monopole)
    double *rpole13;                  // 13n caller order or null
    double *rotm;                     // 9n caller order: rows x, y, z of the local frame (or null)
    int *flip;                        // n: 1 where chkpole inverted the y-dependent components (or null)
    // with the rotation, per sorted site (null: not wanted): the sorted
    // polarization parameters (polarity, thole, pdamp; 3 * npad) and the
    // permanent field zeroed (6 * npad)
    const double *polarity, *thole, *pdamp;
    double *spol, *pfield;
    int nrot;                         // workgroups of the rotation; those past it clear the scaled pairs
    XclearArgs xc;                    // from the tile masks, one thread each
};

__device__ __forceinline__ void dvec(const RotArgs &a, int from, int to, double *d)
{
    d[0] = a.x[to] - a.x[from];
    d[1] = a.y[to] - a.y[from];
    d[2] = a.z[to] - a.z[from];
}

__device__ __forceinline__ double dot3(const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }

__global__ void k_rotpole(RotArgs a)
{
    if ((int)blockIdx.x >= a.nrot) {
        xmask_clear_one(a.xc, ((int)blockIdx.x - a.nrot) * (int)blockDim.x + (int)threadIdx.x);
        return;
    }
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.npad) return;
    const int i = s < a.n ? a.perm[s] : -1;
    if (a.spol) {
        a.spol[s] = i >= 0 ? a.polarity[i] : 0.0;
        a.spol[a.npad + s] = i >= 0 ? a.thole[i] : 0.0;
        a.spol[2 * a.npad + s] = i >= 0 ? a.pdamp[i] : 0.0;
    }
    if (a.pfield)
        for (int c = 0; c < 6; c++) a.pfield[(size_t)c * a.npad + s] = 0.0;
    if (i < 0) {
        for (int c = 0; c < 10; c++) a.smp[c * a.npad + s] = 0.0;
        for (int c = 0; c < 3; c++) a.scp[c * a.npad + s] = 0.0;
        return;
    }
    double p[13];
    for (int c = 0; c < 13; c++) p[c] = a.pole[13 * i + c];
    double val = a.cp ? a.pval0[i] : 0.0;
    if (a.cf) {
        // alterchg.cpp:50-53: pole[0] = mono0 + delta, pval = pval0 + delta
        // (the caller passes mono0 in pole[0])
        const double dq = (double)((const long long *)a.pdelta)[i] * (1.0 / PMQ_FX_CF);
        p[0] += dq;
        val += dq;
    }
    const int frame = a.polaxe[i];
    const int iz = a.zaxis[i] - 1, ix = a.xaxis[i] - 1, ky = a.yaxis[i], iy = abs(ky) - 1;
    // chirality check at Z-then-X sites with a y-axis atom (chkpole.cpp:60-87)
    if (frame == PMQ_FRAME_ZTHENX && ky != 0) {
        double ad[3], bd[3], cd[3];
        dvec(a, iy, i, ad);
        dvec(a, iy, iz, bd);
        dvec(a, iy, ix, cd);
        const double c1 = bd[1] * cd[2] - bd[2] * cd[1];
        const double c2 = cd[1] * ad[2] - cd[2] * ad[1];
        const double c3 = ad[1] * bd[2] - ad[2] * bd[1];
        const double vol = ad[0] * c1 + bd[0] * c2 + cd[0] * c3;
        if ((ky < 0 && vol > 0.0) || (ky > 0 && vol < 0.0)) {
            p[2] = -p[2]; p[5] = -p[5]; p[7] = -p[7]; p[9] = -p[9]; p[11] = -p[11];
            if (a.flip) a.flip[i] = 1;
        } else if (a.flip) {
            a.flip[i] = 0;
        }
    } else if (a.flip) {
        a.flip[i] = 0;
    }
    // rotation matrix rows: ax (x axis), ay, az
    double ax[3] = {1, 0, 0}, ay[3], az[3] = {0, 0, 1}, d[3], r, dot;
    bool planar = false;
    if (frame == PMQ_FRAME_ZONLY) {
        dvec(a, i, iz, d);
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) az[j] = d[j] / r;
        double v[3] = {1, 0, 0};
        dot = az[0];
        if (fabs(dot) > 0.707) { v[0] = 0; v[1] = 1; dot = az[1]; }
        for (int j = 0; j < 3; j++) v[j] -= dot * az[j];
        r = sqrt(dot3(v, v));
        for (int j = 0; j < 3; j++) ax[j] = v[j] / r;
    } else if (frame == PMQ_FRAME_ZTHENX) {
        dvec(a, i, iz, d);
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) az[j] = d[j] / r;
        dvec(a, i, ix, d);
        dot = dot3(d, az);
        for (int j = 0; j < 3; j++) d[j] -= dot * az[j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) ax[j] = d[j] / r;
    } else if (frame == PMQ_FRAME_BISECTOR) {
        double u1[3], u2[3];
        dvec(a, i, iz, u1);
        dvec(a, i, ix, u2);
        const double r1 = sqrt(dot3(u1, u1)), r2 = sqrt(dot3(u2, u2));
        for (int j = 0; j < 3; j++) { u1[j] /= r1; u2[j] /= r2; d[j] = u1[j] + u2[j]; }
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) az[j] = d[j] / r;
        dot = dot3(u2, az);
        for (int j = 0; j < 3; j++) d[j] = u2[j] - dot * az[j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) ax[j] = d[j] / r;
    } else if (frame == PMQ_FRAME_ZBISECT) {
        double u1[3], u2[3];
        dvec(a, i, iz, d);
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) az[j] = d[j] / r;
        dvec(a, i, ix, u1);
        dvec(a, i, iy, u2);
        const double r1 = sqrt(dot3(u1, u1)), r2 = sqrt(dot3(u2, u2));
        for (int j = 0; j < 3; j++) { u1[j] /= r1; u2[j] /= r2; d[j] = u1[j] + u2[j]; }
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) d[j] /= r;
        dot = dot3(d, az);
        if (180.0 - 57.29577951308232088 * acos(dot) < 0.0) {
            planar = true;
            d[0] = u1[1] * u2[2] - u1[2] * u2[1];
            d[1] = u1[2] * u2[0] - u1[0] * u2[2];
            d[2] = u1[0] * u2[1] - u1[1] * u2[0];
            dot = dot3(d, az);
            if (dot < 0.0) { for (int j = 0; j < 3; j++) d[j] = -d[j]; dot = -dot; }
        }
        for (int j = 0; j < 3; j++) d[j] -= dot * az[j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) ax[j] = d[j] / r;
    } else if (frame == PMQ_FRAME_THREEFOLD) {
        double u1[3], u2[3], u3[3];
        dvec(a, i, iz, u1);
        dvec(a, i, ix, u2);
        dvec(a, i, iy, u3);
        const double r1 = sqrt(dot3(u1, u1)), r2 = sqrt(dot3(u2, u2)), r3 = sqrt(dot3(u3, u3));
        for (int j = 0; j < 3; j++) { u1[j] /= r1; u2[j] /= r2; u3[j] /= r3; d[j] = u1[j] + u2[j] + u3[j]; }
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) az[j] = d[j] / r;
        dot = dot3(u1, az);
        for (int j = 0; j < 3; j++) d[j] = u1[j] - dot * az[j];
        r = sqrt(dot3(d, d));
        for (int j = 0; j < 3; j++) ax[j] = d[j] / r;
    }
    ay[0] = ax[2] * az[1] - ax[1] * az[2];
    ay[1] = ax[0] * az[2] - ax[2] * az[0];
    ay[2] = ax[1] * az[0] - ax[0] * az[1];
    if (planar && frame == PMQ_FRAME_ZBISECT) {
        p[1] = 0; p[6] = 0; p[10] = 0;
        p[4] = 0.5 * (p[4] + p[8]);
        p[8] = p[4];
    }
    const double *A[3] = {ax, ay, az};
    double g[13];
    g[0] = p[0];
    for (int h = 0; h < 3; h++) g[1 + h] = p[1] * ax[h] + p[2] * ay[h] + p[3] * az[h];
    for (int h = 0; h < 3; h++)
        for (int k = h; k < 3; k++) {
            double v = 0.0;
            for (int l = 0; l < 3; l++) {
                const double al = A[l][h];
                double t = 0.0;
                for (int m = 0; m < 3; m++) t += A[m][k] * p[4 + 3 * l + m];
                v += al * t;
            }
            g[4 + 3 * h + k] = v;
            g[4 + 3 * k + h] = v;
        }
    const int np = a.npad;
    a.smp[0 * np + s] = a.cp ? val : g[0];
    a.smp[1 * np + s] = g[1];
    a.smp[2 * np + s] = g[2];
    a.smp[3 * np + s] = g[3];
    a.smp[4 * np + s] = g[4];
    a.smp[5 * np + s] = g[5];
    a.smp[6 * np + s] = g[6];
    a.smp[7 * np + s] = g[8];
    a.smp[8 * np + s] = g[9];
    a.smp[9 * np + s] = g[12];
    a.scp[0 * np + s] = a.cp ? a.pcore[i] : 0.0;
    a.scp[1 * np + s] = a.cp ? a.palpha[i] : 0.0;
    a.scp[2 * np + s] = g[0];
    if (a.rpole13)
        for (int c = 0; c < 13; c++) a.rpole13[13 * i + c] = g[c];
    if (a.rotm)
        for (int j = 0; j < 3; j++) {
            a.rotm[9 * (size_t)i + j] = ax[j];
            a.rotm[9 * (size_t)i + 3 + j] = ay[j];
            a.rotm[9 * (size_t)i + 6 + j] = az[j];
        }
}

int launch_rotpole(pmq_ctx *c, bool want_rpole13, bool polar_sites, bool clear_scaled)
{
    RotArgs a;
    a.n = c->n;
    a.npad = c->npad;
    a.perm = c->perm.p;
    a.x = c->x.p; a.y = c->y.p; a.z = c->z.p;
    a.pole = c->pole.p;
    a.polaxe = c->polaxe.p; a.zaxis = c->zaxis.p; a.xaxis = c->xaxis.p; a.yaxis = c->yaxis.p;
    a.cp = c->prm.pentype != PMQ_PEN_NONE;
    a.cf = c->prm.use_chgflx;
    a.pcore = c->pcore.p; a.pval0 = c->pval0.p; a.palpha = c->palpha.p; a.pdelta = c->pdelta.p;
    a.smp = c->smp.p;
    a.scp = c->scp.p;
    a.rpole13 = want_rpole13 ? c->rpole13.p : nullptr;
    a.rotm = c->rotm.p;
    a.flip = c->flip.p;
    a.polarity = c->polarity.p; a.thole = c->thole.p; a.pdamp = c->pdamp.p;
    a.spol = polar_sites ? c->spol.p : nullptr;
    a.pfield = polar_sites ? c->pfield.p : nullptr;
    a.xc = xclear_args(c);
    const int tb = 128;
    a.nrot = (c->npad + tb - 1) / tb;
    const int nx = clear_scaled ? (a.xc.nexcl + tb - 1) / tb : 0;
    hipLaunchKernelGGL(k_rotpole, dim3(a.nrot + nx), dim3(tb), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// ---------------------------------------------------------------------------
// wave-level sums
template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct MpoleArgs {
    int n, npad;
    const int2 *tiles;
    const int *ntiles;
    TileItems items; // pieces of the tiles (tile_pairs.h)
    const double *sx, *sy, *sz, *smp, *scp;
    const int *sflag, *smol, *perm, *iperm;
    DevBox box;
    double off2, f, aewald;
    double *dem, *te, *pot, *aem;
    long long *demq, *teq;   // mixed precision: 64-bit fixed-point gradient and torque (else null)
    long long *potq;         // mixed precision: fixed-point dE/dq of the scaled pairs (else null)
    const unsigned long long *xmask; // per tile: [64,128) row masks of the in-cutoff pairs with an active atom,
                                     // scaled pairs left out
    void *part;              // per tile partial sums (2 sides x PMQ_TILE_NV x 64 of R)
    double *tscal;           // per tile scalars: scalar q of tile t at tscal[q * tstride + t]
    size_t tstride;
    Totals *tot;
    // exclusion corrections
    const ExclPair *excl;
    int nexcl;
};

// R is the arithmetic of the pair terms: double (plain form) or float (mixed
// form, whose displacements are still taken in FP64 before rounding)
// QI: each pair in its quasi-internal frame (pair_mpole.h qi_frame)
template <class R, int PEN, bool CF, bool ANA, bool GRAD, bool VIR, bool EW, bool QI = false>
__global__ __launch_bounds__(64) void k_mpole_tiles(MpoleArgs a)
{
    constexpr bool CP = PEN != 0;
    // side 0: the atoms of block I, side 1: those of block J
    __shared__ double bx[2][PMQ_BLOCK], by[2][PMQ_BLOCK], bz[2][PMQ_BLOCK];
    __shared__ R bm[2][10][PMQ_BLOCK];
    __shared__ R bc[2][3][PMQ_BLOCK]; // core, valence exponent, total charge
    __shared__ int bmol[2][PMQ_BLOCK];
    // FP64 accumulators whatever R is (ds_add_f64 measured faster than
    // ds_add_f32 here)
    constexpr int NACC = ANA ? 8 : CF ? 7 : 6;
    __shared__ double acc[2][NACC][PMQ_BLOCK]; // f(3), t(3), pot, aem
    __shared__ unsigned short ring[TILE_RING];
    const int lane = threadIdx.x;
    const int ntiles = *a.ntiles;
    const int np = a.npad;

    for (int v = blockIdx.x, nitems = ntiles + *a.items.xcnt; v < nitems; v += gridDim.x) {
        TileItem it;
        if (!tile_item(v, ntiles, a.items, it)) continue;
        const int u = it.u;
        const int t = it.t;
        const int2 IJ = a.tiles[t];
        __syncthreads();
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int s = (side ? IJ.y : IJ.x) * PMQ_BLOCK + lane;
            bx[side][lane] = a.sx[s]; by[side][lane] = a.sy[s]; bz[side][lane] = a.sz[s];
            for (int c = 0; c < 10; c++) bm[side][c][lane] = a.smp[c * np + s];
            if (CP) {
                bc[side][0][lane] = a.scp[s]; bc[side][1][lane] = a.scp[np + s]; bc[side][2][lane] = a.scp[2 * np + s];
            }
            if (ANA) bmol[side][lane] = a.smol[s];
            for (int c = 0; c < NACC; c++) acc[side][c][lane] = 0.0;
        }
        TilePairStream ps(a.xmask[(size_t)t * 128 + 64 + lane]);
        __syncthreads();

        double e = 0; // energy sums stay FP64 whatever R is
        R ein = 0;
        Vir6 vir; // FP64 whatever R is
        unsigned nem = 0;
        for (int cnt, ent, b = 0; b < it.b1 && (cnt = ps.next(ring, lane, ent)) > 0; b++) {
            if (b < it.b0) continue;
            if (lane >= cnt) break;
            const int il = ent >> 8, jl = ent & 63;
            double xd = bx[1][jl] - bx[0][il], yd = by[1][jl] - by[0][il], zd = bz[1][jl] - bz[0][il];
            image<double>(a.box, xd, yd, zd);
            const R xr = R(xd), yr = R(yd), zr = R(zd);
            const R r2 = xr * xr + yr * yr + zr * zr;
            Mp<R> mi, mk;
            mi.c = bm[0][0][il]; mi.dx = bm[0][1][il]; mi.dy = bm[0][2][il]; mi.dz = bm[0][3][il];
            mi.qxx = bm[0][4][il]; mi.qxy = bm[0][5][il]; mi.qxz = bm[0][6][il];
            mi.qyy = bm[0][7][il]; mi.qyz = bm[0][8][il]; mi.qzz = bm[0][9][il];
            mk.c = bm[1][0][jl]; mk.dx = bm[1][1][jl]; mk.dy = bm[1][2][jl]; mk.dz = bm[1][3][jl];
            mk.qxx = bm[1][4][jl]; mk.qxy = bm[1][5][jl]; mk.qxz = bm[1][6][jl];
            mk.qyy = bm[1][7][jl]; mk.qyz = bm[1][8][jl]; mk.qzz = bm[1][9][jl];
            PairOut<R> o;
            R bn[6];
            if (EW) ewald_series<R>(r2, R(a.aewald), bn);
            if (QI) {
                // the pair algebra at r = (0, 0, |r|) on the rotated multipoles
                const QiFrame<R> fr = qi_frame<R>(xr, yr, zr, r2);
                const Mp<R> qa = qi_rotate<R>(fr, mi), qb = qi_rotate<R>(fr, mk);
                if (CP) {
                    pair_cp<R, GRAD, (PEN == 0 ? 1 : PEN), EW>(r2, R(0), R(0), fr.len, R(a.f), bc[0][0][il],
                                                               bc[0][2][il], bc[0][1][il], qa, bc[1][0][jl],
                                                               bc[1][2][jl], bc[1][1][jl], qb, o, bn);
                } else {
                    R Rs[6];
                    if (EW) {
                        for (int q = 0; q < 6; q++) Rs[q] = R(a.f) * bn[q];
                    } else {
                        rr_series<R>(r2, R(a.f), Rs);
                    }
                    pair_series<R, GRAD>(Rs, qa, qb, R(0), R(0), fr.len, o);
                }
                if (GRAD) qi_back<R>(fr, o);
            } else if (CP) {
                pair_cp<R, GRAD, (PEN == 0 ? 1 : PEN), EW>(r2, xr, yr, zr, R(a.f), bc[0][0][il], bc[0][2][il],
                                                           bc[0][1][il], mi, bc[1][0][jl], bc[1][2][jl],
                                                           bc[1][1][jl], mk, o, bn);
            } else {
                R Rs[6];
                if (EW) {
                    for (int q = 0; q < 6; q++) Rs[q] = R(a.f) * bn[q];
                } else {
                    rr_series<R>(r2, R(a.f), Rs);
                }
                pair_series<R, GRAD>(Rs, mi, mk, xr, yr, zr, o);
            }
            e += o.e;
            if (ANA) {
                if (o.e != R(0)) nem++;
                atomicAdd(&acc[0][7][il], (double)(R(0.5) * o.e));
                atomicAdd(&acc[1][7][jl], (double)(R(0.5) * o.e));
                if (!EW) {
                    if (bmol[0][il] != bmol[1][jl]) ein += o.e;
                } else if (bmol[0][il] == bmol[1][jl]) {
                    // with Ewald the intermolecular energy is em minus the
                    // molecules' own unscreened pair energies: those go to ein
                    PairOut<R> u;
                    if (CP) {
                        pair_cp<R, false, (PEN == 0 ? 1 : PEN), false>(r2, xr, yr, zr, R(a.f), bc[0][0][il],
                                                                       bc[0][2][il], bc[0][1][il], mi, bc[1][0][jl],
                                                                       bc[1][2][jl], bc[1][1][jl], mk, u);
                    } else {
                        R Rs[6];
                        rr_series<R>(r2, R(a.f), Rs);
                        pair_series<R, false>(Rs, mi, mk, xr, yr, zr, u);
                    }
                    ein += u.e;
                }
            }
            if (GRAD) {
                atomicAdd(&acc[0][0][il], (double)o.fx);
                atomicAdd(&acc[0][1][il], (double)o.fy);
                atomicAdd(&acc[0][2][il], (double)o.fz);
                atomicAdd(&acc[0][3][il], (double)o.tix);
                atomicAdd(&acc[0][4][il], (double)o.tiy);
                atomicAdd(&acc[0][5][il], (double)o.tiz);
                atomicAdd(&acc[1][0][jl], -(double)o.fx);
                atomicAdd(&acc[1][1][jl], -(double)o.fy);
                atomicAdd(&acc[1][2][jl], -(double)o.fz);
                atomicAdd(&acc[1][3][jl], (double)o.tkx);
                atomicAdd(&acc[1][4][jl], (double)o.tky);
                atomicAdd(&acc[1][5][jl], (double)o.tkz);
                if (CF) { atomicAdd(&acc[0][6][il], (double)o.phii); atomicAdd(&acc[1][6][jl], (double)o.phik); }
                if (VIR) vir.pair(xd, yd, zd, (double)o.fx, (double)o.fy, (double)o.fz);
            }
        }
        __syncthreads();
        // partial sums of this tile: side 0 = the i atoms, side 1 = the j atoms
        if (GRAD || ANA) {
            R *P = (R *)a.part + (size_t)u * 2 * PMQ_TILE_NV * 64;
            R *Q = P + PMQ_TILE_NV * 64;
            if (GRAD)
                for (int c = 0; c < 6 + (CF ? 1 : 0); c++) {
                    P[c * 64 + lane] = R(acc[0][c][lane]);
                    Q[c * 64 + lane] = R(acc[1][c][lane]);
                }
            if (ANA) {
                P[7 * 64 + lane] = R(acc[0][7][lane]);
                Q[7 * 64 + lane] = R(acc[1][7][lane]);
            }
        }
        // totals in FP64 whatever R is
        const double ew = wave_sum(e);
        double win = 0;
        if (VIR) vir.wave_sum();
        if (ANA) { win = wave_sum((double)ein); nem = wave_sum(nem); }
        if (lane == 0) {
            double *S = a.tscal + u;
            const size_t ts = a.tstride;
            S[0] = ew;
            if (VIR || ANA) { // the full set is summed when either is on
                vir.store_tile(a.tscal, ts, u);
                S[10 * ts] = win; S[11 * ts] = (double)nem;
            }
        }
    }
}

// scaled pairs, whole (the tiles leave them out), counted in nem when their
// energy is not zero.  The Ewald real-space pair at scale m is
// f (B_n - (1 - m dmp) rr_n) = f (B_n - rr_n) [total multipoles] + m x the
// plain (or charge-penetration damped) pair
template <int PEN, bool CF, bool ANA, bool GRAD, bool VIR>
__global__ __launch_bounds__(64) void k_mpole_excl(MpoleArgs a)
{
    using R = double;
    constexpr bool CP = PEN != 0;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nexcl) return;
    const ExclPair q = a.excl[t];
    const R sc = R(q.m);
    const int si = a.iperm[q.i], sk = a.iperm[q.k];
    const int fi = a.sflag[si], fk = a.sflag[sk];
    if (!((fi | fk) & 2)) return;
    if (sc == R(0) && !(a.aewald > 0.0)) return;
    R xr = a.sx[sk] - a.sx[si], yr = a.sy[sk] - a.sy[si], zr = a.sz[sk] - a.sz[si];
    image<R>(a.box, xr, yr, zr);
    const R r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > a.off2) return;
    const int np = a.npad;
    Mp<R> mi, mk;
    const double *m = a.smp;
    mi.c = m[si]; mi.dx = m[np + si]; mi.dy = m[2 * np + si]; mi.dz = m[3 * np + si];
    mi.qxx = m[4 * np + si]; mi.qxy = m[5 * np + si]; mi.qxz = m[6 * np + si];
    mi.qyy = m[7 * np + si]; mi.qyz = m[8 * np + si]; mi.qzz = m[9 * np + si];
    mk.c = m[sk]; mk.dx = m[np + sk]; mk.dy = m[2 * np + sk]; mk.dz = m[3 * np + sk];
    mk.qxx = m[4 * np + sk]; mk.qxy = m[5 * np + sk]; mk.qxz = m[6 * np + sk];
    mk.qyy = m[7 * np + sk]; mk.qyz = m[8 * np + sk]; mk.qzz = m[9 * np + sk];
    const R sf = sc * R(a.f);
    PairOut<R> o;
    if (CP) {
        pair_cp<R, GRAD, (PEN == 0 ? 1 : PEN)>(r2, xr, yr, zr, sf, a.scp[si], a.scp[2 * np + si], a.scp[np + si], mi,
                                               a.scp[sk], a.scp[2 * np + sk], a.scp[np + sk], mk, o);
    } else {
        R Rs[6];
        rr_series<R>(r2, sf, Rs);
        pair_series<R, GRAD>(Rs, mi, mk, xr, yr, zr, o);
    }
    const R e_scaled = o.e; // the pair's own scaled energy (without Ewald)
    if (a.aewald > 0.0) {
        // f (B_n - rr_n) between the total multipoles
        R bn[6], rr[6], Rs[6];
        ewald_series<R>(r2, R(a.aewald), bn);
        rr_series<R>(r2, R(1), rr);
        for (int j = 0; j < 6; j++) Rs[j] = R(a.f) * (bn[j] - rr[j]);
        Mp<R> ti = mi, tk = mk;
        if (CP) { ti.c = a.scp[2 * np + si]; tk.c = a.scp[2 * np + sk]; }
        pair_series<R, GRAD, true>(Rs, ti, tk, xr, yr, zr, o);
    }
    atomicAdd(&a.tot->em, o.e);
    if (ANA) {
        if (o.e != R(0)) atomicAdd(&a.tot->nem, 1ull);
        atomicAdd(&a.aem[q.i], R(0.5) * o.e);
        atomicAdd(&a.aem[q.k], R(0.5) * o.e);
        // without Ewald the intermolecular pair energies, with it the
        // intramolecular pairs' own (unscreened) energies: einter = em - those
        if (!(a.aewald > 0.0)) {
            if (a.smol[si] != a.smol[sk]) atomicAdd(&a.tot->einter, o.e);
        } else if (a.smol[si] == a.smol[sk]) {
            atomicAdd(&a.tot->einter, e_scaled);
        }
    }
    if (GRAD) {
        grad_add(a.dem, a.demq, 3 * q.i + 0, o.fx); grad_add(a.dem, a.demq, 3 * q.i + 1, o.fy);
        grad_add(a.dem, a.demq, 3 * q.i + 2, o.fz);
        grad_add(a.dem, a.demq, 3 * q.k + 0, -o.fx); grad_add(a.dem, a.demq, 3 * q.k + 1, -o.fy);
        grad_add(a.dem, a.demq, 3 * q.k + 2, -o.fz);
        grad_add(a.te, a.teq, 3 * q.i + 0, o.tix); grad_add(a.te, a.teq, 3 * q.i + 1, o.tiy);
        grad_add(a.te, a.teq, 3 * q.i + 2, o.tiz);
        grad_add(a.te, a.teq, 3 * q.k + 0, o.tkx); grad_add(a.te, a.teq, 3 * q.k + 1, o.tky);
        grad_add(a.te, a.teq, 3 * q.k + 2, o.tkz);
        if (CF) {
            if (a.potq) {
                fx_add(a.potq, q.i, (double)o.phii, PMQ_FX_GRAD);
                fx_add(a.potq, q.k, (double)o.phik, PMQ_FX_GRAD);
            } else {
                atomicAdd(&a.pot[q.i], (double)o.phii);
                atomicAdd(&a.pot[q.k], (double)o.phik);
            }
        }
        if (VIR) {
            Vir6 v;
            v.pair(xr, yr, zr, o.fx, o.fy, o.fz);
            a.tot->add_vir(v);
        }
    }
}

template <int PEN, bool CF, bool ANA, bool GRAD, bool VIR>
static int launch_mpole_t(pmq_ctx *c, const MpoleArgs &a)
{
    const int grid = c->igrid;
    {
        ProfScope ps_(c, "mpole_tiles");
        const bool ew = a.aewald > 0.0;
        const bool qi = c->pair_form == PMQ_PAIR_QI;
#define PMQ_MT(R, CF_, ANA_, EW_)                                                                              \
        do {                                                                                                   \
            if (qi) hipLaunchKernelGGL((k_mpole_tiles<R, PEN, CF_, ANA_, GRAD, VIR, EW_, true>), dim3(grid), dim3(PMQ_BLOCK), 0, c->stream, a); \
            else hipLaunchKernelGGL((k_mpole_tiles<R, PEN, CF_, ANA_, GRAD, VIR, EW_, false>), dim3(grid), dim3(PMQ_BLOCK), 0, c->stream, a); \
        } while (0)
        if (a.demq) { // mixed precision: FP32 pair terms, fixed-point sums
            if (ew) PMQ_MT(float, CF, ANA, true);
            else PMQ_MT(float, CF, ANA, false);
        } else {
            if (ew) PMQ_MT(double, CF, ANA, true);
            else PMQ_MT(double, CF, ANA, false);
        }
#undef PMQ_MT
    }
    PMQ_HIP(hipGetLastError());
    const bool mixed = a.demq != nullptr;
    // the scaled pairs first: the reduction takes their fixed-point share of
    // the gradient and torque (mixed form) into its sums, and the tile
    // scalars in the same launch
    if (a.nexcl > 0)
        hipLaunchKernelGGL((k_mpole_excl<PEN, CF, ANA, GRAD, VIR>), dim3((a.nexcl + 63) / 64), dim3(64), 0,
                           c->stream, a);
    PMQ_HIP(hipGetLastError());
    if (GRAD || ANA) {
        TileReduceArgs r{};
        for (int q = 0; q < 3; q++) {
            if (GRAD) {
                r.out[q] = a.dem; r.out[3 + q] = a.te;
                r.fxq[q] = a.demq; r.fxq[3 + q] = a.teq;
            }
            r.caller_order[q] = r.caller_order[3 + q] = 1;
            r.stride[q] = r.stride[3 + q] = 3;
            r.off[q] = r.off[3 + q] = q;
            r.fxscale[q] = r.fxscale[3 + q] = PMQ_FX_GRAD;
        }
        if (GRAD && CF) r.out[6] = a.pot;
        if (ANA) r.out[7] = a.aem;
        for (int q = 6; q < 8; q++) { r.caller_order[q] = 1; r.stride[q] = 1; r.off[q] = 0; r.fxscale[q] = PMQ_FX_GRAD; }
        return tile_reduce(c, r, 8, mixed, mixed, 0, 0, (VIR || ANA) ? PMQ_TILE_NS : 1);
    }
    return tile_scalars(c, 0, VIR || ANA);
}

template <int PEN, bool CF>
static int launch_mpole_mode(pmq_ctx *c, const MpoleArgs &a, int mode)
{
    switch (mode) {
    case PMQ_ENERGY: return launch_mpole_t<PEN, CF, false, false, false>(c, a);
    case PMQ_ANALYSIS: return launch_mpole_t<PEN, CF, true, false, false>(c, a);
    case PMQ_GRADIENT: return launch_mpole_t<PEN, CF, false, true, false>(c, a);
    default: return launch_mpole_t<PEN, CF, false, true, true>(c, a);
    }
}

int launch_mpole_direct(pmq_ctx *c, int mode, double off2, bool mixed)
{
    MpoleArgs a;
    a.n = c->n;
    a.npad = c->npad;
    a.tiles = c->tiles.p;
    a.ntiles = c->ntiles_eff.p;
    a.sx = c->sx.p; a.sy = c->sy.p; a.sz = c->sz.p;
    a.smp = c->smp.p; a.scp = c->scp.p;
    a.sflag = c->sflag.p; a.smol = c->smol.p; a.perm = c->perm.p; a.iperm = c->iperm.p;
    a.box = c->box;
    a.off2 = off2;
    a.f = c->prm.electric / c->prm.dielec;
    a.dem = c->dem.p; a.te = c->te.p; a.pot = c->pot.p; a.aem = c->aem.p;
    a.demq = mixed ? c->fx_grad.p : nullptr;
    a.teq = mixed ? c->fx_grad.p + 3 * (size_t)c->n : nullptr;
    a.potq = mixed ? c->fx_pot.p : nullptr;
    a.xmask = c->xmask.p;
    a.part = c->tpart.p;
    a.tscal = c->tscal.p;
    a.tstride = item_stride(c);
    a.items = tile_items(c);
    a.tot = c->totals.p;
    a.excl = c->excl.p;
    a.nexcl = (int)c->h_excl.size();
    a.aewald = c->prm.use_ewald ? c->prm.aewald : 0.0;
    const int pen = c->prm.pentype;
    const bool cf = pen != PMQ_PEN_NONE && c->prm.use_chgflx;
    if (pen == PMQ_PEN_NONE) return launch_mpole_mode<0, false>(c, a, mode);
    if (pen == PMQ_PEN_GORDON1) return cf ? launch_mpole_mode<1, true>(c, a, mode) : launch_mpole_mode<1, false>(c, a, mode);
    return cf ? launch_mpole_mode<2, true>(c, a, mode) : launch_mpole_mode<2, false>(c, a, mode);
}

// ---------------------------------------------------------------------------
// torque -> forces on the frame atoms (torque.cpp:42-397)
struct TorqueArgs {
    int n;
    const double *x, *y, *z, *te;
    const int *polaxe, *zaxis, *xaxis, *yaxis;
    double *de;
    long long *deq; // mixed precision: fixed-point accumulation of de (then flushed), else null
    Totals *tot;
    int vir;
};

__device__ __forceinline__ void cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void normalize(double *v)
{
    const double l = sqrt(dot3(v, v));
    v[0] /= l; v[1] /= l; v[2] /= l;
}

// blockIdx.y selects one of two torque sets (the same kernel for both)
__global__ __launch_bounds__(128) void k_torque(TorqueArgs a0, TorqueArgs a1)
{
    const TorqueArgs a = blockIdx.y == 0 ? a0 : a1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double W[6] = {0, 0, 0, 0, 0, 0};
    if (i < a.n && a.polaxe[i] != PMQ_FRAME_NONE) {
        const int frame = a.polaxe[i];
        const double t[3] = {a.te[3 * i], a.te[3 * i + 1], a.te[3 * i + 2]};
        const int ia = a.zaxis[i] - 1, ic = a.xaxis[i] - 1, id = abs(a.yaxis[i]) - 1;
        double u[3] = {a.x[ia] - a.x[i], a.y[ia] - a.y[i], a.z[ia] - a.z[i]}, v[3], w[3];
        const double usiz = sqrt(dot3(u, u));
        double vsiz;
        if (frame != PMQ_FRAME_ZONLY) {
            v[0] = a.x[ic] - a.x[i]; v[1] = a.y[ic] - a.y[i]; v[2] = a.z[ic] - a.z[i];
            vsiz = sqrt(dot3(v, v));
        } else {
            v[0] = 1; v[1] = 0; v[2] = 0;
            vsiz = 1;
            if (fabs(u[0] / usiz) > 0.866) { v[0] = 0; v[1] = 1; }
        }
        if (frame == PMQ_FRAME_ZBISECT || frame == PMQ_FRAME_THREEFOLD) {
            w[0] = a.x[id] - a.x[i]; w[1] = a.y[id] - a.y[i]; w[2] = a.z[id] - a.z[i];
        } else {
            cross(u, v, w);
        }
        const double wsiz = sqrt(dot3(w, w));
        for (int j = 0; j < 3; j++) { u[j] /= usiz; v[j] /= vsiz; w[j] /= wsiz; }
        double fz[3] = {0, 0, 0}, fx[3] = {0, 0, 0}, fy[3] = {0, 0, 0};
        const double dphidu = -dot3(t, u), dphidv = -dot3(t, v), dphidw = -dot3(t, w);
        double uv[3], uw[3], vw[3];
        cross(v, u, uv); normalize(uv);
        cross(w, u, uw); normalize(uw);
        cross(w, v, vw); normalize(vw);
        const double uvcos = dot3(u, v), uvsin = sqrt(1.0 - uvcos * uvcos);
        if (frame == PMQ_FRAME_ZONLY) {
            for (int j = 0; j < 3; j++) fz[j] = uv[j] * dphidv / (usiz * uvsin) + uw[j] * dphidw / usiz;
        } else if (frame == PMQ_FRAME_ZTHENX) {
            for (int j = 0; j < 3; j++) {
                fz[j] = uv[j] * dphidv / (usiz * uvsin) + uw[j] * dphidw / usiz;
                fx[j] = -uv[j] * dphidu / (vsiz * uvsin);
            }
        } else if (frame == PMQ_FRAME_BISECTOR) {
            for (int j = 0; j < 3; j++) {
                fz[j] = uv[j] * dphidv / (usiz * uvsin) + 0.5 * uw[j] * dphidw / usiz;
                fx[j] = -uv[j] * dphidu / (vsiz * uvsin) + 0.5 * vw[j] * dphidw / vsiz;
            }
        } else if (frame == PMQ_FRAME_ZBISECT) {
            double r[3], s[3];
            for (int j = 0; j < 3; j++) r[j] = v[j] + w[j];
            normalize(r);
            cross(u, r, s);
            normalize(s);
            const double dphidr = -dot3(t, r), dphids = -dot3(t, s);
            double ur[3], us[3];
            cross(r, u, ur); normalize(ur);
            cross(s, u, us); normalize(us);
            const double urcos = dot3(u, r), ursin = sqrt(1.0 - urcos * urcos);
            const double vscos = dot3(v, s), vssin = sqrt(1.0 - vscos * vscos);
            const double wscos = dot3(w, s), wssin = sqrt(1.0 - wscos * wscos);
            double t1[3], t2[3];
            for (int j = 0; j < 3; j++) { t1[j] = v[j] - s[j] * vscos; t2[j] = w[j] - s[j] * wscos; }
            normalize(t1);
            normalize(t2);
            const double c1 = dot3(u, t1), c2 = dot3(u, t2);
            const double ut1sin = sqrt(1.0 - c1 * c1), ut2sin = sqrt(1.0 - c2 * c2);
            for (int j = 0; j < 3; j++) {
                fz[j] = ur[j] * dphidr / (usiz * ursin) + us[j] * dphids / usiz;
                fx[j] = (vssin * s[j] - vscos * t1[j]) * dphidu / (vsiz * (ut1sin + ut2sin));
                fy[j] = (wssin * s[j] - wscos * t2[j]) * dphidu / (wsiz * (ut1sin + ut2sin));
            }
        } else if (frame == PMQ_FRAME_THREEFOLD) {
            double p[3];
            for (int j = 0; j < 3; j++) p[j] = u[j] + v[j] + w[j];
            const double psiz = sqrt(dot3(p, p));
            for (int j = 0; j < 3; j++) p[j] /= psiz;
            const double cs[3] = {dot3(u, p), dot3(v, p), dot3(w, p)};
            // axis pairs: (u+v) about w -> fy, (v+w) about u -> fz, (u+w) about v -> fx
            const double *ax[3] = {w, u, v};
            const double *pa[3] = {u, v, u};
            const double *pb[3] = {v, w, w};
            const double sz[3] = {wsiz, usiz, vsiz};
            const double pc[3] = {cs[2], cs[0], cs[1]};
            double *dst[3] = {fy, fz, fx};
            for (int q = 0; q < 3; q++) {
                double r[3], del[3], eps[3];
                for (int j = 0; j < 3; j++) r[j] = pa[q][j] + pb[q][j];
                normalize(r);
                const double c = dot3(r, ax[q]), sn = sqrt(1.0 - c * c);
                const double dr = -dot3(t, r);
                cross(r, ax[q], del);
                normalize(del);
                const double dd = -dot3(t, del);
                cross(del, ax[q], eps);
                for (int j = 0; j < 3; j++)
                    dst[q][j] = del[j] * dr / (sz[q] * sn) + eps[j] * dd * pc[q] / (sz[q] * psiz);
            }
        }
        for (int j = 0; j < 3; j++) {
            grad_add(a.de, a.deq, 3 * ia + j, fz[j]);
            if (frame != PMQ_FRAME_ZONLY) grad_add(a.de, a.deq, 3 * ic + j, fx[j]);
            if (frame == PMQ_FRAME_ZBISECT || frame == PMQ_FRAME_THREEFOLD) grad_add(a.de, a.deq, 3 * id + j, fy[j]);
            grad_add(a.de, a.deq, 3 * i + j, -(fz[j] + fx[j] + fy[j]));
        }
        if (a.vir) {
            const int jz = ia >= 0 ? ia : i, jx = ic >= 0 ? ic : i, jy = id >= 0 ? id : i;
            const double rz[3] = {a.x[jz] - a.x[i], a.y[jz] - a.y[i], a.z[jz] - a.z[i]};
            const double rx[3] = {a.x[jx] - a.x[i], a.y[jx] - a.y[i], a.z[jx] - a.z[i]};
            const double ry[3] = {a.x[jy] - a.x[i], a.y[jy] - a.y[i], a.z[jy] - a.z[i]};
            double M[3][3];
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) M[p][q] = rx[p] * fx[q] + ry[p] * fy[q] + rz[p] * fz[q];
            W[0] = M[0][0];
            W[1] = 0.5 * (M[1][0] + M[0][1]);
            W[2] = 0.5 * (M[2][0] + M[0][2]);
            W[3] = M[1][1];
            W[4] = 0.5 * (M[2][1] + M[1][2]);
            W[5] = M[2][2];
        }
    }
    if (a.vir) {
        __shared__ double red[6][128];
        for (int k = 0; k < 6; k++) red[k][threadIdx.x] = W[k];
        __syncthreads();
        for (int w = 64; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
                for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            atomicAdd(&a.tot->vir[0], red[0][0]);
            atomicAdd(&a.tot->vir[1], red[1][0]); atomicAdd(&a.tot->vir[3], red[1][0]);
            atomicAdd(&a.tot->vir[2], red[2][0]); atomicAdd(&a.tot->vir[6], red[2][0]);
            atomicAdd(&a.tot->vir[4], red[3][0]);
            atomicAdd(&a.tot->vir[5], red[4][0]); atomicAdd(&a.tot->vir[7], red[4][0]);
            atomicAdd(&a.tot->vir[8], red[5][0]);
        }
    }
}

static TorqueArgs torque_args(pmq_ctx *c, int mode, const double *te, double *de, long long *deq)
{
    TorqueArgs a;
    a.n = c->n;
    a.x = c->x.p; a.y = c->y.p; a.z = c->z.p; a.te = te;
    a.polaxe = c->polaxe.p; a.zaxis = c->zaxis.p; a.xaxis = c->xaxis.p; a.yaxis = c->yaxis.p;
    a.de = de;
    a.deq = deq;
    a.tot = c->totals.p;
    a.vir = mode == PMQ_VIRIAL;
    return a;
}

int launch_torque(pmq_ctx *c, int mode, const double *te, double *de, bool mixed)
{
    const TorqueArgs a = torque_args(c, mode, te, de, mixed ? c->fx_grad.p : nullptr);
    hipLaunchKernelGGL(k_torque, dim3((c->n + 127) / 128), dim3(128), 0, c->stream, a, a);
    PMQ_HIP(hipGetLastError());
    if (mixed) return launch_fx_flush(c, c->fx_grad.p, de, 3 * (size_t)c->n, PMQ_FX_GRAD);
    return PMQ_OK;
}

// two torque sets in one launch (mixed: two fixed-point halves of fx_grad,
// flushed together)
int launch_torque_pair(pmq_ctx *c, int mode0, const double *te0, double *de0, int mode1, const double *te1,
                       double *de1, bool mixed)
{
    const size_t n3 = 3 * (size_t)c->n;
    const TorqueArgs a0 = torque_args(c, mode0, te0, de0, mixed ? c->fx_grad.p : nullptr);
    const TorqueArgs a1 = torque_args(c, mode1, te1, de1, mixed ? c->fx_grad.p + n3 : nullptr);
    hipLaunchKernelGGL(k_torque, dim3((c->n + 127) / 128, 2), dim3(128), 0, c->stream, a0, a1);
    PMQ_HIP(hipGetLastError());
    if (mixed) return launch_fx_flush_pair(c, c->fx_grad.p, de0, de1, n3, PMQ_FX_GRAD);
    return PMQ_OK;
}

// ---------------------------------------------------------------------------
// charge flux
struct CfArgs {
    int n, nbond, nangle;
    const double *x, *y, *z;
    const int *ibnd, *iang, *balist;
    const double *bl, *bflx, *anat, *aflx, *abflx;
    double *pdelta;              // 64-bit fixed point (PMQ_FX_CF): the sum does not depend on the order
    const double *pot[2];        // dE/dq of the terms (multipoles, polarization; null: skipped)
    double *de[2];               // their gradients (caller order)
    long long *deq[2];           // mixed form: fixed-point gradients instead (else null)
    Totals *tot;
    int vir;
};

__device__ __forceinline__ void cf_shift(const CfArgs &a, int i, double v)
{
    fx_add((long long *)a.pdelta, i, v, PMQ_FX_CF);
}

__global__ void k_cf_delta(CfArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const double radian = 57.29577951308232088;
    if (t < a.nbond) {
        const int ia = a.ibnd[2 * t], ib = a.ibnd[2 * t + 1];
        const double d[3] = {a.x[ia] - a.x[ib], a.y[ia] - a.y[ib], a.z[ia] - a.z[ib]};
        const double dq = a.bflx[t] * (sqrt(dot3(d, d)) - a.bl[t]);
        cf_shift(a, ia, -dq);
        cf_shift(a, ib, dq);
    } else if (t < a.nbond + a.nangle) {
        const int g = t - a.nbond;
        const int ia = a.iang[3 * g], ib = a.iang[3 * g + 1], ic = a.iang[3 * g + 2];
        const double ab[3] = {a.x[ia] - a.x[ib], a.y[ia] - a.y[ib], a.z[ia] - a.z[ib]};
        const double cb[3] = {a.x[ic] - a.x[ib], a.y[ic] - a.y[ib], a.z[ic] - a.z[ib]};
        const double rab = sqrt(dot3(ab, ab)), rcb = sqrt(dot3(cb, cb));
        if (rab > 0.0 && rcb > 0.0) {
            double cs = dot3(ab, cb) / (rab * rcb);
            cs = fmin(1.0, fmax(-1.0, cs));
            const double dth = (radian * acos(cs) - a.anat[g]) / radian;
            const double dq1 = a.abflx[2 * g] * (rcb - a.bl[a.balist[2 * g + 1]]) + a.aflx[2 * g] * dth;
            const double dq2 = a.abflx[2 * g + 1] * (rab - a.bl[a.balist[2 * g]]) + a.aflx[2 * g + 1] * dth;
            cf_shift(a, ia, dq1);
            cf_shift(a, ib, -(dq1 + dq2));
            cf_shift(a, ic, dq2);
        }
    }
}

__device__ __forceinline__ void cf_add(const CfArgs &a, int set, int i, const double *f, double *W)
{
    grad_add(a.de[set], a.deq[set], 3 * i, f[0]);
    grad_add(a.de[set], a.deq[set], 3 * i + 1, f[1]);
    grad_add(a.de[set], a.deq[set], 3 * i + 2, f[2]);
    // dcflux.cpp:204-223 builds the virial from absolute coordinates
    const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
    W[0] += xi * f[0]; W[1] += yi * f[0]; W[2] += zi * f[0];
    W[3] += yi * f[1]; W[4] += zi * f[1]; W[5] += zi * f[2];
}

__global__ __launch_bounds__(128) void k_cf_grad(CfArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double W[6] = {0, 0, 0, 0, 0, 0};
    for (int set = 0; set < 2; set++) {
        const double *pot = a.pot[set];
        if (!pot) continue;
        if (t < a.nbond) {
            const int ia = a.ibnd[2 * t], ib = a.ibnd[2 * t + 1];
            const double d[3] = {a.x[ia] - a.x[ib], a.y[ia] - a.y[ib], a.z[ia] - a.z[ib]};
            const double pb = a.bflx[t] * (pot[ib] - pot[ia]) / sqrt(dot3(d, d));
            const double fa[3] = {pb * d[0], pb * d[1], pb * d[2]};
            const double fb[3] = {-fa[0], -fa[1], -fa[2]};
            cf_add(a, set, ia, fa, W);
            cf_add(a, set, ib, fb, W);
        } else if (t < a.nbond + a.nangle) {
            const int g = t - a.nbond;
            const int ia = a.iang[3 * g], ib = a.iang[3 * g + 1], ic = a.iang[3 * g + 2];
            const double ab[3] = {a.x[ia] - a.x[ib], a.y[ia] - a.y[ib], a.z[ia] - a.z[ib]};
            const double cb[3] = {a.x[ic] - a.x[ib], a.y[ic] - a.y[ib], a.z[ic] - a.z[ib]};
            const double rab2 = dot3(ab, ab), rcb2 = dot3(cb, cb);
            const double rab = sqrt(rab2), rcb = sqrt(rcb2);
            if (rab > 0.0 && rcb > 0.0) {
                const double dpa = pot[ia] - pot[ib], dpc = pot[ic] - pot[ib];
                const double pb1 = dpa * a.abflx[2 * g], pb2 = dpc * a.abflx[2 * g + 1];
                const double rabc = rab * rcb, dp = dot3(ab, cb);
                const double dra3 = dp / (rab2 * rab * rcb), drc3 = dp / (rab * rcb2 * rcb);
                const double fterm =
                    -rabc / sqrt(rab2 * rcb2 - dp * dp) * (dpa * a.aflx[2 * g] + dpc * a.aflx[2 * g + 1]);
                double fa[3], fc[3], fb[3];
                for (int j = 0; j < 3; j++) {
                    const double fa1 = pb2 * ab[j] / rab, fc1 = pb1 * cb[j] / rcb;
                    const double fa2 = fterm * (cb[j] / rabc - ab[j] * dra3);
                    const double fc2 = fterm * (ab[j] / rabc - cb[j] * drc3);
                    fa[j] = fa1 + fa2;
                    fc[j] = fc1 + fc2;
                    fb[j] = -fa1 - fc1 - fa2 - fc2;
                }
                cf_add(a, set, ia, fa, W);
                cf_add(a, set, ib, fb, W);
                cf_add(a, set, ic, fc, W);
            }
        }
    }
    if (a.vir) {
        __shared__ double red[6][128];
        for (int k = 0; k < 6; k++) red[k][threadIdx.x] = W[k];
        __syncthreads();
        for (int w = 64; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w)
                for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            atomicAdd(&a.tot->vir[0], red[0][0]);
            atomicAdd(&a.tot->vir[1], red[1][0]); atomicAdd(&a.tot->vir[3], red[1][0]);
            atomicAdd(&a.tot->vir[2], red[2][0]); atomicAdd(&a.tot->vir[6], red[2][0]);
            atomicAdd(&a.tot->vir[4], red[3][0]);
            atomicAdd(&a.tot->vir[5], red[4][0]); atomicAdd(&a.tot->vir[7], red[4][0]);
            atomicAdd(&a.tot->vir[8], red[5][0]);
        }
    }
}

static CfArgs cf_args(pmq_ctx *c, int mode)
{
    CfArgs a;
    a.n = c->n; a.nbond = c->nbond; a.nangle = c->nangle;
    a.x = c->x.p; a.y = c->y.p; a.z = c->z.p;
    a.ibnd = c->ibnd.p; a.iang = c->iang.p; a.balist = c->balist.p;
    a.bl = c->bl.p; a.bflx = c->bflx.p; a.anat = c->anat.p; a.aflx = c->aflx.p; a.abflx = c->abflx.p;
    a.pdelta = c->pdelta.p;
    a.pot[0] = a.pot[1] = nullptr;
    a.de[0] = c->dem.p;
    a.de[1] = c->dep.p;
    a.deq[0] = a.deq[1] = nullptr;
    a.tot = c->totals.p;
    a.vir = mode == PMQ_VIRIAL;
    return a;
}

int launch_chgflx_delta(pmq_ctx *c)
{
    int rc = launch_zero(c, c->pdelta.p, c->n);
    if (rc) return rc;
    const int nt = c->nbond + c->nangle;
    if (nt > 0) hipLaunchKernelGGL(k_cf_delta, dim3((nt + 127) / 128), dim3(128), 0, c->stream, cf_args(c, 0));
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int launch_chgflx_grad(pmq_ctx *c, int mode, bool mpole, bool polar, bool mixed)
{
    const int nt = c->nbond + c->nangle;
    const size_t n3 = 3 * (size_t)c->n;
    int rc;
    // the scaled pairs' fixed-point dE/dq first (mixed form)
    if (mixed && mpole && (rc = launch_fx_flush(c, c->fx_pot.p, c->pot.p, c->n, PMQ_FX_GRAD))) return rc;
    if (mixed && polar && (rc = launch_fx_flush(c, c->fx_pot.p + c->n, c->potp.p, c->n, PMQ_FX_GRAD))) return rc;
    CfArgs a = cf_args(c, mode);
    if (mpole) a.pot[0] = c->pot.p;
    if (polar) a.pot[1] = c->potp.p;
    if (mixed) {
        a.deq[0] = c->fx_grad.p;
        a.deq[1] = c->fx_grad.p + n3;
    }
    if (nt > 0 && (mpole || polar))
        hipLaunchKernelGGL(k_cf_grad, dim3((nt + 127) / 128), dim3(128), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    if (mixed && mpole && (rc = launch_fx_flush(c, c->fx_grad.p, c->dem.p, n3, PMQ_FX_GRAD))) return rc;
    if (mixed && polar && (rc = launch_fx_flush(c, c->fx_grad.p + n3, c->dep.p, n3, PMQ_FX_GRAD))) return rc;
    return PMQ_OK;
}

} // namespace pmq
