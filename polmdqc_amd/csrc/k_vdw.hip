// k_vdw.hip -- AMOEBA buffered 14-7 van der Waals.
//
// The reference assigns the parameters (vdw.h, vdwpot.h, kvdw.cpp) and has
// no energy (energy.cpp:268); the form is Tinker's published ehal
// (Halgren, JACS 114, 7827 (1992)):
//   rho = r/R,  E = s eps (1.07/(rho + 0.07))^7 (1.12/(rho^7 + 0.12) - 2),
// r between reduced sites: an atom with a reduction partner sits at
// kred (x_i - x_ired) + x_ired, so a hydrogen's site moves toward its heavy
// atom, and the gradient at that site goes kred to the atom, 1 - kred to the
// partner.  R and eps come from class-pair tables (the 1-4 tables for pairs
// listed 1-4), s is the pair's v2..v5 scale, and the pair is switched off
// between vdwtaper and vdwcut by the quintic switch.
//
// The term has its own pair structure: the reduced sites are sorted and
// tiled at vdwcut (k_neighbors.hip, k_pairs.hip tile_structure), which is
// usually longer than the real-space electrostatic cutoff, and run before
// the electrostatic terms build theirs.  As for the other pair terms, the
// unscaled pairs run in 64 x 64 tiles with per-tile partial sums (FP64, or
// FP32 pairs summed in fixed point), the scaled pairs whole in FP64.
#include "pair_mpole.h"
#include "pmq_internal.h"
#include "tile_pairs.h"

namespace pmq {

struct VdwArgs {
    int n, npad, nv;
    const int2 *tiles;
    const int *ntiles;
    TileItems items; // pieces of the tiles (tile_pairs.h)
    const double *sx, *sy, *sz;          // sorted reduced sites
    const int *perm, *iperm;
    const int *vcls, *vuse;              // caller order: class, reduced site active
    const double *tab;                   // radmin, epsilon, radmin4, epsilon4 (nv x nv each)
    DevBox box;
    double off2;
    Taper taper;
    const unsigned long long *mask;      // 128 per tile: [0,64) in-cutoff rows, scaled pairs left out
    void *part;
    double *tscal;
    size_t tstride;
    const ExclPair *excl;
    int nexcl;
    double *de;                          // gradient at the reduced sites, caller order
    long long *deq;                      // mixed form: its fixed-point accumulator
    Totals *tot;
    int vir;
};

// energy and dE/dr of one pair at radius rv and depth eps (scale included)
template <class T>
__device__ __forceinline__ void hal_pair(T r, T rv, T eps, const Taper &tp, T &e, T &de)
{
    const T rho = r / rv, rho2 = rho * rho, rho6 = rho2 * rho2 * rho2, rho7 = rho6 * rho;
    const T s1 = T(1) / (rho + T(0.07)), s2 = T(1) / (rho7 + T(0.12));
    const T a = T(1.07) * s1, a2 = a * a, t1 = a2 * a2 * a2 * a; // (1.07 / (rho + 0.07))^7
    const T t2 = T(1.12) * s2;
    const T e0 = eps * t1 * (t2 - T(2));
    // dt1/drho = -7 t1 s1, dt2/drho = -7 rho^6 t2 s2
    const T de0 = -T(7) * eps * t1 * (s1 * (t2 - T(2)) + rho6 * t2 * s2) / rv;
    T tv, tdv;
    taper_eval<T>(tp, r, tv, tdv);
    e = e0 * tv;
    de = de0 * tv + e0 * tdv;
}

// the reduced sites, caller order
__global__ void k_vdw_reduce(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, const int *__restrict__ ired,
                             const double *__restrict__ kred, double *__restrict__ vx, double *__restrict__ vy,
                             double *__restrict__ vz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = ired[i];
    if (j == i) {
        vx[i] = x[i]; vy[i] = y[i]; vz[i] = z[i];
    } else {
        const double k = kred[i];
        vx[i] = k * (x[i] - x[j]) + x[j];
        vy[i] = k * (y[i] - y[j]) + y[j];
        vz[i] = k * (z[i] - z[j]) + z[j];
    }
}

template <class T>
__global__ __launch_bounds__(64) void k_vdw_tiles(VdwArgs a)
{
    // side 0: block I, side 1: block J; compacted pair schedule (tile_pairs.h)
    __shared__ double bx[2][64], by[2][64], bz[2][64];
    __shared__ int bc[2][64], bu[2][64];
    __shared__ double acc[2][3][64];
    __shared__ unsigned short ring[TILE_RING];
    const int lane = threadIdx.x, nv = a.nv, nv2 = nv * nv;
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
            const int i = s < a.n ? a.perm[s] : -1;
            bc[side][lane] = i >= 0 ? a.vcls[i] : 0;
            bu[side][lane] = i >= 0 ? a.vuse[i] : 0;
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
            if (r2d > a.off2 || !(bu[0][il] | bu[1][jl])) continue;
            const int q = bc[0][il] * nv + bc[1][jl];
            const double rv = a.tab[q], ep = a.tab[nv2 + q];
            if (ep == 0.0 || !(rv > 0.0)) continue;
            const T r = T(sqrt(r2d));
            T ev, de;
            hal_pair<T>(r, T(rv), T(ep), a.taper, ev, de);
            e += (double)ev;
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
        double ew = e;
        for (int o = 32; o > 0; o >>= 1) ew += __shfl_xor(ew, o, 64);
        if (a.vir) vir.wave_sum();
        if (lane == 0) {
            a.tscal[u] = ew; // scalar 0 of item u
            if (a.vir) vir.store_tile(a.tscal, a.tstride, u);
        }
    }
}

// scaled pairs, whole, in FP64 at their vdW scale (1 for pairs scaled only
// in other terms)
__global__ __launch_bounds__(64) void k_vdw_excl(VdwArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nexcl) return;
    const ExclPair q = a.excl[t];
    if (q.v == 0.0 || !(a.vuse[q.i] | a.vuse[q.k])) return;
    const int si = a.iperm[q.i], sk = a.iperm[q.k];
    double xr = a.sx[sk] - a.sx[si], yr = a.sy[sk] - a.sy[si], zr = a.sz[sk] - a.sz[si];
    image<double>(a.box, xr, yr, zr);
    const double r2 = xr * xr + yr * yr + zr * zr;
    if (r2 > a.off2) return;
    const int nv2 = a.nv * a.nv, c = a.vcls[q.i] * a.nv + a.vcls[q.k];
    const double rv = a.tab[(q.in14 ? 2 * nv2 : 0) + c];
    const double ep = a.tab[(q.in14 ? 3 * nv2 : nv2) + c] * q.v;
    if (ep == 0.0 || !(rv > 0.0)) return;
    const double r = sqrt(r2);
    double e, de;
    hal_pair<double>(r, rv, ep, a.taper, e, de);
    atomicAdd(&a.tot->ev, e);
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

// the reduced sites' gradient onto the atoms: atom i keeps kred_i of its own
// site's (all of it without a partner) and takes 1 - kred_j of each site j
// reduced toward it, in the fixed order of its list (no atomics)
__global__ void k_vdw_split(int n, const int *__restrict__ ired, const double *__restrict__ kred,
                            const int *__restrict__ koff, const int *__restrict__ kid, const double *__restrict__ g,
                            double *__restrict__ de)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double w = ired[i] != i ? kred[i] : 1.0;
    double s[3] = {w * g[3 * i], w * g[3 * i + 1], w * g[3 * i + 2]};
    for (int q = koff[i]; q < koff[i + 1]; q++) {
        const int j = kid[q];
        const double u = 1.0 - kred[j];
        s[0] += u * g[3 * j];
        s[1] += u * g[3 * j + 1];
        s[2] += u * g[3 * j + 2];
    }
    de[3 * i] = s[0];
    de[3 * i + 1] = s[1];
    de[3 * i + 2] = s[2];
}

int vdw_pass(pmq_ctx *c, bool grad, bool mixed, bool vir, bool safe)
{
    const int n = c->n;
    const int tb = 256;
    hipLaunchKernelGGL(k_vdw_reduce, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n, c->x.p, c->y.p, c->z.p,
                       c->ired.p, c->kred.p, c->vx.p, c->vy.p, c->vz.p);
    PMQ_HIP(hipGetLastError());
    // the structure over the reduced sites: plain cells, no preconditioner masks
    const bool pme_cells = c->pme_cells, prec_on = c->prec_on;
    c->pme_cells = false;
    c->prec_on = false;
    const double cut = c->prm.vdwcut, off2 = cut * cut;
    int rc;
    {
        ProfScope ps(c, "neighbors");
        rc = build_neighbors(c, cut, c->vx.p, c->vy.p, c->vz.p, c->vuse.p);
        if (!rc) rc = tile_structure(c, off2, safe);
    }
    c->pme_cells = pme_cells;
    c->prec_on = prec_on;
    if (rc) return rc;
    VdwArgs a;
    a.n = n; a.npad = c->npad; a.nv = c->nvdw;
    a.tiles = c->tiles.p; a.ntiles = c->ntiles_eff.p;
    a.sx = c->sx.p; a.sy = c->sy.p; a.sz = c->sz.p;
    a.perm = c->perm.p; a.iperm = c->iperm.p;
    a.vcls = c->vcls.p; a.vuse = c->vuse.p; a.tab = c->vtab.p;
    a.box = c->box;
    a.off2 = off2;
    a.taper = make_taper(c->prm.vdwtaper, cut);
    a.mask = c->xmask.p;
    a.part = c->tpart.p;
    a.tscal = c->tscal.p;
    a.tstride = item_stride(c);
    a.items = tile_items(c);
    a.excl = c->excl.p;
    a.nexcl = (int)c->h_excl.size();
    a.de = c->gvr.p;
    a.deq = mixed ? c->fx_grad.p : nullptr;
    a.tot = c->totals.p;
    a.vir = vir ? 1 : 0;
    {
        ProfScope ps_(c, "vdw_tiles");
        if (mixed) hipLaunchKernelGGL(k_vdw_tiles<float>, dim3(c->igrid), dim3(64), 0, c->stream, a);
        else hipLaunchKernelGGL(k_vdw_tiles<double>, dim3(c->igrid), dim3(64), 0, c->stream, a);
    }
    PMQ_HIP(hipGetLastError());
    // the scaled pairs first: the reduction takes their fixed-point share
    // (mixed form) into its sums, and the tile scalars in the same launch
    if (a.nexcl) hipLaunchKernelGGL(k_vdw_excl, dim3((a.nexcl + 63) / 64), dim3(64), 0, c->stream, a);
    PMQ_HIP(hipGetLastError());
    if (grad) {
        TileReduceArgs r{};
        for (int q = 0; q < 3; q++) {
            r.out[q] = c->gvr.p;
            r.caller_order[q] = 1;
            r.stride[q] = 3;
            r.off[q] = q;
            r.fxscale[q] = PMQ_FX_GRAD;
            r.fxq[q] = mixed ? c->fx_grad.p : nullptr;
        }
        if ((rc = tile_reduce(c, r, 3, mixed, mixed, 4, 0, vir ? 10 : 1))) return rc; // energy component 4: ev
    } else {
        if ((rc = tile_scalars(c, 4, false))) return rc; // energy component 4: ev
        if (vir && (rc = tile_virial(c))) return rc;
    }
    if (grad) {
        hipLaunchKernelGGL(k_vdw_split, dim3((n + tb - 1) / tb), dim3(tb), 0, c->stream, n, c->ired.p, c->kred.p,
                           c->vkid_off.p, c->vkid.p, c->gvr.p, c->dev.p);
        PMQ_HIP(hipGetLastError());
    }
    return PMQ_OK;
}

} // namespace pmq
