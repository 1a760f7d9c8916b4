// pme_spread.h -- the PME grid geometry and the owner-computes B-spline
// spreading of one brick of the grid, shared by the standalone spreading
// launch and the one fused with the pair-matrix product (k_pme.hip).
#pragma once

#include "pme_common.h"
#include "pmq_internal.h"

namespace pmq {

// the grid as the kernels see it: sizes and A (u = A r, rows K_a times the
// reciprocal box rows)
struct PmeGeom {
    int K0, K1, K2;
    double A[3][3];
};

// fractional multipole: c, A d, A q A^T (off-diagonals doubled)
__device__ __forceinline__ void frac_mpole(const PmeGeom &g, const double *m, double *fm)
{
    fm[0] = m[0];
#pragma unroll
    for (int a = 0; a < 3; a++) fm[1 + a] = g.A[a][0] * m[1] + g.A[a][1] * m[2] + g.A[a][2] * m[3];
    const double Q[3][3] = {{m[4], m[5], m[6]}, {m[5], m[7], m[8]}, {m[6], m[8], m[9]}};
    double F[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            double v = 0;
#pragma unroll
            for (int p = 0; p < 3; p++)
                v += g.A[a][p] * (Q[p][0] * g.A[b][0] + Q[p][1] * g.A[b][1] + Q[p][2] * g.A[b][2]);
            F[a][b] = v;
        }
    fm[4] = F[0][0]; fm[5] = F[1][1]; fm[6] = F[2][2];
    fm[7] = 2 * F[0][1]; fm[8] = 2 * F[0][2]; fm[9] = 2 * F[1][2];
}

// Spreading is owner-computes over bricks of the grid.  Atoms are homed in
// the brick holding the last point of their stencil (floor(u) on each axis);
// every brick is at least order-1 points wide, so the stencil of an atom
// homed in brick h covers only bricks h-1 and h on each axis, and brick b
// receives exactly the atoms homed in b and b+1.  Each grid point is summed
// by one lane and stored once: no atomics and no zeroing of the grid.  With
// PME on, the atom sort orders atoms by the Hilbert index of their home brick
// (home_brick on the same coordinates), so each brick's atoms are one
// contiguous range of the sorted arrays, [bstart, bend).
struct BrickArgs {
    int n, nset;
    int nb[3];
    PmeGeom g;
    const int *bstart, *bend; // sorted-atom range of each brick (lexicographic brick index)
    int npad;
    const double *sx, *sy, *sz;
    const double *src;    // sorted SoA sources: 10 multipole comps (mode 0), 3 dipole comps x nset (mode 1)
                          // or one scalar per atom (mode 2: dispersion C6 coefficients)
    const double *charge; // mode 0: total monopole per sorted atom (charge penetration), or null
    void *grid;           // nset grids of the kernel's arithmetic type
    size_t gstride;
    const double *skip;   // nonzero: return at once (or null)
};

// One workgroup of 4 waves per brick; lane = grid point (bricks of more than
// 64 points are done in chunks of 64).  The atoms homed in the 8 candidate
// bricks are staged in LDS, 64 at a time (one thread each: stencil start,
// B-spline weights, fractional coefficients); wave w sums candidates w, w+4,
// ... for its lane's point, and the four partial sums are added in wave
// order, so the grid is bit-reproducible.
constexpr int SPREAD_CH = 64; // one wave: candidate j is prepared by lane j of each wave

// N values read or written by one LDS instruction (two or three for 32 and
// 48 bytes)
template <class T, int N>
struct alignas(N * sizeof(T) >= 16 ? 16 : N * sizeof(T)) LdsVec {
    T v[N];
};

// T: the arithmetic of the weights and sums (double in the plain form, float
// for the electrostatic grids of the mixed form); GT: the grid's type
// the body of workgroup wg of an nwg-workgroup launch (nwg a multiple of 8;
// 256 lanes)
template <int ORD, int MODE, class T, class GT = T>
__device__ __forceinline__ void spread_brick(const BrickArgs &a, int wg, int nwg)
{
    if (a.skip && *a.skip != 0.0) return;
    constexpr int ND = MODE == 0 ? 2 : MODE == 1 ? 1 : 0;
    constexpr int NC = MODE == 0 ? 10 : MODE == 1 ? 6 : 1;
    constexpr int NS = MODE == 1 ? 2 : 1;
    // per candidate j: th[axis][j][k] the weight and its derivatives at
    // stencil offset k (one vector read; the lanes of a wave, at different k
    // for the same j, read different banks), cf[j] the source coefficients,
    // si0[j] the stencil origin
    constexpr int DP = ND == 2 ? 4 : ND + 1;
    constexpr int NCP = MODE == 0 ? 12 : MODE == 1 ? 8 : 1;
    __shared__ LdsVec<T, DP> th[3][SPREAD_CH][ORD];
    __shared__ LdsVec<T, NCP> cf[SPREAD_CH];
    __shared__ int4 si0[SPREAD_CH];
    __shared__ T red[4][NS][64];
    GT *grid = (GT *)a.grid;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int K0 = a.g.K0, K1 = a.g.K1, K2 = a.g.K2;
    const int b = xcd_item(wg, nwg / PMQ_NXCD);
    if (b >= a.nb[0] * a.nb[1] * a.nb[2]) return;
    const int bx = b % a.nb[0], by = (b / a.nb[0]) % a.nb[1], bz = b / (a.nb[0] * a.nb[1]);
    const int lo0 = brick_lo(bx, K0, a.nb[0]), lo1 = brick_lo(by, K1, a.nb[1]), lo2 = brick_lo(bz, K2, a.nb[2]);
    const int s0 = brick_lo(bx + 1, K0, a.nb[0]) - lo0, s1 = brick_lo(by + 1, K1, a.nb[1]) - lo1,
              s2 = brick_lo(bz + 1, K2, a.nb[2]) - lo2;
    const int bsz = s0 * s1 * s2;
    const int nset = MODE == 1 ? a.nset : 1; // 1 or 2
    // the atoms homed in bricks b + (0|1, 0|1, 0|1), each brick once
    int hs[8], hc[8], ntot = 0;
#pragma unroll
    for (int h = 0; h < 8; h++) {
        const int dx = h & 1, dy = (h >> 1) & 1, dz = h >> 2;
        hs[h] = 0;
        hc[h] = 0;
        if ((dx && a.nb[0] == 1) || (dy && a.nb[1] == 1) || (dz && a.nb[2] == 1)) continue;
        const int hx = (bx + dx) % a.nb[0], hy = (by + dy) % a.nb[1], hz = (bz + dz) % a.nb[2];
        const int hb = (hz * a.nb[1] + hy) * a.nb[0] + hx;
        hs[h] = a.bstart[hb];
        hc[h] = a.bend[hb] - hs[h];
        ntot += hc[h];
    }
    for (int pbase = 0; pbase < bsz; pbase += 64) {
        const int p = pbase + lane;
        const bool pon = p < bsz;
        const int lx = p % s0, ly = (p / s0) % s1, lz = p / (s0 * s1);
        const int gx = lo0 + lx, gy = lo1 + ly, gz = lo2 + lz;
        T acc[NS];
#pragma unroll
        for (int q = 0; q < NS; q++) acc[q] = T(0);
        for (int base = 0; base < ntot; base += SPREAD_CH) {
            const int cnt = min(SPREAD_CH, ntot - base);
            __syncthreads();
            // candidate preparation split over the four waves: wave ax < 3
            // forms the axis-ax weights of candidate base + lane, wave 3 its
            // source coefficients
            if (lane < cnt) {
                // position of candidate base + lane in the concatenated home-brick ranges
                int e = 0, off = 0;
                const int idx = base + lane;
#pragma unroll
                for (int h = 0; h < 8; h++) {
                    if (idx >= off && idx < off + hc[h]) e = hs[h] + idx - off;
                    off += hc[h];
                }
                const int s = e;
                if (wv < 3) {
                    // the per-axis steps of atom_frac (same grid_coord, so the
                    // same stencil origin as the gathers and the brick sort)
                    const double *Ar = wv == 0 ? a.g.A[0] : wv == 1 ? a.g.A[1] : a.g.A[2];
                    const int Ka = wv == 0 ? K0 : wv == 1 ? K1 : K2;
                    const double u = grid_coord(Ar, a.sx[s], a.sy[s], a.sz[s]);
                    const double fl = floor(u);
                    int i0 = ((int)fl - ORD + 1) % Ka;
                    if (i0 < 0) i0 += Ka;
                    // weights in FP64 whatever T is (FP64 costs what FP32
                    // does here and keeps the mixed form's energy error down)
                    double t[PMQ_MAX_ORDER][4];
                    bspline_weights<double, ND>(u - fl, ORD, t);
#pragma unroll
                    for (int k = 0; k < ORD; k++)
#pragma unroll
                        for (int d = 0; d <= ND; d++) th[wv][lane][k].v[d] = t[k][d];
                    (wv == 0 ? si0[lane].x : wv == 1 ? si0[lane].y : si0[lane].z) = i0;
                } else if (MODE == 0) {
                    double m[10], fm[10];
#pragma unroll
                    for (int c = 0; c < 10; c++) m[c] = a.src[c * a.npad + s];
                    if (a.charge) m[0] = a.charge[s];
                    frac_mpole(a.g, m, fm);
#pragma unroll
                    for (int c = 0; c < 10; c++) cf[lane].v[c] = T(fm[c]);
                } else if (MODE == 2) {
                    cf[lane].v[0] = T(a.src[s]);
                } else {
                    for (int set = 0; set < nset; set++) {
                        const double *d = a.src + (size_t)set * 3 * a.npad;
                        const double ux = d[s], uy = d[a.npad + s], uz = d[2 * a.npad + s];
#pragma unroll
                        for (int r = 0; r < 3; r++)
                            cf[lane].v[3 * set + r] = T(a.g.A[r][0] * ux + a.g.A[r][1] * uy + a.g.A[r][2] * uz);
                    }
                }
            }
            __syncthreads();
            if (!pon) continue;
            for (int j = wv; j < cnt; j += 4) {
                const int4 o = si0[j];
                int kx = gx - o.x, ky = gy - o.y, kz = gz - o.z;
                if (kx < 0) kx += K0;
                if (ky < 0) ky += K1;
                if (kz < 0) kz += K2;
                if (kx >= ORD || ky >= ORD || kz >= ORD) continue;
                const LdsVec<T, DP> X = th[0][j][kx], Y = th[1][j][ky], Z = th[2][j][kz];
                const LdsVec<T, NCP> C = cf[j];
                if (MODE == 0) {
                    const T x0 = X.v[0], x1 = X.v[1], x2 = X.v[2 % DP];
                    const T y0 = Y.v[0], y1 = Y.v[1], y2 = Y.v[2 % DP];
                    const T z0 = Z.v[0], z1 = Z.v[1], z2 = Z.v[2 % DP];
                    acc[0] += C.v[0] * x0 * y0 * z0 + C.v[1] * x1 * y0 * z0 + C.v[2] * x0 * y1 * z0
                            + C.v[3] * x0 * y0 * z1 + C.v[4] * x2 * y0 * z0 + C.v[5] * x0 * y2 * z0
                            + C.v[6] * x0 * y0 * z2 + C.v[7] * x1 * y1 * z0 + C.v[8] * x1 * y0 * z1
                            + C.v[9] * x0 * y1 * z1;
                } else if (MODE == 2) {
                    acc[0] += C.v[0] * X.v[0] * Y.v[0] * Z.v[0];
                } else {
                    const T x0 = X.v[0], x1 = X.v[1 % DP];
                    const T y0 = Y.v[0], y1 = Y.v[1 % DP];
                    const T z0 = Z.v[0], z1 = Z.v[1 % DP];
                    const T w100 = x1 * y0 * z0, w010 = x0 * y1 * z0, w001 = x0 * y0 * z1;
#pragma unroll
                    for (int q = 0; q < NS; q++)
                        acc[q] += C.v[(3 * q) % NCP] * w100 + C.v[(3 * q + 1) % NCP] * w010
                                + C.v[(3 * q + 2) % NCP] * w001;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NS; q++) red[wv][q][lane] = acc[q];
        __syncthreads();
        if (wv == 0 && pon)
            for (int q = 0; q < nset; q++)
                grid[q * a.gstride + ((size_t)gz * K1 + gy) * K0 + gx] =
                    GT(((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) + red[3][q][lane]);
    }
}

} // namespace pmq
