// k_pme.hip -- smooth particle-mesh Ewald for multipoles and induced dipoles.
//
// Algorithm (Essmann et al. 1995; Sagui, Pedersen and Darden 2004): every
// site's multipole is expressed in fractional grid coordinates u = A r
// (A = K * reciprocal box), spread onto the K0 x K1 x K2 grid with order-p
// cardinal B-splines and their derivatives, transformed with rocFFT,
// multiplied by the reciprocal kernel
//     G(m) = exp(-pi^2 h^2 / a^2) / (pi V h^2 |b1 b2 b3|^2)
// and transformed back to the potential; the potential and its derivatives
// (up to third order) are then interpolated back at every site.
//
// Grids: real grid index (g3*K1 + g2)*K0 + g1 (g1 fastest), complex half
// grid ((K0/2+1) along g1).  Spreading is owner-computes over grid bricks
// (LDS accumulation, no global atomics); gathering uses one lane group per
// atom.
#include <rocfft/rocfft.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstring>

#include "pair_matrix.h"
#include "pme_common.h"
#include "pme_spread.h"
#include "pmq_internal.h"

namespace pmq {

// ---------------------------------------------------------------------------
// host setup: A matrix, B-spline moduli, reciprocal kernel, FFT plans
static double bspline_m(int n, double x)
{
    if (x <= 0.0 || x >= n) return 0.0;
    if (n == 2) return 1.0 - fabs(x - 1.0);
    return (x * bspline_m(n - 1, x) + (n - x) * bspline_m(n - 1, x - 1.0)) / (n - 1);
}

static int fft_check(pmq_ctx *c, rocfft_status s, const char *what)
{
    if (s != rocfft_status_success) {
        c->err = std::string("rocFFT: ") + what + " failed";
        return PMQ_ERR_HIP;
    }
    return PMQ_OK;
}

static void destroy_plans(pmq_ctx *c)
{
    void **plans[8] = {&c->plan_f1, &c->plan_b1, &c->plan_f2, &c->plan_b2,
                       &c->plan_f1s, &c->plan_b1s, &c->plan_f2s, &c->plan_b2s};
    for (auto p : plans)
        if (*p) {
            rocfft_plan_destroy((rocfft_plan)*p);
            *p = nullptr;
        }
    if (c->fft_info) {
        rocfft_execution_info_destroy((rocfft_execution_info)c->fft_info);
        c->fft_info = nullptr;
    }
}

void pme_release(pmq_ctx *c)
{
    destroy_plans(c);
    if (c->pme_side) {
        (void)hipStreamSynchronize(c->pme_side);
        (void)hipStreamDestroy(c->pme_side);
        c->pme_side = nullptr;
    }
    if (c->pme_fork) (void)hipEventDestroy(c->pme_fork);
    if (c->pme_join) (void)hipEventDestroy(c->pme_join);
    c->pme_fork = c->pme_join = nullptr;
}

// brick partition of one grid axis for spreading: bricks at least
// max(4, order - 1) points wide
static int brick_count(int K, int order) { return std::max(1, K / std::max(4, order - 1)); }

// psi(k) (2 pi^1.5 a^3 / 3)^-1: the Fourier transform of the long-range
// part (1 - g(a r)) r^-6 of dispersion, g(x) = (1 + x^2 + x^4/2) e^{-x^2},
// b = k / (2a)
static double disp_psi_unit(double b)
{
    return 1.77245385090551602730 * b * b * b * erfc(b) + (0.5 - b * b) * exp(-b * b);
}

int pme_setup(pmq_ctx *c)
{
    const pmq_params &p = c->prm;
    const bool ew = p.use_ewald, dew = c->have_disp && p.use_dewald;
    if (ew && dew && (p.dfft1 != p.nfft1 || p.dfft2 != p.nfft2 || p.dfft3 != p.nfft3 || p.dbsorder != p.bsorder)) {
        c->err = "PME: dispersion and multipole PME must share the grid and B-spline order";
        return PMQ_ERR_UNSUPPORTED;
    }
    const int K[3] = {ew ? p.nfft1 : p.dfft1, ew ? p.nfft2 : p.dfft2, ew ? p.nfft3 : p.dfft3};
    const int order = ew ? p.bsorder : p.dbsorder;
    if (K[0] < 8 || K[1] < 8 || K[2] < 8 || order < 4 || order > PMQ_MAX_ORDER) {
        c->err = "PME: grid sizes must be >= 8 and the B-spline order 4..6";
        return PMQ_ERR_ARG;
    }
    if (c->box.kind != PMQ_BOX_ORTHO && c->box.kind != PMQ_BOX_MONO && c->box.kind != PMQ_BOX_TRI) {
        c->err = "PME: Ewald summation needs an orthogonal, monoclinic or triclinic box";
        return PMQ_ERR_ARG;
    }
    const double alpha = ew ? p.aewald : 0.0, dalpha = dew ? p.dewald : 0.0;
    const bool same_grid = K[0] == c->K[0] && K[1] == c->K[1] && K[2] == c->K[2] && c->bsorder == order;
    bool same_box = same_grid && c->pme_alpha == alpha && c->pme_dalpha == dalpha && c->pme_dbox.kind == c->box.kind;
    for (int a = 0; a < 3 && same_box; a++)
        for (int b = 0; b < 3; b++) same_box = same_box && c->pme_dbox.recip[a][b] == c->box.recip[a][b];
    PMQ_HIP(c->pme_frac.alloc(40 * (size_t)c->npad)); // follows the atom count
    if (same_box && c->plan_f1) return PMQ_OK;
    c->K[0] = K[0]; c->K[1] = K[1]; c->K[2] = K[2];
    c->bsorder = order;
    for (int a = 0; a < 3; a++) c->pme_nb[a] = brick_count(K[a], order);
    PMQ_HIP(c->pme_bstart.alloc((size_t)c->pme_nb[0] * c->pme_nb[1] * c->pme_nb[2]));
    PMQ_HIP(c->pme_bend.alloc((size_t)c->pme_nb[0] * c->pme_nb[1] * c->pme_nb[2]));
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) c->pmeA[a][b] = K[a] * c->box.recip[a][b];
    // B-spline moduli, zeros replaced by the mean of their neighbours
    std::vector<double> bsmod[3];
    for (int a = 0; a < 3; a++) {
        bsmod[a].assign(K[a], 0.0);
        for (int m = 0; m < K[a]; m++) {
            double re = 0, im = 0;
            for (int k = 0; k < order - 1; k++) {
                const double w = bspline_m(order, k + 1.0);
                const double arg = 2.0 * M_PI * m * k / K[a];
                re += w * cos(arg);
                im += w * sin(arg);
            }
            bsmod[a][m] = re * re + im * im;
        }
        for (int m = 0; m < K[a]; m++)
            if (bsmod[a][m] < 1e-7) bsmod[a][m] = 0.5 * (bsmod[a][(m - 1 + K[a]) % K[a]] + bsmod[a][(m + 1) % K[a]]);
    }
    const double (*A)[3] = c->pmeA;
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
                     + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    const double V = (double)K[0] * K[1] * K[2] / det;
    const int H0 = K[0] / 2 + 1;
    // electrostatics: G = exp(-pi^2 h^2/a^2) / (pi V h^2 bsmod), m = 0 dropped;
    // dispersion: Gd = -psi(2 pi |h|) / (V bsmod), m = 0 kept.  Index m maps
    // to the frequency m or m - K (m > K/2).  In a skewed cell G(m) and G(-m)
    // differ where a component sits at K/2 (that index stands for +K/2 on both
    // sides, while the cross terms of h^2 change sign), and the real part of
    // a full complex convolution equals a convolution with their mean, which
    // is what the tables hold: the half-complex transforms then give the real
    // potential of the full complex sum (and the same result whichever sign
    // K/2 is given).
    std::vector<double> G((size_t)H0 * K[1] * K[2]), Gd(dew ? G.size() : 0);
    const double a2 = alpha * alpha;
    const double dpre = 2.0 * pow(M_PI, 1.5) * dalpha * dalpha * dalpha / 3.0;
    auto kernels = [&](int m1, int m2, int m3, double *ge, double *gd) {
        const int f1 = m1 > K[0] / 2 ? m1 - K[0] : m1;
        const int f2 = m2 > K[1] / 2 ? m2 - K[1] : m2;
        const int f3 = m3 > K[2] / 2 ? m3 - K[2] : m3;
        double h[3];
        for (int j = 0; j < 3; j++) h[j] = f1 * A[0][j] / K[0] + f2 * A[1][j] / K[1] + f3 * A[2][j] / K[2];
        const double h2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
        const double bs = bsmod[0][m1] * bsmod[1][m2] * bsmod[2][m3];
        *gd = dew ? -dpre * disp_psi_unit(M_PI * sqrt(h2) / dalpha) / (V * bs) : 0.0;
        *ge = (ew && (m1 || m2 || m3)) ? exp(-M_PI * M_PI * h2 / a2) / (M_PI * V * h2 * bs) : 0.0;
    };
    for (int m3 = 0; m3 < K[2]; m3++)
        for (int m2 = 0; m2 < K[1]; m2++)
            for (int m1 = 0; m1 < H0; m1++) {
                const size_t id = ((size_t)m3 * K[1] + m2) * H0 + m1;
                double ge, gd, ge2, gd2;
                kernels(m1, m2, m3, &ge, &gd);
                kernels((K[0] - m1) % K[0], (K[1] - m2) % K[1], (K[2] - m3) % K[2], &ge2, &gd2);
                G[id] = ge == ge2 ? ge : 0.5 * (ge + ge2);
                if (dew) Gd[id] = gd == gd2 ? gd : 0.5 * (gd + gd2);
            }
    {
        std::vector<double> bsm;
        for (int a = 0; a < 3; a++) bsm.insert(bsm.end(), bsmod[a].begin(), bsmod[a].end());
        PMQ_HIP(c->pme_bsmod.alloc(bsm.size()));
        PMQ_HIP(hipMemcpyAsync(c->pme_bsmod.p, bsm.data(), bsm.size() * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
        PMQ_HIP(hipStreamSynchronize(c->stream)); // bsm is a host temporary
    }
    PMQ_HIP(c->pme_g.alloc(G.size()));
    PMQ_HIP(hipMemcpyAsync(c->pme_g.p, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<float> Gs(G.begin(), G.end());
    PMQ_HIP(c->pme_g_s.alloc(Gs.size()));
    PMQ_HIP(hipMemcpyAsync(c->pme_g_s.p, Gs.data(), Gs.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (dew) {
        PMQ_HIP(c->pme_gd.alloc(Gd.size()));
        PMQ_HIP(hipMemcpyAsync(c->pme_gd.p, Gd.data(), Gd.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    PMQ_HIP(hipStreamSynchronize(c->stream)); // G, Gd are host temporaries
    const size_t nreal = (size_t)K[0] * K[1] * K[2], ncplx = (size_t)H0 * K[1] * K[2];
    PMQ_HIP(c->qgrid.alloc(2 * nreal));
    PMQ_HIP(c->cgrid.alloc(2 * 2 * ncplx));
    // the second stream of pme_mpole_fields (never created under capture)
    if (!c->pme_side) PMQ_HIP(hipStreamCreateWithFlags(&c->pme_side, hipStreamNonBlocking));
    if (!c->pme_fork) PMQ_HIP(hipEventCreateWithFlags(&c->pme_fork, hipEventDisableTiming));
    if (!c->pme_join) PMQ_HIP(hipEventCreateWithFlags(&c->pme_join, hipEventDisableTiming));
    if (!same_grid || !c->plan_f1) {
        destroy_plans(c);
        int rc;
        if ((rc = fft_check(c, rocfft_setup(), "setup"))) return rc;
        const size_t len[3] = {(size_t)K[0], (size_t)K[1], (size_t)K[2]};
        rocfft_plan pf1, pb1, pf2, pb2;
        if ((rc = fft_check(c, rocfft_plan_create(&pf1, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                                  rocfft_precision_double, 3, len, 1, nullptr), "plan"))) return rc;
        c->plan_f1 = pf1;
        if ((rc = fft_check(c, rocfft_plan_create(&pb1, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                                  rocfft_precision_double, 3, len, 1, nullptr), "plan"))) return rc;
        c->plan_b1 = pb1;
        if ((rc = fft_check(c, rocfft_plan_create(&pf2, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                                  rocfft_precision_double, 3, len, 2, nullptr), "plan"))) return rc;
        c->plan_f2 = pf2;
        if ((rc = fft_check(c, rocfft_plan_create(&pb2, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                                  rocfft_precision_double, 3, len, 2, nullptr), "plan"))) return rc;
        c->plan_b2 = pb2;
        rocfft_plan sp[4];
        const rocfft_transform_type tt[4] = {rocfft_transform_type_real_forward, rocfft_transform_type_real_inverse,
                                             rocfft_transform_type_real_forward, rocfft_transform_type_real_inverse};
        for (int q = 0; q < 4; q++)
            if ((rc = fft_check(c, rocfft_plan_create(&sp[q], rocfft_placement_notinplace, tt[q], rocfft_precision_single,
                                                      3, len, q < 2 ? 1 : 2, nullptr), "plan"))) return rc;
        c->plan_f1s = sp[0]; c->plan_b1s = sp[1]; c->plan_f2s = sp[2]; c->plan_b2s = sp[3];
        c->pme_fft32 = -1; // to be timed
        c->pme_own64 = c->pme_own32 = -1;
        if ((rc = fft_own_setup(c))) return rc;
        size_t wmax = 0;
        for (rocfft_plan pl : {pf1, pb1, pf2, pb2, sp[0], sp[1], sp[2], sp[3]}) {
            size_t w = 0;
            rocfft_plan_get_work_buffer_size(pl, &w);
            wmax = std::max(wmax, w);
        }
        PMQ_HIP(c->fft_work.alloc(wmax));
        rocfft_execution_info info;
        if ((rc = fft_check(c, rocfft_execution_info_create(&info), "info"))) return rc;
        c->fft_info = info;
        if ((rc = fft_check(c, rocfft_execution_info_set_stream(info, c->stream), "stream"))) return rc;
        if (wmax && (rc = fft_check(c, rocfft_execution_info_set_work_buffer(info, c->fft_work.p, wmax), "work"))) return rc;
    }
    c->pme_dbox = c->box;
    c->pme_alpha = alpha;
    c->pme_dalpha = dalpha;
    return PMQ_OK;
}

// ---------------------------------------------------------------------------
// kernels
//
// Gathering: one group of G lanes serves one atom and reduces with butterfly
// shuffles.  Planes (G = 8): lane kz < order sums the order x order points of
// stencil plane kz, the others idle; rows (G = 32, or 64 for order 6): lane
// (ky, kz) sums one row of `order` points.  Planes form the B-spline weights
// once per plane instead of once per row, rows give four times the lanes
// for small systems.  Atoms go to workgroups in sorted (spatial) order,
// each XCD taking a contiguous eighth, so the grid points a workgroup reads
// are shared with its neighbours in the same L2.
static PmeGeom geom(const pmq_ctx *c)
{
    PmeGeom g;
    g.K0 = c->K[0]; g.K1 = c->K[1]; g.K2 = c->K[2];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) g.A[a][b] = c->pmeA[a][b];
    return g;
}


// stencil start and fractional offset along each axis
__device__ __forceinline__ void atom_frac(const PmeGeom &g, double px, double py, double pz, double *w, int *i0, int ord)
{
    const double p[3] = {px, py, pz};
    const int K[3] = {g.K0, g.K1, g.K2};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double u = grid_coord(g.A[a], p[0], p[1], p[2]);
        const double fl = floor(u);
        w[a] = u - fl;
        int s = ((int)fl - ord + 1) % K[a];
        if (s < 0) s += K[a];
        i0[a] = s;
    }
}

// weight set of one axis at stencil index k (selected without dynamic
// register indexing)
template <int ORD, int ND, class T>
__device__ __forceinline__ void pick(const T th[PMQ_MAX_ORDER][4], int k, T *out)
{
#pragma unroll
    for (int d = 0; d <= ND; d++) out[d] = T(0);
#pragma unroll
    for (int q = 0; q < ORD; q++)
        if (q == k) {
#pragma unroll
            for (int d = 0; d <= ND; d++) out[d] = th[q][d];
        }
}


// the standalone spreading launch (body in pme_spread.h)
template <int ORD, int MODE, class T, class GT = T>
__global__ __launch_bounds__(256) void k_pme_spread_brick(BrickArgs a)
{
    spread_brick<ORD, MODE, T, GT>(a, blockIdx.x, gridDim.x);
}

template <class T>
__global__ void k_pme_convolve(int ncplx, int nset, const T *__restrict__ G, T *__restrict__ cg, const double *skip)
{
    if (skip && *skip != 0.0) return;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)ncplx * nset) return;
    const T g = G[t % ncplx];
    cg[2 * t] *= g;
    cg[2 * t + 1] *= g;
}

// this lane's share of the potential derivatives at one site, in the
// fractional order 0; 1,2,3; 11,22,33,12,13,23; 111,222,333,112,113,122,133,
// 223,233,123 (NOUT = 4, 10 or 20): the plane kz of the stencil, all ky and
// kx; idle lanes (`on` false) give zeros
template <int ORD, int NOUT, class T, class GT>
__device__ __forceinline__ void gather_part(const PmeGeom &g, const GT *grid, const T tx[PMQ_MAX_ORDER][4],
                                            const T ty[PMQ_MAX_ORDER][4], const T *tz, const int *i0, int kz,
                                            bool on, double *o)
{
#pragma unroll
    for (int j = 0; j < NOUT; j++) o[j] = 0.0;
    if (!on) return;
    int gz = i0[2] + kz;
    if (gz >= g.K2) gz -= g.K2;
    T acc[NOUT];
#pragma unroll
    for (int j = 0; j < NOUT; j++) acc[j] = T(0);
#pragma unroll
    for (int ky = 0; ky < ORD; ky++) {
        int gy = i0[1] + ky;
        if (gy >= g.K1) gy -= g.K1;
        const GT *row = grid + ((size_t)gz * g.K1 + gy) * g.K0;
        T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
        for (int kx = 0; kx < ORD; kx++) {
            int gx = i0[0] + kx;
            if (gx >= g.K0) gx -= g.K0;
            const T q = T(row[gx]);
            s0 += q * tx[kx][0];
            s1 += q * tx[kx][1];
            if (NOUT > 4) s2 += q * tx[kx][2];
            if (NOUT > 10) s3 += q * tx[kx][3];
        }
        const T *y = ty[ky];
        acc[0] += s0 * y[0] * tz[0];
        acc[1] += s1 * y[0] * tz[0];
        acc[2] += s0 * y[1] * tz[0];
        acc[3] += s0 * y[0] * tz[1];
        if (NOUT > 4) {
            acc[4] += s2 * y[0] * tz[0];
            acc[5] += s0 * y[2] * tz[0];
            acc[6] += s0 * y[0] * tz[2];
            acc[7] += s1 * y[1] * tz[0];
            acc[8] += s1 * y[0] * tz[1];
            acc[9] += s0 * y[1] * tz[1];
        }
        if (NOUT > 10) {
            acc[10] += s3 * y[0] * tz[0];
            acc[11] += s0 * y[3] * tz[0];
            acc[12] += s0 * y[0] * tz[3];
            acc[13] += s2 * y[1] * tz[0];
            acc[14] += s2 * y[0] * tz[1];
            acc[15] += s1 * y[2] * tz[0];
            acc[16] += s1 * y[0] * tz[2];
            acc[17] += s0 * y[2] * tz[1];
            acc[18] += s0 * y[1] * tz[2];
            acc[19] += s1 * y[1] * tz[1];
        }
    }
#pragma unroll
    for (int j = 0; j < NOUT; j++) o[j] = acc[j];
}

// the row variant: lane (ky, kz) sums the `order` points of one stencil row
template <int ORD, int NOUT, class T, class GT>
__device__ __forceinline__ void gather_row(const PmeGeom &g, const GT *grid, const T tx[PMQ_MAX_ORDER][4],
                                           const T *ty, const T *tz, const int *i0, int ky, int kz, bool on,
                                           double *o)
{
#pragma unroll
    for (int j = 0; j < NOUT; j++) o[j] = 0.0;
    if (!on) return;
    int gy = i0[1] + ky, gz = i0[2] + kz;
    if (gy >= g.K1) gy -= g.K1;
    if (gz >= g.K2) gz -= g.K2;
    const GT *row = grid + ((size_t)gz * g.K1 + gy) * g.K0;
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int kx = 0; kx < ORD; kx++) {
        int gx = i0[0] + kx;
        if (gx >= g.K0) gx -= g.K0;
        const T q = T(row[gx]);
        s0 += q * tx[kx][0];
        s1 += q * tx[kx][1];
        if (NOUT > 4) s2 += q * tx[kx][2];
        if (NOUT > 10) s3 += q * tx[kx][3];
    }
    o[0] = s0 * ty[0] * tz[0];
    o[1] = s1 * ty[0] * tz[0];
    o[2] = s0 * ty[1] * tz[0];
    o[3] = s0 * ty[0] * tz[1];
    if (NOUT > 4) {
        o[4] = s2 * ty[0] * tz[0];
        o[5] = s0 * ty[2] * tz[0];
        o[6] = s0 * ty[0] * tz[2];
        o[7] = s1 * ty[1] * tz[0];
        o[8] = s1 * ty[0] * tz[1];
        o[9] = s0 * ty[1] * tz[1];
    }
    if (NOUT > 10) {
        o[10] = s3 * ty[0] * tz[0];
        o[11] = s0 * ty[3] * tz[0];
        o[12] = s0 * ty[0] * tz[3];
        o[13] = s2 * ty[1] * tz[0];
        o[14] = s2 * ty[0] * tz[1];
        o[15] = s1 * ty[2] * tz[0];
        o[16] = s1 * ty[0] * tz[2];
        o[17] = s0 * ty[2] * tz[1];
        o[18] = s0 * ty[1] * tz[2];
        o[19] = s1 * ty[1] * tz[1];
    }
}

// one site's share for lane l of a G-lane group: G = 8, the plane kz = l;
// otherwise the row (ky, kz) = (l % ORD, l / ORD)
template <int ORD, int G, int NOUT, class T, class GT>
__device__ __forceinline__ void gather_share(const PmeGeom &g, const GT *grid, const T tx[PMQ_MAX_ORDER][4],
                                             const T ty[PMQ_MAX_ORDER][4], const T *ty1, const T *tz,
                                             const int *i0, int ky, int kz, bool on, double *o)
{
    if constexpr (G == 8) gather_part<ORD, NOUT, T, GT>(g, grid, tx, ty, tz, i0, kz, on, o);
    else gather_row<ORD, NOUT, T, GT>(g, grid, tx, ty1, tz, i0, ky, kz, on, o);
}

template <int G, int NOUT>
__device__ __forceinline__ void group_reduce(double *o)
{
#pragma unroll
    for (int j = 0; j < NOUT; j++) {
        double v = o[j];
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        o[j] = v;
    }
}

// fractional derivatives -> Cartesian phi, g[3], H (xx,yy,zz,xy,xz,yz), T[10]
// (xxx,yyy,zzz,xxy,xxz,xyy,xzz,yyz,yzz,xyz)
template <int NOUT>
__device__ __forceinline__ void frac_to_cart(const PmeGeom &g, const double *f, double *phi, double *grad, double *H,
                                             double *T)
{
    *phi = f[0];
#pragma unroll
    for (int x = 0; x < 3; x++) grad[x] = g.A[0][x] * f[1] + g.A[1][x] * f[2] + g.A[2][x] * f[3];
    if (NOUT > 4) {
        const double F2[3][3] = {{f[4], f[7], f[8]}, {f[7], f[5], f[9]}, {f[8], f[9], f[6]}};
        const int hx[6] = {0, 1, 2, 0, 0, 1}, hy[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            double v = 0;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) v += g.A[a][hx[q]] * g.A[b][hy[q]] * F2[a][b];
            H[q] = v;
        }
    }
    if (NOUT > 10) {
        double F3[3][3][3];
        const int idx[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                                {0, 1, 1}, {0, 2, 2}, {1, 1, 2}, {1, 2, 2}, {0, 1, 2}};
#pragma unroll
        for (int q = 0; q < 10; q++) {
            const int a = idx[q][0], b = idx[q][1], d = idx[q][2];
            const double v = f[10 + q];
            F3[a][b][d] = F3[a][d][b] = F3[b][a][d] = F3[b][d][a] = F3[d][a][b] = F3[d][b][a] = v;
        }
#pragma unroll
        for (int q = 0; q < 10; q++) {
            const int x = idx[q][0], y = idx[q][1], z = idx[q][2];
            double v = 0;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
#pragma unroll
                    for (int d = 0; d < 3; d++) v += g.A[a][x] * g.A[b][y] * g.A[d][z] * F3[a][b][d];
            T[q] = v;
        }
    }
}

// common lane setup of the gather kernels
#define GATHER_SETUP(ND, T)                                                                                  \
    constexpr int G = GL;                                                                                    \
    const int slot = xcd_item(blockIdx.x, gridDim.x / PMQ_NXCD) * (blockDim.x / G) + threadIdx.x / G;        \
    const int l = threadIdx.x % G;                                                                           \
    const bool valid = slot < a.n;                                                                           \
    const bool on = valid && l < (G == 8 ? ORD : ORD * ORD);                                                 \
    const int s = valid ? slot : 0;                                                                          \
    const int ky = G == 8 ? 0 : l % ORD, kz = G == 8 ? (l < ORD ? l : 0) : (l / ORD) % ORD;                  \
    double w[3];                                                                                             \
    int i0[3];                                                                                               \
    atom_frac(a.g, a.sx[s], a.sy[s], a.sz[s], w, i0, ORD);                                                   \
    T tx[PMQ_MAX_ORDER][4], ty[PMQ_MAX_ORDER][4], ty1[4], tz[4];                                             \
    {                                                                                                        \
        /* weights in FP64 whatever T is, then rounded to T */                                               \
        double bx[PMQ_MAX_ORDER][4], by[PMQ_MAX_ORDER][4], bzf[PMQ_MAX_ORDER][4], bz[4], by1[4];             \
        bspline_weights<double, ND>(w[0], ORD, bx);                                                          \
        bspline_weights<double, ND>(w[1], ORD, by);                                                          \
        bspline_weights<double, ND>(w[2], ORD, bzf);                                                         \
        pick<ORD, ND, double>(bzf, kz, bz);                                                                  \
        pick<ORD, ND, double>(by, ky, by1);                                                                  \
        _Pragma("unroll") for (int k = 0; k < ORD; k++)                                                      \
            _Pragma("unroll") for (int d = 0; d <= ND; d++) { tx[k][d] = T(bx[k][d]); ty[k][d] = T(by[k][d]); } \
        _Pragma("unroll") for (int d = 0; d <= ND; d++) { tz[d] = T(bz[d]); ty1[d] = T(by1[d]); }          \
    }

struct GatherFracArgs {
    int n, nset;
    PmeGeom g;
    const double *sx, *sy, *sz;
    const void *grid;      // nset grids (double, or float in the mixed form)
    size_t gstride;
    double *frac;          // n x nset x 20 fractional derivatives, sorted atom order
};

// this lane's share of the Cartesian potential gradient from its share of
// the fractional derivatives (f[1..3]; the map of frac_to_cart)
template <int G>
__device__ __forceinline__ void frac_grad(const PmeGeom &g, const double *f, double *gc)
{
#pragma unroll
    for (int x = 0; x < 3; x++) gc[x] = g.A[0][x] * f[1] + g.A[1][x] * f[2] + g.A[2][x] * f[3];
}

// Sums over the G lanes of a group of N values each, scattered: lane l
// ends with the totals of components l, l + G, l + 2G, ... (< N) in
// o[0], o[1], ...  Each stage halves the components a lane keeps and trades
// the other half with the lane `h` away, G - 1 shuffles per G components
// instead of log2(G) per component for a full all-reduce (20 values over
// 32 lanes: 31 shuffles instead of 100).  A fixed order, so the sums are
// reproducible.
template <int G, int N>
__device__ __forceinline__ void group_reduce_scatter(double *o, int l)
{
    constexpr int Q = (N + G - 1) / G;
    double w[Q][G];
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int r = 0; r < G; r++) w[q][r] = q * G + r < N ? o[q * G + r] : 0.0;
#pragma unroll
    for (int h = G / 2; h > 0; h >>= 1) {
        const bool hi = (l & h) != 0;
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
            for (int k = 0; k < h; k++) {
                const double keep = hi ? w[q][h + k] : w[q][k], send = hi ? w[q][k] : w[q][h + k];
                w[q][k] = keep + __shfl_xor(send, h, 64);
            }
    }
#pragma unroll
    for (int q = 0; q < Q; q++) o[q] = w[q][0];
}

// the 20 fractional potential derivatives of every atom in each of nset
// grids; lane l of the group stores components l, l + G, l + 2G < 20
template <int ORD, int GL, class T, class GT>
__global__ __launch_bounds__(256) void k_pme_gather_frac(GatherFracArgs a)
{
    GATHER_SETUP(3, T)
    for (int set = 0; set < a.nset; set++) {
        double fp[20];
        gather_share<ORD, G, 20, T, GT>(a.g, (const GT *)a.grid + set * a.gstride, tx, ty, ty1, tz, i0, ky, kz, on, fp);
        group_reduce_scatter<G, 20>(fp, l);
#pragma unroll
        for (int q = 0; q * G < 20; q++)
            if (valid && q * G + l < 20) a.frac[((size_t)s * a.nset + set) * 20 + q * G + l] = fp[q];
    }
}

struct MpolePostArgs {
    int n, npad;
    PmeGeom g;
    const double *smp, *charge;
    const int *perm;
    const double *frac;      // n x 20
    double f, aewald;
    int grad, polar;
    double *dem, *te;        // caller order
    double *pot;             // charge flux: += dE/dq (reciprocal potential and self term), caller order, or null
    double *cphi;            // 10 * npad sorted (phi, g, H)
    double *fieldd, *fieldp; // 3 * npad sorted (+= reciprocal + self field)
    Totals *tot;
    int vir;                 // the strain term of the dipoles and quadrupoles (with tot)
    double *aem;             // analysis: per-atom partition (caller order, +=), or null
};

// the strain term of a dipole d and quadrupole Q (the reference's traceless
// x1/3 form) in a reciprocal potential of gradient g and Hessian H: the
// Cartesian components stay fixed while the reciprocal vectors strain,
// which adds -s [d_b g_a + 2 Q_bc H_ac] (symmetrized) to the virial
__device__ __forceinline__ void vir_cterm(Vir6 &v, const double *d, const double Q[3][3], const double *g,
                                          const double H[3][3], double s)
{
    double W[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
            W[p][q] = d[q] * g[p] + 2.0 * (Q[q][0] * H[p][0] + Q[q][1] * H[p][1] + Q[q][2] * H[p][2]);
    v.xx -= s * W[0][0];
    v.yy -= s * W[1][1];
    v.zz -= s * W[2][2];
    v.xy -= 0.5 * s * (W[0][1] + W[1][0]);
    v.xz -= 0.5 * s * (W[0][2] + W[2][0]);
    v.yz -= 0.5 * s * (W[1][2] + W[2][1]);
}

// permanent multipoles in their own reciprocal potential (one thread per
// sorted atom): energy (f/2) U, gradient, torque, the self energy, and the
// reciprocal + self field the induced dipoles see
__global__ __launch_bounds__(256) void k_pme_mpole_post(MpolePostArgs a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    Vir6 vir;
    if (s < a.n) {
        double fp[20];
        for (int j = 0; j < 20; j++) fp[j] = a.frac[(size_t)s * 20 + j];
        double phi, gr[3], H[6], T[10];
        frac_to_cart<20>(a.g, fp, &phi, gr, H, T);
        double m[10];
        for (int c = 0; c < 10; c++) m[c] = a.smp[c * a.npad + s];
        if (a.charge) m[0] = a.charge[s];
        const double Hm[3][3] = {{H[0], H[3], H[4]}, {H[3], H[1], H[5]}, {H[4], H[5], H[2]}};
        const double Q[3][3] = {{m[4], m[5], m[6]}, {m[5], m[7], m[8]}, {m[6], m[8], m[9]}};
        double U = m[0] * phi + m[1] * gr[0] + m[2] * gr[1] + m[3] * gr[2];
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) U += Q[p][q] * Hm[p][q];
        // self energy -f a/sqrt(pi) [c^2 + 2a^2/3 d.d + 8a^4/5 q:q]
        const double al = a.aewald, a2 = al * al;
        const double dd = m[1] * m[1] + m[2] * m[2] + m[3] * m[3];
        const double qq = m[4] * m[4] + m[7] * m[7] + m[9] * m[9] + 2.0 * (m[5] * m[5] + m[6] * m[6] + m[8] * m[8]);
        e = 0.5 * a.f * U - a.f * al / 1.77245385090551602730 * (m[0] * m[0] + 2.0 * a2 / 3.0 * dd + 1.6 * a2 * a2 * qq);
        if (a.vir && a.tot) vir_cterm(vir, m + 1, Q, gr, Hm, a.f);
        a.cphi[s] = phi;
        for (int q = 0; q < 3; q++) a.cphi[(1 + q) * a.npad + s] = gr[q];
        for (int q = 0; q < 6; q++) a.cphi[(4 + q) * a.npad + s] = H[q];
        if (a.polar) {
            const double self = 4.0 / 3.0 * a2 * al / 1.77245385090551602730;
            for (int q = 0; q < 3; q++) {
                const double v = -gr[q] + self * m[1 + q];
                a.fieldd[q * a.npad + s] += v;
                a.fieldp[q * a.npad + s] += v;
            }
        }
        if (a.grad) {
            const double T3[3][3][3] = {{{T[0], T[3], T[4]}, {T[3], T[5], T[9]}, {T[4], T[9], T[6]}},
                                        {{T[3], T[5], T[9]}, {T[5], T[1], T[7]}, {T[9], T[7], T[8]}},
                                        {{T[4], T[9], T[6]}, {T[9], T[7], T[8]}, {T[6], T[8], T[2]}}};
            const int i = a.perm[s];
            for (int x = 0; x < 3; x++) {
                double v = m[0] * gr[x] + m[1] * Hm[0][x] + m[2] * Hm[1][x] + m[3] * Hm[2][x];
                for (int p = 0; p < 3; p++)
                    for (int q = 0; q < 3; q++) v += Q[p][q] * T3[p][q][x];
                a.dem[3 * i + x] += a.f * v;
            }
            // torque d x E - 2 eps:(q H), E = -g
            double QH[3][3];
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) QH[p][q] = Q[p][0] * Hm[0][q] + Q[p][1] * Hm[1][q] + Q[p][2] * Hm[2][q];
            const double tq[3] = {-(m[2] * gr[2] - m[3] * gr[1]) - 2.0 * (QH[1][2] - QH[2][1]),
                                  -(m[3] * gr[0] - m[1] * gr[2]) - 2.0 * (QH[2][0] - QH[0][2]),
                                  -(m[1] * gr[1] - m[2] * gr[0]) - 2.0 * (QH[0][1] - QH[1][0])};
            for (int x = 0; x < 3; x++) a.te[3 * i + x] += a.f * tq[x];
            // (f/2) M.phi and the self energy in the (total) charge of i
            if (a.pot) a.pot[i] += a.f * phi - 2.0 * a.f * al / 1.77245385090551602730 * m[0];
        }
    }
    if (a.aem && s < a.n) {
        // analysis: the site's reciprocal and self energy, and one self term
        // per multipole site in nem
        a.aem[a.perm[s]] += e;
        bool site = false;
        for (int c = 0; c < 10; c++) site = site || a.smp[c * a.npad + s] != 0.0;
        if (a.charge) site = site || a.charge[s] != 0.0;
        if (site) atomicAdd(&a.tot->nem, 1ull);
    }
    if (a.tot) {
        for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(&a.tot->em, e);
        if (a.vir) {
            vir.wave_sum();
            if ((threadIdx.x & 63) == 0) a.tot->add_vir(vir);
        }
    }
}

struct GatherFieldArgs {
    int n, npad, nset;
    PmeGeom g;
    const double *sx, *sy, *sz;
    const void *grid;   // double, or float in the mixed form
    size_t gstride;
    const double *dip;  // nset x 3 x npad dipoles (for the self term)
    double self;        // 4/3 a^3 / sqrt(pi)
    double *field;      // nset x 3 x npad (+=; read only with the PCG step)
    const double *skip; // nonzero: return at once (or null)
    PcgVec pcg;         // EPI: the PCG step E1 on each finished site (pcg_steps.h)
};

// reciprocal + self field of the induced-dipole sets at every site, added
// to field; with EPI (two sets, dip = z) the sum, T z at the site, goes to
// the PCG step E1 instead: lane c < 6 of the atom's group takes component c,
// its loads issued before the stencil sums
template <int ORD, int GL, class T, class GT, bool EPI>
__global__ __launch_bounds__(256) void k_pme_gather_field(GatherFieldArgs a)
{
    if (a.skip && *a.skip != 0.0) return;
    GATHER_SETUP(1, T)
    if (EPI) {
        static_assert(!EPI || GL >= 6, "one lane per component");
        const bool mine = valid && l < 6;
        PcgE1In in{0.0, 0.0, 0.0, 0.0};
        double fr = 0.0;
        if (mine) {
            in = pcg_e1_load(a.pcg, s, l);
            fr = a.field[(size_t)l * a.npad + s];
        }
        double beta[2];
        pcg_e1_begin(a.pcg, beta);
        double fp0[4], fp1[4];
        gather_share<ORD, G, 4, T, GT>(a.g, (const GT *)a.grid, tx, ty, ty1, tz, i0, ky, kz, on, fp0);
        gather_share<ORD, G, 4, T, GT>(a.g, (const GT *)a.grid + a.gstride, tx, ty, ty1, tz, i0, ky, kz, on, fp1);
        // the Cartesian gradient is linear in the fractional one, so each
        // lane maps its share first and the group sums the six components
        // scattered: lane c < 6 ends with component c (set c / 3, axis c % 3)
        double gc[6];
        frac_grad<G>(a.g, fp0, gc);
        frac_grad<G>(a.g, fp1, gc + 3);
        group_reduce_scatter<G, 6>(gc, l);
        double pv0 = 0.0, pv1 = 0.0;
        if (mine) {
            const double f = fr + (-gc[0] + a.self * in.z);
            (l < 3 ? pv0 : pv1) = pcg_e1_apply(a.pcg, s, l, f, in, beta);
        }
        pcg_e1_end(a.pcg, pv0, pv1);
        return;
    }
    for (int set = 0; set < a.nset; set++) {
        double fp[4];
        gather_share<ORD, G, 4, T, GT>(a.g, (const GT *)a.grid + set * a.gstride, tx, ty, ty1, tz, i0, ky, kz, on, fp);
        double gc[3];
        frac_grad<G>(a.g, fp, gc);
        group_reduce_scatter<G, 3>(gc, l);
        if (valid && l < 3) {
            const size_t o = (size_t)set * 3 * a.npad + (size_t)l * a.npad + s;
            a.field[o] += -gc[0] + a.self * a.dip[o];
        }
    }
}

struct PolGradPostArgs {
    int n, npad;
    PmeGeom g;
    const double *smp, *charge;
    const double *cphi;          // permanent potential (phi, g, H)
    const double *ud, *up;       // sorted dipoles
    const int *perm;
    const double *frac;          // n x 2 x 20: potentials of mu_d, mu_p
    double f, self;
    double *dep, *tep;           // caller order
    double *potp;                // charge flux: += (f/2)(phi_d + phi_p) per atom (caller order), or null
    int vir;                     // the strain terms of the cross energies into tot
    Totals *tot;
};

// reciprocal-space polarization gradient (one thread per sorted atom):
//   (f/2) [(mu_d+mu_p).H_M + mu_d.H_p + mu_p.H_d] on the induced dipoles,
//   (f/2) gradient and torque of the permanent multipoles in phi_d + phi_p,
//   and the self-term torque d x (f/2) self (mu_d + mu_p)
__device__ __forceinline__ void polgrad_site(const PolGradPostArgs &a, int s, Vir6 &vir);

__global__ __launch_bounds__(256) void k_pme_polgrad_post(PolGradPostArgs a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    Vir6 vir;
    if (s < a.n) polgrad_site(a, s, vir);
    if (a.vir) {
        vir.wave_sum();
        if ((threadIdx.x & 63) == 0) a.tot->add_vir(vir);
    }
}

// one site of k_pme_polgrad_post; with a.vir also the strain terms of the
// reciprocal cross energies (1/2)[E(mu_d + mu_p, M) + E(mu_d, mu_p)]:
// (f/2) [C(mu in phi_M) + C(M in phi_mu) + C(mu_d in phi_p) + C(mu_p in phi_d)]
__device__ __forceinline__ void polgrad_site(const PolGradPostArgs &a, int s, Vir6 &vir)
{
    // the transform is linear: phi_d + phi_p in one pass, H_d on its own
    double fs[20], fd[10];
    for (int j = 0; j < 20; j++) fs[j] = a.frac[(size_t)s * 40 + j] + a.frac[(size_t)s * 40 + 20 + j];
    for (int j = 0; j < 10; j++) fd[j] = a.frac[(size_t)s * 40 + j];
    double phs, g[3], Hs[6], T[10], phd, gd[3], Hd[6];
    frac_to_cart<20>(a.g, fs, &phs, g, Hs, T);
    frac_to_cart<10>(a.g, fd, &phd, gd, Hd, nullptr);
    const int np = a.npad;
    const double ud[3] = {a.ud[s], a.ud[np + s], a.ud[2 * np + s]};
    const double up[3] = {a.up[s], a.up[np + s], a.up[2 * np + s]};
    double HM[6];
    for (int q = 0; q < 6; q++) HM[q] = a.cphi[(4 + q) * np + s];
    auto full = [](const double *h, double M[3][3]) {
        M[0][0] = h[0]; M[1][1] = h[1]; M[2][2] = h[2];
        M[0][1] = M[1][0] = h[3]; M[0][2] = M[2][0] = h[4]; M[1][2] = M[2][1] = h[5];
    };
    double Hm[3][3], Hdm[3][3], H[3][3], Hpm[3][3];
    full(HM, Hm);
    full(Hd, Hdm);
    full(Hs, H);
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) Hpm[p][q] = H[p][q] - Hdm[p][q];
    double m[10];
    for (int c = 0; c < 10; c++) m[c] = a.smp[c * np + s];
    if (a.charge) m[0] = a.charge[s];
    const double Q[3][3] = {{m[4], m[5], m[6]}, {m[5], m[7], m[8]}, {m[6], m[8], m[9]}};
    const double T3[3][3][3] = {{{T[0], T[3], T[4]}, {T[3], T[5], T[9]}, {T[4], T[9], T[6]}},
                                {{T[3], T[5], T[9]}, {T[5], T[1], T[7]}, {T[9], T[7], T[8]}},
                                {{T[4], T[9], T[6]}, {T[9], T[7], T[8]}, {T[6], T[8], T[2]}}};
    const int i = a.perm[s];
    const double hf = 0.5 * a.f;
    for (int x = 0; x < 3; x++) {
        double v = 0.0;
        for (int p = 0; p < 3; p++) v += (ud[p] + up[p]) * Hm[p][x] + ud[p] * Hpm[p][x] + up[p] * Hdm[p][x];
        v += m[0] * g[x] + m[1] * H[0][x] + m[2] * H[1][x] + m[3] * H[2][x];
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) v += Q[p][q] * T3[p][q][x];
        a.dep[3 * i + x] += hf * v;
    }
    double QH[3][3];
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) QH[p][q] = Q[p][0] * H[0][q] + Q[p][1] * H[1][q] + Q[p][2] * H[2][q];
    // torque in phi_d + phi_p, plus the self term d x (self (mu_d + mu_p))
    const double E[3] = {-g[0] + a.self * (ud[0] + up[0]), -g[1] + a.self * (ud[1] + up[1]),
                         -g[2] + a.self * (ud[2] + up[2])};
    const double tq[3] = {m[2] * E[2] - m[3] * E[1] - 2.0 * (QH[1][2] - QH[2][1]),
                          m[3] * E[0] - m[1] * E[2] - 2.0 * (QH[2][0] - QH[0][2]),
                          m[1] * E[1] - m[2] * E[0] - 2.0 * (QH[0][1] - QH[1][0])};
    for (int x = 0; x < 3; x++) a.tep[3 * i + x] += hf * tq[x];
    // the reciprocal cross energy (f/2) M.(phi_d + phi_p) in the charge of i
    if (a.potp) a.potp[i] += hf * phs;
    if (a.vir) {
        const double gM[3] = {a.cphi[np + s], a.cphi[2 * np + s], a.cphi[3 * np + s]};
        const double gp[3] = {g[0] - gd[0], g[1] - gd[1], g[2] - gd[2]};
        const double mu[3] = {ud[0] + up[0], ud[1] + up[1], ud[2] + up[2]};
        const double Z[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        vir_cterm(vir, mu, Z, gM, Z, hf);
        vir_cterm(vir, m + 1, Q, g, H, hf);
        vir_cterm(vir, ud, Z, gp, Z, hf);
        vir_cterm(vir, up, Z, gd, Z, hf);
    }
}

// ---------------------------------------------------------------------------
// VIRIAL: k-space sum of a reciprocal energy, (coef) sum_m G |rho(m)|^2
// [2 h_a h_b (pi^2/a^2 + 1/h^2) - delta_ab] for the electrostatic kernel
// G = exp(-pi^2 h^2/a^2)/(pi V h^2 bsmod); for dispersion (G = -psi(b)/(V
// bsmod), b = pi |h|/a) sum_m |rho|^2 [-G delta_ab + 3 pre (pi/a)^2
// (sqrt(pi) b erfc(b) - exp(-b^2)) h_a h_b / (V bsmod)].  One thread per
// half-complex entry: interior columns stand for m and -m (weight 2), and
// every entry takes the mean of the weight at m and at -m, as the energy's
// kernel table does (they differ where a component sits at K/2 in a skewed
// cell).
struct VirKArgs {
    PmeGeom g;
    int H0;
    double a;              // Ewald coefficient of this kernel
    double V;
    const double *bsmod;   // K0, K1, K2 moduli
    const double *cg;      // nset complex half grids (interleaved re, im)
    size_t cstride;        // complex entries per set
    int nset;
    double coef[2];        // weight of each set (energy prefactor included)
    int disp;
    Totals *tot;
};

__device__ __forceinline__ void virk_weight(const VirKArgs &a, int m1, int m2, int m3, double W[6])
{
    const int K[3] = {a.g.K0, a.g.K1, a.g.K2};
    const int m[3] = {m1, m2, m3};
    double h[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int f = m[d] > K[d] / 2 ? m[d] - K[d] : m[d];
#pragma unroll
        for (int j = 0; j < 3; j++) h[j] += f * a.g.A[d][j] / K[d];
    }
    const double h2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
    const double bs = a.bsmod[m1] * a.bsmod[K[0] + m2] * a.bsmod[K[0] + K[1] + m3];
    double diag, off;
    const double pi = 3.14159265358979323846, sqrtpi = 1.77245385090551602730;
    if (!a.disp) {
        if (h2 == 0.0) {
            for (int q = 0; q < 6; q++) W[q] = 0.0;
            return;
        }
        const double G = exp(-pi * pi * h2 / (a.a * a.a)) / (pi * a.V * h2 * bs);
        diag = -G;
        off = G * 2.0 * (pi * pi / (a.a * a.a) + 1.0 / h2);
    } else {
        const double pre = 2.0 * pi * sqrtpi * a.a * a.a * a.a / 3.0;
        const double b = pi * sqrt(h2) / a.a;
        const double G = -pre * (sqrtpi * b * b * b * erfc(b) + (0.5 - b * b) * exp(-b * b)) / (a.V * bs);
        diag = -G;
        off = 3.0 * pre * pi * pi / (a.a * a.a) * (sqrtpi * b * erfc(b) - exp(-b * b)) / (a.V * bs);
    }
    W[0] = off * h[0] * h[0] + diag; W[1] = off * h[0] * h[1]; W[2] = off * h[0] * h[2];
    W[3] = off * h[1] * h[1] + diag; W[4] = off * h[1] * h[2]; W[5] = off * h[2] * h[2] + diag;
}

__global__ __launch_bounds__(256) void k_pme_vir_ksum(VirKArgs a)
{
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    Vir6 v; // accumulated with the sign of a strain derivative (added to the virial as is)
    if (id < a.cstride) {
        const int m1 = (int)(id % a.H0), m2 = (int)((id / a.H0) % a.g.K1), m3 = (int)(id / ((size_t)a.H0 * a.g.K1));
        double rho2 = 0.0;
        for (int set = 0; set < a.nset; set++) {
            const double re = a.cg[2 * (set * a.cstride + id)], im = a.cg[2 * (set * a.cstride + id) + 1];
            rho2 += a.coef[set] * (re * re + im * im);
        }
        double Wp[6], Wm[6];
        virk_weight(a, m1, m2, m3, Wp);
        virk_weight(a, (a.g.K0 - m1) % a.g.K0, (a.g.K1 - m2) % a.g.K1, (a.g.K2 - m3) % a.g.K2, Wm);
        const double wt = (m1 == 0 || 2 * m1 == a.g.K0) ? 1.0 : 2.0;
        const double s = 0.5 * wt * rho2;
        v.xx = s * (Wp[0] + Wm[0]); v.xy = s * (Wp[1] + Wm[1]); v.xz = s * (Wp[2] + Wm[2]);
        v.yy = s * (Wp[3] + Wm[3]); v.yz = s * (Wp[4] + Wm[4]); v.zz = s * (Wp[5] + Wm[5]);
    }
    v.wave_sum();
    __shared__ double red[4][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = v.xx; red[w][1] = v.xy; red[w][2] = v.xz; red[w][3] = v.yy; red[w][4] = v.yz; red[w][5] = v.zz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Vir6 t;
        t.xx = red[0][0] + red[1][0] + red[2][0] + red[3][0];
        t.xy = red[0][1] + red[1][1] + red[2][1] + red[3][1];
        t.xz = red[0][2] + red[1][2] + red[2][2] + red[3][2];
        t.yy = red[0][3] + red[1][3] + red[2][3] + red[3][3];
        t.yz = red[0][4] + red[1][4] + red[2][4] + red[3][4];
        t.zz = red[0][5] + red[1][5] + red[2][5] + red[3][5];
        a.tot->add_vir(t);
    }
}

// combined source of the polarization virial: the permanent multipoles
// (total charges with charge penetration) plus both induced-dipole sets
__global__ void k_pme_vir_src(int npad, const double *smp, const double *charge, const double *uind, double *out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    for (int c = 0; c < 10; c++) {
        double v = smp[(size_t)c * npad + s];
        if (c == 0 && charge) v = charge[s];
        if (c >= 1 && c <= 3) v += uind[(size_t)(c - 1) * npad + s] + uind[(size_t)(c + 2) * npad + s];
        out[(size_t)c * npad + s] = v;
    }
}

// ---------------------------------------------------------------------------
// launchers
static int vir_ksum(pmq_ctx *c, int nset, double coef0, double coef1, bool disp, double alpha);

static int fft(pmq_ctx *c, bool forward, int nset, bool fp32)
{
    rocfft_plan pl = fp32 ? (rocfft_plan)(forward ? (nset == 1 ? c->plan_f1s : c->plan_f2s)
                                                  : (nset == 1 ? c->plan_b1s : c->plan_b2s))
                          : (rocfft_plan)(forward ? (nset == 1 ? c->plan_f1 : c->plan_f2)
                                                  : (nset == 1 ? c->plan_b1 : c->plan_b2));
    void *in[1] = {forward ? (void *)c->qgrid.p : (void *)c->cgrid.p}; // read as float by the FP32 plans
    void *out[1] = {forward ? (void *)c->cgrid.p : (void *)c->qgrid.p};
    return fft_check(c, rocfft_execute(pl, in, out, (rocfft_execution_info)c->fft_info), "execute");
}

// FP32 grids and FFTs: the mixed form when its timing chose them
static bool grid32(const pmq_ctx *c) { return c->pme_fp32 && c->pme_fft32 == 1; }

// G: a kernel table of the plain form (null = the electrostatic one)
static int pme_potential(pmq_ctx *c, int nset, const double *G = nullptr, const double *skip = nullptr)
{
    int rc;
    const bool fp32 = grid32(c) && !G;
    // one timed launch of the family = one whole convolution, either engine
    ProfScope ps_(c, nset == 2 ? "pme_fft" : "pme_fft1");
    if ((fp32 ? c->pme_own32 : c->pme_own64) == 1) {
        return fft_own_convolve(c, nset, fp32 ? (const void *)c->pme_g_s.p : (const void *)(G ? G : c->pme_g.p), fp32,
                                skip);
    }
    if ((rc = fft(c, true, nset, fp32))) return rc;
    const int H0 = c->K[0] / 2 + 1;
    const int ncplx = H0 * c->K[1] * c->K[2];
    const size_t tot = (size_t)ncplx * nset;
    const dim3 grid((unsigned)((tot + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL(k_pme_convolve<float>, grid, dim3(256), 0, c->stream, ncplx, nset, c->pme_g_s.p,
                           (float *)c->cgrid.p, skip);
    else
        hipLaunchKernelGGL(k_pme_convolve<double>, grid, dim3(256), 0, c->stream, ncplx, nset, G ? G : c->pme_g.p,
                           c->cgrid.p, skip);
    PMQ_HIP(hipGetLastError());
    return fft(c, false, nset, fp32);
}

static size_t grid_size(const pmq_ctx *c) { return (size_t)c->K[0] * c->K[1] * c->K[2]; }

// lanes per atom in the gathers: 8 (one stencil plane per lane, the weights
// formed once per plane) when the system is large enough to fill the chip
// that way, else one stencil row per lane (ORD^2 rows in 32 or 64 lanes)
constexpr int PMQ_GATHER_PLANE_MIN = 12288;
constexpr int gather_row_lanes(int ord) { return ord * ord <= 32 ? 32 : 64; }
static dim3 lanes_grid(int n, int lanes) { return dim3(xcd_grid(((size_t)n * lanes + 255) / 256)); }

#define PMQ_GATHER_LAUNCH(KERNEL, ORD, GL, ...) \
    hipLaunchKernelGGL((KERNEL<ORD, GL __VA_ARGS__>), lanes_grid(c->n, GL), dim3(256), 0, c->stream, args)
static bool gather_planes(const pmq_ctx *c)
{
    return c->gather_lanes == PMQ_GATHER_PLANES || (c->gather_lanes == PMQ_GATHER_AUTO && c->n >= PMQ_GATHER_PLANE_MIN);
}

#define PMQ_ORDER_DISPATCH(KERNEL, ...)                                                                    \
    if (gather_planes(c)) {                                                                                \
        switch (c->bsorder) {                                                                              \
        case 4: PMQ_GATHER_LAUNCH(KERNEL, 4, 8, __VA_ARGS__); break;                                       \
        case 5: PMQ_GATHER_LAUNCH(KERNEL, 5, 8, __VA_ARGS__); break;                                       \
        default: PMQ_GATHER_LAUNCH(KERNEL, 6, 8, __VA_ARGS__); break;                                      \
        }                                                                                                  \
    } else {                                                                                               \
        switch (c->bsorder) {                                                                              \
        case 4: PMQ_GATHER_LAUNCH(KERNEL, 4, gather_row_lanes(4), __VA_ARGS__); break;                     \
        case 5: PMQ_GATHER_LAUNCH(KERNEL, 5, gather_row_lanes(5), __VA_ARGS__); break;                     \
        default: PMQ_GATHER_LAUNCH(KERNEL, 6, gather_row_lanes(6), __VA_ARGS__); break;                    \
        }                                                                                                  \
    }

// the sorted-atom range of every brick comes with this evaluation's atom
// sort (k_gather, when the sort followed the spreading bricks)
static int pme_brick_ranges(pmq_ctx *c)
{
    if (!c->pme_cells) {
        c->err = "PME: the atoms were not sorted by spreading brick";
        return PMQ_ERR_STATE;
    }
    return PMQ_OK;
}

// FP32 grids for the electrostatic modes of the mixed form; dispersion
// (mode 2) stays FP64
static BrickArgs brick_args(pmq_ctx *c, const double *src, const double *charge, int nset, const double *skip)
{
    BrickArgs args;
    args.n = c->n; args.nset = nset;
    for (int a = 0; a < 3; a++) args.nb[a] = c->pme_nb[a];
    args.g = geom(c);
    args.bstart = c->pme_bstart.p; args.bend = c->pme_bend.p;
    args.npad = c->npad;
    args.sx = c->sx.p; args.sy = c->sy.p; args.sz = c->sz.p;
    args.src = src;
    args.charge = charge;
    args.skip = skip;
    args.grid = c->qgrid.p;
    args.gstride = grid_size(c);
    return args;
}

static int spread(pmq_ctx *c, int mode, const double *src, const double *charge, int nset,
                  const double *skip = nullptr)
{
    const bool fp32 = c->pme_fp32 && mode != 2, g32 = grid32(c) && mode != 2;
    const BrickArgs args = brick_args(c, src, charge, nset, skip);
    if (mode == 1 && nset != 2) {
        c->err = "PME: dipole spreading takes two sets";
        return PMQ_ERR_ARG;
    }
    const dim3 grid(xcd_grid((size_t)c->pme_nb[0] * c->pme_nb[1] * c->pme_nb[2]));
    ProfScope ps_(c, mode == 1 ? "pme_spread" : "pme_spread_mpole");
#define PMQ_SPREAD(ORD, MODE)                                                                                  \
    do {                                                                                                       \
        if (g32) hipLaunchKernelGGL((k_pme_spread_brick<ORD, MODE, float, float>), grid, dim3(256), 0, c->stream, args); \
        else if (fp32) hipLaunchKernelGGL((k_pme_spread_brick<ORD, MODE, float, double>), grid, dim3(256), 0, c->stream, args); \
        else hipLaunchKernelGGL((k_pme_spread_brick<ORD, MODE, double, double>), grid, dim3(256), 0, c->stream, args); \
    } while (0)
    switch (c->bsorder * 3 + mode) {
    case 12: PMQ_SPREAD(4, 0); break;
    case 13: PMQ_SPREAD(4, 1); break;
    case 14: PMQ_SPREAD(4, 2); break;
    case 15: PMQ_SPREAD(5, 0); break;
    case 16: PMQ_SPREAD(5, 1); break;
    case 17: PMQ_SPREAD(5, 2); break;
    case 18: PMQ_SPREAD(6, 0); break;
    case 19: PMQ_SPREAD(6, 1); break;
    default: PMQ_SPREAD(6, 2); break;
    }
#undef PMQ_SPREAD
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// fractional potential derivatives of nset grids into pme_frac
static int gather_frac(pmq_ctx *c, int nset, const void *grid = nullptr)
{
    GatherFracArgs args;
    args.n = c->n; args.nset = nset; args.g = geom(c);
    args.sx = c->sx.p; args.sy = c->sy.p; args.sz = c->sz.p;
    args.grid = grid ? grid : c->qgrid.p;
    args.gstride = grid_size(c);
    args.frac = c->pme_frac.p;
    ProfScope ps_(c, "pme_gather");
    if (grid32(c)) { PMQ_ORDER_DISPATCH(k_pme_gather_frac, , float, float) }
    else if (c->pme_fp32) { PMQ_ORDER_DISPATCH(k_pme_gather_frac, , float, double) }
    else { PMQ_ORDER_DISPATCH(k_pme_gather_frac, , double, double) }
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// the mixed form's FFT precision: rocFFT's FP32 transforms are not faster
// than its FP64 ones at every grid, so both are timed once per grid (three
// forward + inverse pairs of the two-set transform each) and the faster
// kept (FP32 wins ties)
// The transforms of the grid, timed once per grid on the evaluation's own
// stream (two-set convolutions, the shape of every PCG iteration; two
// rounds, the best of each): for FP64 grids the direct LDS transforms
// against rocFFT; in the mixed form also FP32 grids (the faster of the two
// engines), kept when they beat FP64.
static int time_convolution(pmq_ctx *c, float *ms)
{
    hipEvent_t e[2];
    for (auto &x : e) PMQ_HIP(hipEventCreate(&x));
    int rc = PMQ_OK;
    for (int it = 0; it < 4 && !rc; it++) {
        if (it == 1) PMQ_HIP(hipEventRecord(e[0], c->stream));
        rc = pme_potential(c, 2);
    }
    if (!rc) {
        PMQ_HIP(hipEventRecord(e[1], c->stream));
        PMQ_HIP(hipEventSynchronize(e[1]));
        PMQ_HIP(hipEventElapsedTime(ms, e[0], e[1]));
    }
    for (auto &x : e) (void)hipEventDestroy(x);
    return rc;
}

static int choose_fft(pmq_ctx *c)
{
    if (c->pme_own64 >= 0 && (!c->pme_fp32 || c->pme_fft32 >= 0)) return PMQ_OK;
    const bool fp32_form = c->pme_fp32;
    const bool lib = c->fft_engine != PMQ_FFT_DIRECT, own = c->fft_engine != PMQ_FFT_LIBRARY;
    // candidates: (FP32 grids?, direct transforms?); two rounds, best of each
    struct Cand { bool fp32; int own; float best; };
    Cand cand[4];
    int nc = 0;
    for (int p = 0; p < (fp32_form ? 2 : 1); p++) {
        const bool fits = fft_own_fits(c, p == 1);
        if (lib || !fits) cand[nc++] = Cand{p == 1, 0, 1e30f};
        if (own && fits) cand[nc++] = Cand{p == 1, 1, 1e30f};
    }
    int rc = PMQ_OK;
    if (nc > 1 || fp32_form)
        for (int round = 0; round < 2 && !rc; round++)
            for (int q = 0; q < nc && !rc; q++) {
                c->pme_fp32 = cand[q].fp32;
                if (cand[q].fp32) { c->pme_fft32 = 1; c->pme_own32 = cand[q].own; }
                else c->pme_own64 = cand[q].own;
                float ms = 0.0f;
                if (!(rc = time_convolution(c, &ms))) cand[q].best = std::min(cand[q].best, ms);
            }
    c->pme_fp32 = fp32_form;
    if (rc) return rc;
    float b64 = 1e30f, b32 = 1e30f;
    for (int q = 0; q < nc; q++) {
        if (!cand[q].fp32 && (cand[q].best < b64 || b64 == 1e30f)) { b64 = cand[q].best; c->pme_own64 = cand[q].own; }
        if (cand[q].fp32 && (cand[q].best < b32 || b32 == 1e30f)) { b32 = cand[q].best; c->pme_own32 = cand[q].own; }
    }
    if (fp32_form) c->pme_fft32 = c->fft_grids == 1 ? 1 : c->fft_grids == 2 ? 0 : (b32 <= b64 ? 1 : 0);
    return PMQ_OK;
}

// fn's launches on pme_side, forked from the evaluation's stream here and
// joined by side_join, so that they run beside what the evaluation's stream
// takes in between (which must not touch what fn reads or writes); the FFT
// engine's stream follows.  Inline on the evaluation's stream when side is
// false or profiling is on (the library's HIP-event timings assume one
// stream).
template <class F> static int side_launch(pmq_ctx *c, bool side, F &&fn)
{
    c->pme_forked = false;
    side = side && !c->profiling && c->pme_side && c->pme_fork && c->pme_join;
    if (!side) return fn();
    int rc;
    hipStream_t main = c->stream;
    PMQ_HIP(hipEventRecord(c->pme_fork, main));
    PMQ_HIP(hipStreamWaitEvent(c->pme_side, c->pme_fork, 0));
    c->stream = c->pme_side;
    rc = c->fft_info ? fft_check(c, rocfft_execution_info_set_stream((rocfft_execution_info)c->fft_info, c->pme_side),
                                 "stream")
                     : PMQ_OK;
    if (!rc) rc = fn();
    c->stream = main;
    if (c->fft_info) {
        const int r2 = fft_check(c, rocfft_execution_info_set_stream((rocfft_execution_info)c->fft_info, main), "stream");
        if (!rc) rc = r2;
    }
    const hipError_t e = hipEventRecord(c->pme_join, c->pme_side);
    if (!rc && e != hipSuccess) {
        c->err = hipGetErrorString(e);
        rc = PMQ_ERR_HIP;
    }
    c->pme_forked = true;
    return rc;
}

// the evaluation's stream waits for the launches side_launch forked
static int side_join(pmq_ctx *c)
{
    if (!c->pme_forked) return PMQ_OK;
    c->pme_forked = false;
    PMQ_HIP(hipStreamWaitEvent(c->stream, c->pme_join, 0));
    return PMQ_OK;
}

// the permanent multipoles' fractional potential (spread, transform,
// gather into pme_frac), and with disp the dispersion potential after it
// (spread and transform, left in qgrid for pme_disp's gather); side: beside
// the real-space pair kernels launched between here and pme_mpole (none of
// them touches the grids or pme_frac)
int pme_mpole_fields(pmq_ctx *c, bool side, bool disp)
{
    int rc;
    c->pme_forked = false;
    if ((rc = choose_fft(c))) return rc;
    if ((rc = pme_brick_ranges(c))) return rc;
    const double *charge = c->prm.pentype != PMQ_PEN_NONE ? c->scp.p + 2 * (size_t)c->npad : nullptr;
    rc = side_launch(c, side, [&]() {
        int r;
        if (!(r = spread(c, 0, c->smp.p, charge, 1)) && !(r = pme_potential(c, 1))) r = gather_frac(c, 1);
        if (!r && disp && !(r = spread(c, 2, c->shp.p, nullptr, 1))) r = pme_potential(c, 1, c->pme_gd.p);
        return r;
    });
    c->mpole_frac_ready = rc == PMQ_OK;
    c->disp_pot_ready = rc == PMQ_OK && disp;
    return rc;
}

int pme_mpole(pmq_ctx *c, int mode, bool mpole, bool polar)
{
    int rc;
    if ((rc = side_join(c))) return rc;
    if (!c->mpole_frac_ready && (rc = pme_mpole_fields(c, false))) return rc;
    c->mpole_frac_ready = false;
    const double *charge = c->prm.pentype != PMQ_PEN_NONE ? c->scp.p + 2 * (size_t)c->npad : nullptr;
    MpolePostArgs args;
    args.n = c->n; args.npad = c->npad; args.g = geom(c);
    args.smp = c->smp.p; args.charge = charge;
    args.perm = c->perm.p;
    args.frac = c->pme_frac.p;
    args.f = c->prm.electric / c->prm.dielec;
    args.aewald = c->prm.aewald;
    args.grad = mpole && (mode == PMQ_GRADIENT || mode == PMQ_VIRIAL);
    args.polar = polar;
    args.dem = c->dem.p; args.te = c->te.p;
    args.pot = args.grad && c->prm.pentype != PMQ_PEN_NONE && c->prm.use_chgflx ? c->pot.p : nullptr;
    args.cphi = c->cphi.p; args.fieldd = c->pfield.p; args.fieldp = c->pfield.p + 3 * (size_t)c->npad;
    args.tot = mpole ? c->totals.p : nullptr;
    args.vir = mode == PMQ_VIRIAL ? 1 : 0;
    args.aem = mpole && mode == PMQ_ANALYSIS ? c->aem.p : nullptr;
    {
        ProfScope ps_(c, "pme_gather");
        hipLaunchKernelGGL(k_pme_mpole_post, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, args);
    }
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// reciprocal + self field of the two spread dipole sets, added to field
// (or with pcg, the PCG step E1 on the total)
static int dipole_gather(pmq_ctx *c, const double *dip, double *field, const double *skip, PcgVec *pcg)
{
    int rc;
    const size_t gs = grid_size(c);
    if ((rc = pme_potential(c, 2, nullptr, skip))) return rc;
    GatherFieldArgs args;
    args.n = c->n; args.npad = c->npad; args.nset = 2; args.g = geom(c);
    args.sx = c->sx.p; args.sy = c->sy.p; args.sz = c->sz.p;
    args.grid = c->qgrid.p;
    args.gstride = gs;
    args.dip = dip;
    args.self = 4.0 / 3.0 * c->prm.aewald * c->prm.aewald * c->prm.aewald / 1.77245385090551602730;
    args.field = field;
    args.skip = skip;
    if (pcg) {
        const unsigned nb = lanes_grid(c->n, gather_planes(c) ? 8 : gather_row_lanes(c->bsorder)).x;
        if ((size_t)nb * 2 > c->pcg_part1_cap) {
            c->err = "PCG: partial-sum buffer too small for the field gather";
            return PMQ_ERR_STATE;
        }
        pcg->n1 = (int)nb;
    }
    args.pcg = pcg ? *pcg : PcgVec{};
    {
        ProfScope ps_(c, "pme_gather_field");
        if (pcg) {
            if (grid32(c)) { PMQ_ORDER_DISPATCH(k_pme_gather_field, , float, float, true) }
            else if (c->pme_fp32) { PMQ_ORDER_DISPATCH(k_pme_gather_field, , float, double, true) }
            else { PMQ_ORDER_DISPATCH(k_pme_gather_field, , double, double, true) }
        } else {
            if (grid32(c)) { PMQ_ORDER_DISPATCH(k_pme_gather_field, , float, float, false) }
            else if (c->pme_fp32) { PMQ_ORDER_DISPATCH(k_pme_gather_field, , float, double, false) }
            else { PMQ_ORDER_DISPATCH(k_pme_gather_field, , double, double, false) }
        }
    }
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int pme_dipole_field(pmq_ctx *c, const double *dip, double *field, const double *skip, PcgVec *pcg)
{
    // dip and field: two sets of 3 x npad, contiguous
    int rc;
    if ((rc = spread(c, 1, dip, nullptr, 2, skip))) return rc;
    return dipole_gather(c, dip, field, skip, pcg);
}

// The mutual field's two independent first steps in one launch: workgroups
// [0, nspread) spread the two induced-dipole sets onto the grids (bricks in
// the standalone launch's XCD-aware order; nspread is a multiple of 8, so
// the rest keep their own order on the XCDs), the rest form the pair-matrix
// product T p.  Both read p only and write disjoint outputs; together they
// fill the chip where each alone is bound by its chain of dependent loads.
template <int ORD, class T, class GT, int BK, class V>
__global__ __launch_bounds__(256) void k_spread_spmv(BrickArgs sa, SpmvArgs ma, int nspread)
{
    // the first launch of a PCG iteration: block 0 tests the last one
    if (ma.check && ma.pcg.iter >= 2 && blockIdx.x == 0) pcg_converged(ma.pcg, ma.pcg.iter);
    if ((int)blockIdx.x < nspread) spread_brick<ORD, 1, T, GT>(sa, blockIdx.x, nspread);
    else spmv_rows<SPMV_G, BK, V>(ma, (int)blockIdx.x - nspread);
}

// field = T dip (real space, the pair matrix) + the reciprocal and self
// field of dip; skip: the solve's converged flag (or null); pcg: a PCG
// iteration (dip = z): the first launch runs the convergence test first,
// the gather the step E1 on the total instead of storing it; or null
int pme_mutual_field(pmq_ctx *c, const double *dip, double *field, const double *skip, PcgVec *pcg)
{
    const bool fp32 = c->pme_fp32, g32 = grid32(c);
    const int bk = spmv_box(c);
    if (fp32 != c->tmixed || bk == ROW_BOX_NONE) {
        // forms the fused launch is not built for: the two launches
        int rc;
        if ((rc = pair_matrix_apply(c, field, skip, pcg, pcg ? PCG_APPLY_CHECK : PCG_APPLY_NONE))) return rc;
        return pme_dipole_field(c, dip, field, skip, pcg);
    }
    const BrickArgs sa = brick_args(c, dip, nullptr, 2, skip);
    const SpmvArgs ma = spmv_args(c, field, skip, pcg, pcg ? 1 : 0);
    const int nspread = (int)xcd_grid((size_t)c->pme_nb[0] * c->pme_nb[1] * c->pme_nb[2]);
    const int nspmv = (int)(((size_t)c->npad * SPMV_G + 255) / 256);
    const dim3 grid(nspread + nspmv);
    {
        ProfScope ps_(c, "pol_spmv_spread");
#define PMQ_SS(ORD, T, GT, BK, V) hipLaunchKernelGGL((k_spread_spmv<ORD, T, GT, BK, V>), grid, dim3(256), 0, c->stream, sa, ma, nspread)
#define PMQ_SS_BOX(ORD, T, GT, V)                                                                              \
        do {                                                                                                   \
            if (bk == ROW_BOX_ORTHO) PMQ_SS(ORD, T, GT, ROW_BOX_ORTHO, V);                                   \
            else PMQ_SS(ORD, T, GT, ROW_BOX_ANY, V);                                                          \
        } while (0)
#define PMQ_SS_PREC(ORD)                                                                                       \
        do {                                                                                                   \
            if (g32) PMQ_SS_BOX(ORD, float, float, float);                                                    \
            else if (fp32) PMQ_SS_BOX(ORD, float, double, float);                                             \
            else PMQ_SS_BOX(ORD, double, double, double);                                                     \
        } while (0)
        switch (c->bsorder) {
        case 4: PMQ_SS_PREC(4); break;
        case 5: PMQ_SS_PREC(5); break;
        default: PMQ_SS_PREC(6); break;
        }
#undef PMQ_SS_PREC
#undef PMQ_SS_BOX
#undef PMQ_SS
    }
    PMQ_HIP(hipGetLastError());
    return dipole_gather(c, dip, field, skip, pcg);
}

// dispersion: energy (1/2) C phi + a^6/12 C^2 and gradient C grad(phi) of
// every atom (one lane group per atom); the gradient is added in caller
// order (each atom written once), the energy summed per wave into edsp
struct DispPostArgs {
    int n, npad;
    PmeGeom g;
    const double *sx, *sy, *sz, *cs; // sorted positions, C6 coefficients
    const int *perm;
    const double *grid;
    double self;                     // a^6 / 12
    int grad;
    double *de;                      // caller order (+=)
    Totals *tot;
};

template <int ORD, int GL>
__global__ __launch_bounds__(256) void k_pme_disp_gather(DispPostArgs a)
{
    GATHER_SETUP(1, double)
    double fp[4];
    gather_share<ORD, G, 4, double, double>(a.g, a.grid, tx, ty, ty1, tz, i0, ky, kz, on, fp);
    group_reduce<G, 4>(fp);
    double e = 0.0;
    if (valid && l == 0) {
        double phi, gr[3];
        frac_to_cart<4>(a.g, fp, &phi, gr, nullptr, nullptr);
        const double C = a.cs[s];
        e = 0.5 * C * phi + a.self * C * C;
        if (a.grad) {
            const int i = a.perm[s];
            for (int q = 0; q < 3; q++) a.de[3 * i + q] += C * gr[q];
        }
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&a.tot->edsp, e);
}

// need_bin: the atoms are not yet binned by brick in this evaluation (no
// multipole PME ran before)
int pme_disp(pmq_ctx *c, bool grad, bool need_bin, bool vir)
{
    int rc;
    if ((rc = choose_fft(c))) return rc;
    if (need_bin && (rc = pme_brick_ranges(c))) return rc;
    if (c->disp_pot_ready) {
        // computed after the multipoles' potential (pme_mpole_fields)
        if ((rc = side_join(c))) return rc;
        c->disp_pot_ready = false;
    } else {
        if ((rc = spread(c, 2, c->shp.p, nullptr, 1))) return rc;
        if ((rc = pme_potential(c, 1, c->pme_gd.p))) return rc;
    }
    DispPostArgs args;
    args.n = c->n; args.npad = c->npad; args.g = geom(c);
    args.sx = c->sx.p; args.sy = c->sy.p; args.sz = c->sz.p; args.cs = c->shp.p;
    args.perm = c->perm.p;
    args.grid = c->qgrid.p;
    const double a2 = c->prm.dewald * c->prm.dewald;
    args.self = a2 * a2 * a2 / 12.0;
    args.grad = grad ? 1 : 0;
    args.de = c->dedsp.p;
    args.tot = c->totals.p;
    {
        ProfScope ps_(c, "pme_gather");
        PMQ_ORDER_DISPATCH(k_pme_disp_gather, )
    }
    PMQ_HIP(hipGetLastError());
    if (vir) {
        // the k-space sum of (1/2) sum_m G_d |rho|^2, from a second spreading
        // of the C6 coefficients (dispersion grids are FP64 in both forms)
        if ((rc = spread(c, 2, c->shp.p, nullptr, 1))) return rc;
        if ((rc = fft(c, true, 1, false))) return rc;
        if ((rc = vir_ksum(c, 1, 0.5, 0.0, true, c->prm.dewald))) return rc;
    }
    return PMQ_OK;
}

// the reciprocal potential grids the solve carries (k_pcg_prec): p then u,
// two sets each, in the element type of the field grids
bool pme_carry_grids(pmq_ctx *c, void **qg, void **gp, void **gu, size_t *pts, bool *g32)
{
    if (!c->prm.use_ewald || !c->qgrid.p) return false;
    const size_t gs = grid_size(c);
    if (c->phi_acc.alloc(4 * gs) != hipSuccess) return false; // sized for FP64 grids
    const bool f = grid32(c);
    *qg = c->qgrid.p;
    *gp = c->phi_acc.p;
    *gu = f ? (void *)((float *)c->phi_acc.p + 2 * gs) : (void *)(c->phi_acc.p + 2 * gs);
    *pts = 2 * gs;
    *g32 = f;
    return true;
}

// the induced dipoles' fractional potential (gather into pme_frac, from
// the grids the solve carried when it did, else spread and transform).
// Measured on the second stream beside the real-space gradient tiles
// (side = true): slower (957 -> 909 evals/s mixed on the 12,288-atom box),
// the gather being too short to pay for the fork and join
static int pme_polar_grad_fields(pmq_ctx *c, bool side)
{
    c->pme_forked = false;
    void *qg, *gp, *gu = nullptr;
    size_t pts;
    bool g32;
    if (c->phi_carried && !pme_carry_grids(c, &qg, &gp, &gu, &pts, &g32)) {
        c->err = "PME: the carried potential grids are gone";
        return PMQ_ERR_STATE;
    }
    return side_launch(c, side, [&]() {
        // with carried grids, the potential of the solve's dipoles (k_pcg_prec)
        if (gu) return gather_frac(c, 2, gu);
        int r;
        if (!(r = spread(c, 1, c->uind.p, nullptr, 2)) && !(r = pme_potential(c, 2))) r = gather_frac(c, 2);
        return r;
    });
}

int pme_polar_grad(pmq_ctx *c, bool vir)
{
    int rc;
    if (c->pme_forked) {
        if ((rc = side_join(c))) return rc;
    } else if ((rc = pme_polar_grad_fields(c, false))) {
        return rc;
    }
    PolGradPostArgs args;
    args.n = c->n; args.npad = c->npad; args.g = geom(c);
    args.smp = c->smp.p;
    args.charge = c->prm.pentype != PMQ_PEN_NONE ? c->scp.p + 2 * (size_t)c->npad : nullptr;
    args.cphi = c->cphi.p;
    args.ud = c->uind.p; args.up = c->uind.p + 3 * (size_t)c->npad;
    args.perm = c->perm.p;
    args.frac = c->pme_frac.p;
    args.f = c->prm.electric / c->prm.dielec;
    args.self = 4.0 / 3.0 * c->prm.aewald * c->prm.aewald * c->prm.aewald / 1.77245385090551602730;
    args.dep = c->dep.p;
    args.tep = c->tep.p;
    args.potp = c->prm.pentype != PMQ_PEN_NONE && c->prm.use_chgflx ? c->potp.p : nullptr;
    args.vir = vir ? 1 : 0;
    args.tot = c->totals.p;
    {
        ProfScope ps_(c, "pme_gather");
        hipLaunchKernelGGL(k_pme_polgrad_post, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, args);
    }
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

static double cell_volume(const pmq_ctx *c)
{
    const double (*A)[3] = c->pmeA;
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
                     + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    return (double)c->K[0] * c->K[1] * c->K[2] / det;
}

// the k-space virial sum of the nset transformed grids in cgrid (FP64)
static int vir_ksum(pmq_ctx *c, int nset, double coef0, double coef1, bool disp, double alpha)
{
    VirKArgs k;
    k.g = geom(c);
    k.H0 = c->K[0] / 2 + 1;
    k.a = alpha;
    k.V = cell_volume(c);
    k.bsmod = c->pme_bsmod.p;
    k.cg = c->cgrid.p;
    k.cstride = (size_t)k.H0 * c->K[1] * c->K[2];
    k.nset = nset;
    k.coef[0] = coef0;
    k.coef[1] = coef1;
    k.disp = disp ? 1 : 0;
    k.tot = c->totals.p;
    hipLaunchKernelGGL(k_pme_vir_ksum, dim3((unsigned)((k.cstride + 255) / 256)), dim3(256), 0, c->stream, k);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// VIRIAL, after the solve: the k-space sums of the reciprocal multipole
// energy (f/2) S(M) and of the polarization's cross energies
// (1/2)[E(mu_d + mu_p, M) + E(mu_d, mu_p)] = (f/4)[S(M + mu_d + mu_p) - S(M)
// - S(mu_d) - S(mu_p)], S(X) the k-space sum of X's grid.  Grids and
// transforms in FP64 whatever the precision form (the virial is an FP64
// quantity of both forms); the atoms are binned by brick already.
int pme_virial(pmq_ctx *c, bool mpole, bool polar)
{
    int rc;
    const bool keep32 = c->pme_fp32;
    c->pme_fp32 = false;
    const double f = c->prm.electric / c->prm.dielec, a = c->prm.aewald;
    const double *charge = c->prm.pentype != PMQ_PEN_NONE ? c->scp.p + 2 * (size_t)c->npad : nullptr;
    const double cm = (mpole ? 1.0 : 0.0) - (polar ? 0.5 : 0.0);
    rc = PMQ_OK;
    if (cm != 0.0 && !rc) {
        if (!(rc = spread(c, 0, c->smp.p, charge, 1)) && !(rc = fft(c, true, 1, false)))
            rc = vir_ksum(c, 1, 0.5 * f * cm, 0.0, false, a);
    }
    if (polar && !rc) {
        const int np = c->npad;
        if (c->pme_vsrc.alloc(10 * (size_t)np) != hipSuccess) {
            c->err = "PME virial: out of device memory";
            rc = PMQ_ERR_HIP;
        } else {
            hipLaunchKernelGGL(k_pme_vir_src, dim3((np + 255) / 256), dim3(256), 0, c->stream, np, c->smp.p, charge,
                               c->uind.p, c->pme_vsrc.p);
            if (!(rc = spread(c, 0, c->pme_vsrc.p, nullptr, 1)) && !(rc = fft(c, true, 1, false)))
                rc = vir_ksum(c, 1, 0.25 * f, 0.0, false, a);
            if (!rc && !(rc = spread(c, 1, c->uind.p, nullptr, 2)) && !(rc = fft(c, true, 2, false)))
                rc = vir_ksum(c, 2, -0.25 * f, -0.25 * f, false, a);
        }
    }
    c->pme_fp32 = keep32;
    return rc;
}

} // namespace pmq
