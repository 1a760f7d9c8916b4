// pcg_steps.h -- the vector steps of the induced-dipole PCG, run at the end
// of the operator's last launch and inside the preconditioner's launch.
//
// The solve (k_polar.hip polar_solve) applies the operator A = 1/alpha - T
// to the preconditioned residual z rather than to the search direction p,
// and carries A p and z by linearity:
//   E1, where (T z)_t is complete (the PME field gather, or the row product
//       without PME):
//         p = z + beta p,   v = A p = A z + beta v,   partials of p.v;
//         the last block:   step = r.z / p.v
//   E2, the preconditioner's launch (P r = udiag a r + max(a, 1e-8) T_s a r,
//       linear in r; the diagonal preconditioner P r = a r):
//         z -= step P v,   r -= step v,   u += step p,   partials of r.z, r.r;
//         the last block:   beta = new r.z / r.z and the convergence test
// This is the textbook iteration (p = z + beta p; step = r.z / p.A p;
// u += step p; r -= step A p; z = P r) with A p and P r formed from A z and
// P v, so that no launch reads a value that another block of the same
// launch updates: one iteration is the operator's launches plus one.
//
// Reductions: a launch stores one set of partial sums per block; the
// launches that need the totals add them again in every block (the same
// fixed order everywhere, so every block gets the same bits):
//   the launch ending in E1: r.z of E2 (k-1) for beta
//   E2: p.v of E1 for the step length
// The convergence test of iteration k - 1 (r.r of E2 (k-1)) is block 0's
// of the first launch of iteration k, whose other blocks carry on: when it
// passes, the rest of iteration k and all later launches return at once
// (one first launch of the batch's spare iteration is the price of not
// testing in every block), and after each batch one single-block launch
// tests the last iteration.  No block of a launch waits on another and
// nothing but the launch boundary orders the partials.
#pragma once

#include "pmq_internal.h"

namespace pmq {

// scalar slots of the solve (16 doubles).  r.z of iteration j sits in
// slot pcg_rz_slot(j): the launch ending in E1 writes r.z (k-1) while it
// still reads r.z (k-2).
enum {
    PCG_RZ = 0,      // r.z, two sets (even iterations)
    PCG_STEP = 2,    // step lengths, two sets (block 0 of E2 records them)
    PCG_BETA = 4,    // direction updates, two sets (block 0 of E1 records them)
    PCG_RR = 6,      // r.r, two sets
    PCG_DONE = 8,    // nonzero once converged (the launches of later iterations return at once)
    PCG_ITERS = 9,   // iterations at convergence
    PCG_EPS = 10,    // debye x rms residual of the last tested iteration
    PCG_RZ_ODD = 11, // r.z, two sets (odd iterations)
};

__host__ __device__ __forceinline__ int pcg_rz_slot(int j) { return (j & 1) ? PCG_RZ_ODD : PCG_RZ; }

struct PcgVec {
    int npad;
    int iter;          // this iteration, 1, 2, ... (0: the initial residual)
    int prec;          // the sparse preconditioner (else the diagonal one)
    const double *spol;
    double udiag;
    double *u, *r, *z, *p, *v; // 2 x 3 x npad, set-major
    double *pks;       // z packed with the positions (PMQ_PK per atom, slots 3..8): the row product's input
    double *pkw;       // a v (a r at the start) packed likewise: the preconditioner's input
    double *part1;     // E1: p.v of the two sets, 2 per block
    int n1;            // blocks of the launch ending in E1
    double *part2;     // E2: r.z, r.z, r.r, r.r, 4 per block
    int n2;            // blocks of the E2 launch
    double *sc;
    double npolar, poleps;
};

constexpr int PCG_TB = 256; // lanes per block of every launch that stores or adds partials
constexpr int PCG_RB = 8;   // blocks' partials each lane loads at a time

// Up to four sums over the block (every lane gets them): wave shuffles,
// then the waves in order.
__device__ __forceinline__ void pcg_block_sum4(double v[4])
{
    __shared__ double red[4][PCG_TB / 64];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
    __syncthreads(); // red may still be read by an earlier call
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) red[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double s = red[q][0];
#pragma unroll
        for (int k = 1; k < PCG_TB / 64; k++) s += red[q][k];
        v[q] = s;
    }
}

// totals of components q0 .. q0+NQ-1 of n blocks' partials (stride values
// per block), in every lane: lane l adds blocks l, l + PCG_TB, ... in
// order, then the block sum
template <int NQ>
__device__ __forceinline__ void pcg_totals(const double *part, int stride, int q0, int n, double out[4])
{
    double s[4] = {0, 0, 0, 0};
    for (int b0 = 0; b0 < n; b0 += PCG_TB * PCG_RB) {
        double t[PCG_RB][NQ];
#pragma unroll
        for (int k = 0; k < PCG_RB; k++) {
            const int b = b0 + k * PCG_TB + (int)threadIdx.x;
#pragma unroll
            for (int q = 0; q < NQ; q++) t[k][q] = b < n ? part[(size_t)b * stride + q0 + q] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < PCG_RB; k++)
#pragma unroll
            for (int q = 0; q < NQ; q++) s[q] += t[k][q];
    }
    pcg_block_sum4(s);
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = s[q];
}

// the convergence test of iteration k - 1 (debye sqrt(max r.r / npolar) <
// poleps), run by one block (block 0 of the first launch of iteration
// k >= 2, or the check after a batch): it records the test and sets
// PCG_DONE, on which the later launches return at once
__device__ __forceinline__ void pcg_converged(const PcgVec &a, int k)
{
    double rr[4];
    pcg_totals<2>(a.part2, 4, 2, a.n2, rr);
    const double eps = 4.80321 * sqrt(fmax(rr[0], rr[1]) / a.npolar);
    if (threadIdx.x == 0) {
        a.sc[PCG_EPS] = eps;
        a.sc[PCG_RR] = rr[0];
        a.sc[PCG_RR + 1] = rr[1];
        if (eps < a.poleps) {
            a.sc[PCG_ITERS] = (double)(k - 1);
            a.sc[PCG_DONE] = 1.0;
        }
    }
}

// start of E1: r.z of iteration k - 1 (recorded by block 0) and beta
__device__ __forceinline__ void pcg_e1_begin(const PcgVec &a, double beta[2])
{
    double rz[4];
    pcg_totals<2>(a.part2, 4, 0, a.n2, rz);
    const int sn = pcg_rz_slot(a.iter - 1), so = pcg_rz_slot(a.iter - 2);
#pragma unroll
    for (int set = 0; set < 2; set++) {
        const double old = a.iter > 1 ? a.sc[so + set] : 0.0;
        beta[set] = old != 0.0 ? rz[set] / old : 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.sc[sn] = rz[0];
        a.sc[sn + 1] = rz[1];
        a.sc[PCG_BETA] = beta[0];
        a.sc[PCG_BETA + 1] = beta[1];
    }
}

// E1 at component c (0..5) of site s, given f = (T z) there; returns the
// term of p.v.  pcg_e1_load issues its loads (so a kernel can start them
// early), pcg_e1_apply the rest.
struct PcgE1In {
    double z, pol, p, v;
};

__device__ __forceinline__ PcgE1In pcg_e1_load(const PcgVec &a, int s, int c)
{
    const size_t t = (size_t)c * a.npad + s;
    return PcgE1In{a.z[t], a.spol[s], a.p[t], a.v[t]};
}

__device__ __forceinline__ double pcg_e1_apply(const PcgVec &a, int s, int c, double f, const PcgE1In &in,
                                               const double beta[2])
{
    const size_t t = (size_t)c * a.npad + s;
    const double az = in.pol != 0.0 ? in.z / in.pol - f : 0.0;
    double pn = in.z, vn = az;
    if (a.iter > 1) {
        const double b = c < 3 ? beta[0] : beta[1];
        pn = in.z + b * in.p;
        vn = az + b * in.v;
    }
    a.p[t] = pn;
    a.v[t] = vn;
    if (a.prec) a.pkw[(size_t)PMQ_PK * s + 3 + c] = in.pol * vn;
    return pn * vn;
}

__device__ __forceinline__ double pcg_e1(const PcgVec &a, int s, int c, double f, const double beta[2])
{
    return pcg_e1_apply(a, s, c, f, pcg_e1_load(a, s, c), beta);
}

// end of E1: the block's p.v partials (every lane passes its sums)
__device__ __forceinline__ void pcg_e1_end(const PcgVec &a, double pv0, double pv1)
{
    double v[4] = {pv0, pv1, 0.0, 0.0};
    pcg_block_sum4(v);
    if (threadIdx.x == 0) {
        a.part1[(size_t)blockIdx.x * 2] = v[0];
        a.part1[(size_t)blockIdx.x * 2 + 1] = v[1];
    }
}

// start of E2 (iter >= 1): the step lengths r.z / p.v (block 0 records them)
__device__ __forceinline__ void pcg_e2_begin(const PcgVec &a, double step[2])
{
    double pv[4];
    pcg_totals<2>(a.part1, 2, 0, a.n1, pv);
    const int so = pcg_rz_slot(a.iter - 1);
#pragma unroll
    for (int set = 0; set < 2; set++) step[set] = pv[set] != 0.0 ? a.sc[so + set] / pv[set] : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.sc[PCG_STEP] = step[0];
        a.sc[PCG_STEP + 1] = step[1];
    }
}

// E2 at component c of site s, given q = (P v) there (iter >= 1), or
// q = (P r) for the initial residual (iter 0: z = q); rz and rr receive the
// site's terms of the new r.z and r.r
__device__ __forceinline__ void pcg_e2(const PcgVec &a, int s, int c, double q, const double step[2], double &rz,
                                       double &rr)
{
    const size_t t = (size_t)c * a.npad + s;
    double zn, rn;
    if (a.iter == 0) {
        zn = q;
        rn = a.r[t];
    } else {
        const double st = c < 3 ? step[0] : step[1];
        zn = a.z[t] - st * q;
        rn = a.r[t] - st * a.v[t];
        a.r[t] = rn;
        a.u[t] += st * a.p[t];
    }
    a.z[t] = zn;
    a.pks[(size_t)PMQ_PK * s + 3 + c] = zn;
    rz = rn * zn;
    rr = rn * rn;
}

// end of E2: the block's r.z and r.r partials
__device__ __forceinline__ void pcg_e2_end(const PcgVec &a, double rz0, double rz1, double rr0, double rr1)
{
    double v[4] = {rz0, rz1, rr0, rr1};
    pcg_block_sum4(v);
    if (threadIdx.x < 4) a.part2[(size_t)blockIdx.x * 4 + threadIdx.x] = v[threadIdx.x];
}

} // namespace pmq
