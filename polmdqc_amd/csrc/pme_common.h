// pme_common.h -- cardinal B-spline weights of one atom along one axis.
//
// For fractional offset w in [0,1) the order-n stencil has n nonzero
// weights theta_k = M_n(w + n - 1 - k), k = 0..n-1.  They are built by the
// recursion M_{m+1}(x) = [x M_m(x) + (m+1-x) M_m(x-1)] / m, and the d-th
// derivative is the d-th backward difference of the order n-d weights
// (M_n' (x) = M_{n-1}(x) - M_{n-1}(x-1)).
#pragma once

#include <hip/hip_runtime.h>

#define PMQ_MAX_ORDER 6

namespace pmq {

// XCD-aware order of a launch of gridDim.x = 8 * per workgroups: the chip
// deals workgroups to its 8 XCDs round robin (w % 8), so workgroup w takes
// item (w % 8) * per + w / 8 of the launch's items: each XCD, with its own
// L2, works on one contiguous eighth (spatially compact in the sorted order)
constexpr int PMQ_NXCD = 8;
__device__ __forceinline__ int xcd_item(int w, int per) { return (w % PMQ_NXCD) * per + w / PMQ_NXCD; }
__host__ __device__ __forceinline__ unsigned xcd_grid(size_t items)
{
    return (unsigned)(((items + PMQ_NXCD - 1) / PMQ_NXCD) * PMQ_NXCD);
}

// grid coordinate u = A p of one axis (A row = K times a reciprocal box
// row), with explicit fused multiply-adds so every kernel that bins or
// stencils an atom gets the same bits
__device__ __forceinline__ double grid_coord(const double *A, double px, double py, double pz)
{
    return fma(A[0], px, fma(A[1], py, A[2] * pz));
}

// spreading bricks: axis a is cut at brick_lo(b) = floor(b K / nb)
__device__ __host__ __forceinline__ int brick_lo(int b, int K, int nb) { return (int)(((long long)b * K) / nb); }

// Morton interleave of three 10-bit cell or brick indices
__device__ __forceinline__ unsigned morton3(unsigned x, unsigned y, unsigned z)
{
    auto spread = [](unsigned v) {
        v &= 0x3ff;
        v = (v | (v << 16)) & 0x030000ff;
        v = (v | (v << 8)) & 0x0300f00f;
        v = (v | (v << 4)) & 0x030c30c3;
        v = (v | (v << 2)) & 0x09249249;
        return v;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

// the grid's brick layout, handed to the atom sort
struct BrickGeom {
    double A[3][3];
    int K[3], nb[3];
};

// home brick of a point: per axis the brick holding floor(u) mod K
__device__ __forceinline__ void home_brick(const double (*A)[3], const int *K, const int *nb, double px, double py,
                                           double pz, int *hb)
{
    for (int a = 0; a < 3; a++) {
        int h = (int)floor(grid_coord(A[a], px, py, pz)) % K[a];
        if (h < 0) h += K[a];
        int b = (int)(((long long)h * nb[a]) / K[a]);
        while (b + 1 < nb[a] && brick_lo(b + 1, K[a], nb[a]) <= h) b++;
        while (b > 0 && brick_lo(b, K[a], nb[a]) > h) b--;
        hb[a] = b;
    }
}

template <class R, int ND>
__device__ __forceinline__ void bspline_weights(R w, int n, R th[PMQ_MAX_ORDER][4])
{
    // b_m[k] = M_m(w + m - 1 - k), kept for orders n-3 .. n
    R b[PMQ_MAX_ORDER + 1];
    R keep[4][PMQ_MAX_ORDER + 3];
    for (int j = 0; j < PMQ_MAX_ORDER + 1; j++) b[j] = R(0);
    for (int q = 0; q < 4; q++)
        for (int j = 0; j < PMQ_MAX_ORDER + 3; j++) keep[q][j] = R(0);
    b[0] = R(1);
    for (int m = 1; m <= n; m++) {
        if (m >= n - ND) {
            // keep[n - m][k + (n - m)] = b_m[k] so differences index without bounds checks
            const int q = n - m;
            for (int k = 0; k < m; k++) keep[q][k + q] = b[k];
        }
        if (m == n) break;
        R nb[PMQ_MAX_ORDER + 1];
        const R inv = R(1) / R(m);
        for (int k = 0; k <= m; k++) {
            const R left = k > 0 ? b[k - 1] : R(0);
            const R right = k < m ? b[k] : R(0);
            nb[k] = ((w + R(m - k)) * left + (R(k + 1) - w) * right) * inv;
        }
        for (int k = 0; k <= m; k++) b[k] = nb[k];
    }
    for (int k = 0; k < n; k++) {
        // keep[q][k + q - j] = b_{n-q}[k - j]; out of range entries are zero
        th[k][0] = keep[0][k];
        if (ND >= 1) th[k][1] = keep[1][k + 1 - 1] - keep[1][k + 1];
        if (ND >= 2) th[k][2] = keep[2][k] - R(2) * keep[2][k + 1] + keep[2][k + 2];
        if (ND >= 3) th[k][3] = keep[3][k] - R(3) * keep[3][k + 1] + R(3) * keep[3][k + 2] - keep[3][k + 3];
    }
}

} // namespace pmq
