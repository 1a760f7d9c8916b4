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

// Position of cell (x, y, z) along the 3-D Hilbert curve through a cube of
// 2^bits cells a side (bits <= 10), by Skilling's transpose method (J.
// Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 381
// (2004)): undo the curve's rotations and reflections level by level, Gray
// code the result, then interleave the three words' bits, x highest.
// Cells next to each other on the curve share a face, so any run of cells
// along it stays compact, which is what the atom sort's 64-atom blocks need
// (a Morton order jumps across the box between its octants).
__device__ __host__ __forceinline__ unsigned hilbert3(unsigned x, unsigned y, unsigned z, int bits)
{
    unsigned X[3] = {x, y, z};
    for (unsigned q = 1u << (bits - 1); q > 1u; q >>= 1) {
        const unsigned p = q - 1u;
        for (int i = 0; i < 3; i++) {
            if (X[i] & q) {
                X[0] ^= p;
            } else {
                const unsigned t = (X[0] ^ X[i]) & p;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    unsigned t = 0;
    for (unsigned q = 1u << (bits - 1); q > 1u; q >>= 1)
        if (X[2] & q) t ^= q - 1u;
    unsigned h = 0;
    for (int b = bits - 1; b >= 0; b--)
        for (int i = 0; i < 3; i++) h = (h << 1) | (((X[i] ^ t) >> b) & 1u);
    return h;
}

// bits of the Hilbert cube holding n cells on the longest axis
__host__ __forceinline__ int hilbert_bits(int n)
{
    int bits = 1;
    while ((1 << bits) < n) bits++;
    return bits;
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
