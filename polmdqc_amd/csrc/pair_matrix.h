// pair_matrix.h -- the dipole-field operator T as a row-owned sparse
// matrix: one stored entry per pair and row, its product with two dipole
// sets, and the launch-independent body of that product (the standalone
// product kernel and the one fused with the PME spreading both run it).
#pragma once

#include "pcg_steps.h"
#include "pmq_internal.h"

namespace pmq {

// One stored entry: the pair tensor is T = R2 r r^T - R1 I (r the minimum
// image vector from the row's atom to the column's), so an entry keeps the
// two scalars (R1, R2), 8 (float) or 16 (double) bytes, beside its column in
// a separate int array shared by the full and the preconditioner matrices;
// the products rebuild r from the sorted positions.
template <class V>
struct alignas(2 * sizeof(V)) TEnt {
    V r1, r2;
};

// The geometry a product needs to rebuild each entry's pair vector: the
// sorted positions and the periodic box (the minimum image convention of the
// fill).
struct RowGeom {
    const double *sx, *sy, *sz;
    DevBox box;
    double len[3], inv[3]; // rectangular box: edge lengths and their inverses
};

// box forms the products are built for: none, rectangular (nearest image by
// rounding, the same image the fill's wrapping finds), any other (the
// fill's own image function)
enum { ROW_BOX_NONE = 0, ROW_BOX_ORTHO = 1, ROW_BOX_ANY = 2 };

template <int BK>
__device__ __forceinline__ void row_image(const RowGeom &geo, double &x, double &y, double &z)
{
    if constexpr (BK == ROW_BOX_ORTHO) {
        x -= geo.len[0] * rint(x * geo.inv[0]);
        y -= geo.len[1] * rint(y * geo.inv[1]);
        z -= geo.len[2] * rint(z * geo.inv[2]);
    } else if constexpr (BK == ROW_BOX_ANY) {
        image<double>(geo.box, x, y, z);
    }
}

// T u for one entry (column k, scalars R1 R2) added to acc, both dipole sets:
// T u = R2 r (r . u) - R1 u with r rebuilt from the positions
__device__ __forceinline__ void entry_apply(double r1, double r2, double xr, double yr, double zr, const double d[6],
                                            double acc[6])
{
#pragma unroll
    for (int set = 0; set < 2; set++) {
        const double ux = d[3 * set], uy = d[3 * set + 1], uz = d[3 * set + 2];
        const double ru = r2 * (xr * ux + yr * uy + zr * uz);
        acc[3 * set + 0] += ru * xr - r1 * ux;
        acc[3 * set + 1] += ru * yr - r1 * uy;
        acc[3 * set + 2] += ru * zr - r1 * uz;
    }
}

// this lane's share (entries e0 + g, e0 + g + G, ... < e1) of T u for both
// dipole sets of the row whose atom sits at (xi, yi, zi), U entries at a
// time: their columns and scalars are loaded together, then the packed
// records of those columns (position and both dipoles, 80 bytes as five
// 16-byte loads) together, so a lane waits on two loads per U entries and
// each entry touches one record rather than nine arrays.  Entries past e1
// are clamped to e1 - 1 and weighted by zero.
template <int G, int U, int BK, class V>
__device__ __forceinline__ void row_apply(const int *cols, const TEnt<V> *ent, int e0, int e1, int g,
                                          const RowGeom &geo, double xi, double yi, double zi, const double *pk,
                                          double acc[6])
{
    for (int base = e0 + g; base < e1; base += G * U) {
        int k[U];
        TEnt<V> r[U];
        double wgt[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int e = min(base + q * G, e1 - 1);
            wgt[q] = base + q * G < e1 ? 1.0 : 0.0;
            k[q] = cols[e];
            r[q] = ent[e];
        }
        double2 rec[U][5];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const double2 *cr = reinterpret_cast<const double2 *>(pk + (size_t)PMQ_PK * k[q]);
#pragma unroll
            for (int h = 0; h < 5; h++) rec[q][h] = cr[h];
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            double xr = rec[q][0].x - xi, yr = rec[q][0].y - yi, zr = rec[q][1].x - zi;
            row_image<BK>(geo, xr, yr, zr);
            const double d[6] = {rec[q][1].y, rec[q][2].x, rec[q][2].y, rec[q][3].x, rec[q][3].y, rec[q][4].x};
            entry_apply(wgt[q] * (double)r[q].r1, wgt[q] * (double)r[q].r2, xr, yr, zr, d, acc);
        }
    }
}

// entries per lane taken together by the product and the preconditioner,
// and the lanes per row of each
constexpr int SPMV_U = 2;
constexpr int PREC_U = 2;
constexpr int SPMV_G = 16;
constexpr int PREC_G = 8;
// each row's six outputs (two dipole sets x xyz) are written by its lanes
// g < 6 after the shuffle reduction, and the reduction needs a power of two
static_assert(SPMV_G >= 6 && (SPMV_G & (SPMV_G - 1)) == 0 && SPMV_G <= 64, "SpMV lanes per row");
static_assert(PREC_G >= 6 && (PREC_G & (PREC_G - 1)) == 0 && PREC_G <= 64, "preconditioner lanes per row");

// the product's arguments; ent points to TEnt<V> of the stored type
struct SpmvArgs {
    int npad;
    const int *rowptr;
    size_t cap;
    const void *ent;
    const int *col;
    RowGeom geo;
    const double *pk;  // the source vector packed with the positions (PMQ_PK per atom)
    double *dst;       // 2 x 3 x npad (overwritten; not written with the PCG step)
    const double *skip; // nonzero: return at once (or null)
    PcgVec pcg;        // EPI: the PCG step E1 on each finished row (pcg_steps.h)
    int check;         // the first launch of a PCG iteration: its convergence test first
};

// dst = T src for both dipole sets, the rows of 256-lane workgroup `wg`
// (G lanes per row), FP64 arithmetic; with EPI the rows' values go to the
// PCG step E1 instead (T src is then the whole operator: no PME part)
template <int G, int BK, class V, bool EPI = false>
__device__ __forceinline__ void spmv_rows(const SpmvArgs &a, int wg)
{
    if (a.skip && *a.skip != 0.0) return;
    const int tid = wg * 256 + (int)threadIdx.x;
    const int row = tid / G, g = tid % G;
    const int np = a.npad;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (row < np) {
        const int e1 = (int)min((size_t)a.rowptr[row + 1], a.cap);
        row_apply<G, SPMV_U, BK, V>(a.col, (const TEnt<V> *)a.ent, a.rowptr[row], e1, g, a.geo, a.geo.sx[row],
                                    a.geo.sy[row], a.geo.sz[row], a.pk, acc);
    }
#pragma unroll
    for (int c = 0; c < 6; c++)
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
    double beta[2] = {0.0, 0.0}, pv[2] = {0.0, 0.0};
    if (EPI) pcg_e1_begin(a.pcg, beta);
    if (row < np && g < 6) {
        double v = acc[0];
#pragma unroll
        for (int c = 1; c < 6; c++)
            if (g == c) v = acc[c];
        if (EPI) pv[g / 3] = pcg_e1(a.pcg, row, g, v, beta);
        else a.dst[(size_t)g * np + row] = v;
    }
    if (EPI) pcg_e1_end(a.pcg, pv[0], pv[1]);
}

// the product's arguments and box form of a context (k_pairs.hip)
SpmvArgs spmv_args(pmq_ctx *c, double *dst, const double *skip, const PcgVec *pcg = nullptr, int check = 0);
int spmv_box(const pmq_ctx *c);

} // namespace pmq
