// k_neighbors.hip -- spatial sort and the 64x64 block-tile list.
//
// Every evaluation: (1) key each atom by the Hilbert index of its cell
// (cells about 4 A wide in the box's fractional frame, or in the bounding box
// of a non-periodic system; with PME the grid's spreading bricks) with the
// atom index in the low 32 bits, so the sort is total and deterministic; (2)
// radix-sort the keys; (3) gather the permutation and bound each 64-atom
// block (one launch, a wave per block); (4) keep every block pair (I <= J)
// whose bounds come within the cutoff (flags, a device scan, the list in
// tile_slot order).  Runs of cells along a Hilbert curve are compact
// (consecutive cells share a face), so 64 consecutive atoms fill a small box:
// on a 12,288-atom water box with 4.1 A bricks and a 7 A cutoff the blocks'
// mean extent is 9.7 x 10.2 x 9.6 A and 4,503 block pairs are kept, against
// 16.0 x 11.9 x 9.5 A and 7,675 pairs for a Morton order of the same bricks.
//
// A block pair is kept when the minimum-image gap between the two bounding
// boxes is within the cutoff.  For orthogonal boxes and free clusters the
// gap is per axis; for monoclinic and triclinic boxes it is a sphere test
// over the 27 nearest lattice images; truncated-octahedron and dodecahedron
// boxes keep every pair (their image() is not a lattice minimum image).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "pme_common.h"
#include "pmq_internal.h"

namespace pmq {

DevBox make_box(const pmq_box *b)
{
    DevBox d{};
    d.kind = b ? b->kind : PMQ_BOX_NONE;
    if (!b || b->kind == PMQ_BOX_NONE) return d;
    const double radian = 57.29577951308232088;
    d.xbox = b->xbox;
    d.ybox = b->ybox;
    d.zbox = b->zbox;
    d.beta_sin = 1; d.beta_cos = 0; d.gamma_sin = 1; d.gamma_cos = 0; d.beta_term = 0; d.gamma_term = 1;
    if (b->kind == PMQ_BOX_TRI) {
        const double ac = cos(b->alpha / radian);
        d.beta_sin = sin(b->beta / radian);
        d.beta_cos = cos(b->beta / radian);
        d.gamma_sin = sin(b->gamma / radian);
        d.gamma_cos = cos(b->gamma / radian);
        d.beta_term = (ac - d.beta_cos * d.gamma_cos) / d.gamma_sin;
        d.gamma_term = sqrt(d.beta_sin * d.beta_sin - d.beta_term * d.beta_term);
    } else if (b->kind == PMQ_BOX_MONO) {
        d.beta_sin = sin(b->beta / radian);
        d.beta_cos = cos(b->beta / radian);
        d.gamma_term = d.beta_sin;
    }
    // lattice vectors as columns (lattice.cpp), reciprocal rows
    const double ar[3] = {d.xbox, 0, 0};
    const double br[3] = {d.ybox * d.gamma_cos, d.ybox * d.gamma_sin, 0};
    const double cr[3] = {d.zbox * d.beta_cos, d.zbox * d.beta_term, d.zbox * d.gamma_term};
    for (int a = 0; a < 3; a++) {
        d.lvec[a][0] = ar[a];
        d.lvec[a][1] = br[a];
        d.lvec[a][2] = cr[a];
    }
    const double vol = ar[0] * (br[1] * cr[2] - br[2] * cr[1]) - ar[1] * (br[0] * cr[2] - br[2] * cr[0])
                     + ar[2] * (br[0] * cr[1] - br[1] * cr[0]);
    // recip rows = (b x c, c x a, a x b) / vol
    d.recip[0][0] = (br[1] * cr[2] - cr[1] * br[2]) / vol;
    d.recip[0][1] = (br[2] * cr[0] - cr[2] * br[0]) / vol;
    d.recip[0][2] = (br[0] * cr[1] - cr[0] * br[1]) / vol;
    d.recip[1][0] = (cr[1] * ar[2] - ar[1] * cr[2]) / vol;
    d.recip[1][1] = (cr[2] * ar[0] - ar[2] * cr[0]) / vol;
    d.recip[1][2] = (cr[0] * ar[1] - ar[0] * cr[1]) / vol;
    d.recip[2][0] = (ar[1] * br[2] - br[1] * ar[2]) / vol;
    d.recip[2][1] = (ar[2] * br[0] - br[2] * ar[0]) / vol;
    d.recip[2][2] = (ar[0] * br[1] - br[0] * ar[1]) / vol;
    return d;
}

// bounding box of a free cluster (one workgroup)
__global__ __launch_bounds__(1024) void k_bbox(int n, const double *__restrict__ x, const double *__restrict__ y,
                                               const double *__restrict__ z, double *__restrict__ out)
{
    __shared__ double s[6][1024];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int a = 0; a < 3; a++) { lo[a] = fmin(lo[a], v[a]); hi[a] = fmax(hi[a], v[a]); }
    }
    for (int a = 0; a < 3; a++) { s[a][threadIdx.x] = lo[a]; s[3 + a][threadIdx.x] = hi[a]; }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; a++) {
                s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// fractional coordinate in [0,1) of atom i for the cell grid
__device__ __forceinline__ void frac_of(const DevBox &b, const double *bbox, double px, double py, double pz,
                                        double *f)
{
    if (b.kind == PMQ_BOX_NONE) {
        const double p[3] = {px, py, pz};
        for (int a = 0; a < 3; a++) {
            const double ext = fmax(bbox[3 + a] - bbox[a], 1e-6);
            f[a] = fmin(fmax((p[a] - bbox[a]) / ext, 0.0), 0.999999);
        }
    } else {
        for (int a = 0; a < 3; a++) {
            double v = b.recip[a][0] * px + b.recip[a][1] * py + b.recip[a][2] * pz;
            v -= floor(v);
            f[a] = v >= 1.0 ? 0.0 : v;
        }
    }
}

__global__ void k_cell_keys(int n, DevBox b, const double *__restrict__ bbox, int ncx, int ncy, int ncz, int bits,
                            const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, unsigned long long *__restrict__ keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f[3];
    frac_of(b, bbox, x[i], y[i], z[i], f);
    const unsigned cx = min((int)(f[0] * ncx), ncx - 1);
    const unsigned cy = min((int)(f[1] * ncy), ncy - 1);
    const unsigned cz = min((int)(f[2] * ncz), ncz - 1);
    const unsigned m = hilbert3(cx, cy, cz, bits);
    keys[i] = ((unsigned long long)m << 32) | (unsigned)i;
}

// with PME on: the cells are the spreading bricks (home_brick, the same
// arithmetic the spreading kernels use), so every brick's atoms end up
// contiguous in the sorted order; the bricks' sorted-atom ranges are
// cleared here (empty) and set by k_gather
__global__ void k_brick_keys(int n, BrickGeom g, int bits, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, unsigned long long *__restrict__ keys,
                             int *__restrict__ bstart, int *__restrict__ bend)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbricks = g.nb[0] * g.nb[1] * g.nb[2];
    for (int q = i; q < nbricks; q += gridDim.x * blockDim.x) bstart[q] = bend[q] = 0;
    if (i >= n) return;
    int hb[3];
    home_brick(g.A, g.K, g.nb, x[i], y[i], z[i], hb);
    keys[i] = ((unsigned long long)hilbert3(hb[0], hb[1], hb[2], bits) << 32) | (unsigned)i;
}

// wave (64-lane) minimum and maximum
__device__ __forceinline__ double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// One wave per 64-atom block: the permutation, sorted positions and per-atom
// flags (bit0 valid, bit1 active); with PME bricks (bstart non-null) the
// sorted-atom range [bstart, bend) of every non-empty brick, from where the
// brick part of the sorted keys changes (the brick's lexicographic index from
// the atom's home brick); then the block's bounds: axis-aligned in the
// (wrapped) fractional-cell frame, as a Cartesian center and half extent,
// and the span of the stored (unwrapped) coordinates of every slot, padding
// included, which the tile mask kernel's one-image test reads
__global__ __launch_bounds__(PMQ_BLOCK) void k_gather(int n, int npad, const unsigned long long *__restrict__ skeys,
                                                      const double *__restrict__ x, const double *__restrict__ y,
                                                      const double *__restrict__ z, const int *__restrict__ usef,
                                                      const int *__restrict__ mol, int *__restrict__ perm,
                                                      int *__restrict__ iperm, double *__restrict__ sx,
                                                      double *__restrict__ sy, double *__restrict__ sz,
                                                      int *__restrict__ sflag, int *__restrict__ smol, BrickGeom g,
                                                      int *__restrict__ bstart, int *__restrict__ bend, DevBox b,
                                                      double *__restrict__ blk)
{
    const int I = blockIdx.x, lane = threadIdx.x;
    const int s = I * PMQ_BLOCK + lane;
    double q[3] = {0.0, 0.0, 0.0};
    if (s < n) {
        const unsigned long long key = skeys[s];
        const int i = (int)(key & 0xffffffffu);
        perm[s] = i;
        iperm[i] = s;
        q[0] = x[i];
        q[1] = y[i];
        q[2] = z[i];
        sx[s] = q[0];
        sy[s] = q[1];
        sz[s] = q[2];
        if (bstart) {
            const unsigned h = (unsigned)(key >> 32);
            const bool first = s == 0 || (unsigned)(skeys[s - 1] >> 32) != h;
            const bool last = s == n - 1 || (unsigned)(skeys[s + 1] >> 32) != h;
            if (first || last) {
                int hb[3];
                home_brick(g.A, g.K, g.nb, x[i], y[i], z[i], hb);
                const int b = (hb[2] * g.nb[1] + hb[1]) * g.nb[0] + hb[0];
                if (first) bstart[b] = s;
                if (last) bend[b] = s + 1;
            }
        }
        sflag[s] = 1 | (usef[i] ? 2 : 0);
        smol[s] = mol[i];
    } else if (s < npad) {
        perm[s] = -1;
        sx[s] = sy[s] = sz[s] = 0.0;
        sflag[s] = 0;
        smol[s] = -1;
    }
    // the block's bounds
    const bool valid = s < n;
    double p[3] = {q[0], q[1], q[2]};
    if (valid && b.kind != PMQ_BOX_NONE) {
        // wrap into the primary cell so the block is compact
        double f[3];
        for (int a = 0; a < 3; a++) {
            f[a] = b.recip[a][0] * p[0] + b.recip[a][1] * p[1] + b.recip[a][2] * p[2];
            f[a] -= floor(f[a]);
        }
        for (int a = 0; a < 3; a++) p[a] = b.lvec[a][0] * f[0] + b.lvec[a][1] * f[1] + b.lvec[a][2] * f[2];
    }
    double lo[3], hi[3], rlo[3], rhi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = wave_min(valid ? p[a] : 1e300);
        hi[a] = wave_max(valid ? p[a] : -1e300);
        rlo[a] = wave_min(q[a]);
        rhi[a] = wave_max(q[a]);
    }
    if (lane < 3) {
        const double l0 = lane == 0 ? lo[0] : lane == 1 ? lo[1] : lo[2];
        const double h0 = lane == 0 ? hi[0] : lane == 1 ? hi[1] : hi[2];
        blk[PMQ_BLK_STRIDE * I + lane] = 0.5 * (l0 + h0);
        blk[PMQ_BLK_STRIDE * I + 3 + lane] = 0.5 * (h0 - l0);
        blk[PMQ_BLK_STRIDE * I + 6 + lane] = lane == 0 ? rlo[0] : lane == 1 ? rlo[1] : rlo[2];
        blk[PMQ_BLK_STRIDE * I + 9 + lane] = lane == 0 ? rhi[0] : lane == 1 ? rhi[1] : rhi[2];
    }
}

__device__ __forceinline__ void tri_index(long long t, int nb, int &I, int &J)
{
    // t enumerates I <= J row by row: row I has nb - I entries
    // solve for I with the quadratic, then fix rounding
    const double dn = 2.0 * nb + 1.0;
    int i = (int)floor((dn - sqrt(dn * dn - 8.0 * (double)t)) * 0.5);
    if (i < 0) i = 0;
    while (i > 0 && (long long)i * nb - (long long)i * (i - 1) / 2 > t) i--;
    while ((long long)(i + 1) * nb - (long long)(i + 1) * i / 2 <= t) i++;
    I = i;
    J = (int)(t - ((long long)i * nb - (long long)i * (i - 1) / 2)) + i;
}

// whether block pair t (I <= J, row by row) comes within the cutoff
__device__ __forceinline__ int tile_keep(long long t, int nblk, const DevBox &b, double cut,
                                         const double *__restrict__ blk)
{
    int I, J;
    tri_index(t, nblk, I, J);
    bool keep = true;
    const double *bi = blk + PMQ_BLK_STRIDE * I, *bj = blk + PMQ_BLK_STRIDE * J;
    double d[3] = {bj[0] - bi[0], bj[1] - bi[1], bj[2] - bi[2]};
    if (b.kind == PMQ_BOX_NONE || b.kind == PMQ_BOX_ORTHO) {
        if (b.kind == PMQ_BOX_ORTHO) {
            const double L[3] = {b.xbox, b.ybox, b.zbox};
            for (int a = 0; a < 3; a++) d[a] -= L[a] * rint(d[a] / L[a]);
        }
        double g2 = 0.0;
        for (int a = 0; a < 3; a++) {
            const double g = fmax(0.0, fabs(d[a]) - bi[3 + a] - bj[3 + a]);
            g2 += g * g;
        }
        keep = g2 <= cut * cut * (1.0 + 1e-12) + 1e-12;
    } else if (b.kind == PMQ_BOX_MONO || b.kind == PMQ_BOX_TRI) {
        // fractional wrap of the center offset, then the nearest of 27 images
        double f[3];
        for (int a = 0; a < 3; a++) {
            f[a] = b.recip[a][0] * d[0] + b.recip[a][1] * d[1] + b.recip[a][2] * d[2];
            f[a] -= rint(f[a]);
        }
        double best = 1e300;
        for (int u = -1; u <= 1; u++)
            for (int v = -1; v <= 1; v++)
                for (int w = -1; w <= 1; w++) {
                    const double g[3] = {f[0] + u, f[1] + v, f[2] + w};
                    double r2 = 0.0;
                    for (int a = 0; a < 3; a++) {
                        const double c = b.lvec[a][0] * g[0] + b.lvec[a][1] * g[1] + b.lvec[a][2] * g[2];
                        r2 += c * c;
                    }
                    best = fmin(best, r2);
                }
        const double ri = sqrt(bi[3] * bi[3] + bi[4] * bi[4] + bi[5] * bi[5]);
        const double rj = sqrt(bj[3] * bj[3] + bj[4] * bj[4] + bj[5] * bj[5]);
        keep = sqrt(best) - ri - rj <= cut * (1.0 + 1e-12) + 1e-9;
    }
    return keep ? 1 : 0;
}

__global__ void k_tile_flags(int nblk, long long npairs, DevBox b, double cut, const double *__restrict__ blk,
                             int *__restrict__ flag)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    flag[t] = tile_keep(t, nblk, b, cut, blk);
}

// the listed tiles (tile_slot order); the last pair's thread also stores the
// count and clears the count of further tile pieces (tile_pairs.h)
__global__ void k_tile_write(int nblk, long long npairs, const int *__restrict__ flag, const int *__restrict__ scan,
                             int2 *__restrict__ tiles, int *__restrict__ count)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == npairs - 1) {
        count[0] = scan[t] + flag[t];
        count[1] = 0; // the tile masks count the further pieces of split tiles here
    }
    if (t >= npairs || !flag[t]) return;
    int I, J;
    tri_index(t, nblk, I, J);
    tiles[tile_slot(I, J, nblk, scan[t])] = make_int2(I, J);
}

int build_neighbors(pmq_ctx *c, double cutoff, const double *px, const double *py, const double *pz,
                    const int *use)
{
    const int n = c->n, npad = c->npad, nblk = c->nblk;
    hipStream_t st = c->stream;
    const DevBox &b = c->box;
    if (!px) {
        px = c->x.p;
        py = c->y.p;
        pz = c->z.p;
    }
    if (!use) use = c->usef.p;
    if (b.kind == PMQ_BOX_NONE) {
        hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, st, n, px, py, pz, c->bbox.p);
    }
    // cells about 4 A wide (at most 1024 per axis)
    int nc[3];
    if (b.kind == PMQ_BOX_NONE) {
        // cell counts from the host-side estimate of the extent
        for (int a = 0; a < 3; a++) nc[a] = 64;
    } else {
        const double len[3] = {b.xbox, b.ybox, b.zbox};
        for (int a = 0; a < 3; a++) nc[a] = std::max(1, std::min(1024, (int)(len[a] / 4.0)));
    }
    const int tb = 256;
    int mx = 1;
    for (int a = 0; a < 3; a++) mx = std::max(mx, c->pme_cells ? c->pme_nb[a] : nc[a]);
    const int bits = hilbert_bits(mx);
    BrickGeom g{};
    int *bstart = nullptr, *bend = nullptr;
    if (c->pme_cells) {
        for (int a = 0; a < 3; a++) {
            for (int q = 0; q < 3; q++) g.A[a][q] = c->pmeA[a][q];
            g.K[a] = c->K[a];
            g.nb[a] = c->pme_nb[a];
        }
        bstart = c->pme_bstart.p;
        bend = c->pme_bend.p;
        if (!bstart || !bend) {
            c->err = "build_neighbors: PME bricks not set up";
            return PMQ_ERR_STATE;
        }
        hipLaunchKernelGGL(k_brick_keys, dim3((n + tb - 1) / tb), dim3(tb), 0, st, n, g, bits, px, py, pz,
                           c->keys.p, bstart, bend);
    } else {
        hipLaunchKernelGGL(k_cell_keys, dim3((n + tb - 1) / tb), dim3(tb), 0, st, n, b, c->bbox.p, nc[0], nc[1],
                           nc[2], bits, px, py, pz, c->keys.p);
    }
    // The low 32 bits of a key hold the atom index and keys are written in
    // index order, so a stable sort on the Hilbert bits alone gives the same
    // order as a sort on all 64: only bits [32, 32 + 3 b) are passed, b the
    // bits of the largest cell (or brick) index on any axis.
    const unsigned end_bit = 32u + 3u * (unsigned)bits;
    size_t bytes = 0;
    PMQ_HIP(rocprim::radix_sort_keys(nullptr, bytes, c->keys.p, c->keys_sorted.p, (size_t)n, 32u, end_bit, st));
    if (bytes > c->scratch_sort_bytes) {
        c->err = "build_neighbors: sort scratch too small";
        return PMQ_ERR_STATE;
    }
    bytes = c->scratch_sort_bytes;
    PMQ_HIP(rocprim::radix_sort_keys(c->scratch_sort.p, bytes, c->keys.p, c->keys_sorted.p, (size_t)n, 32u, end_bit,
                                     st));
    if (nblk * PMQ_BLOCK != npad) {
        c->err = "build_neighbors: the padded atom count is not a whole number of blocks";
        return PMQ_ERR_STATE;
    }
    hipLaunchKernelGGL(k_gather, dim3(nblk), dim3(PMQ_BLOCK), 0, st, n, npad, c->keys_sorted.p, px, py, pz, use,
                       c->molcule.p, c->perm.p, c->iperm.p, c->sx.p, c->sy.p, c->sz.p, c->sflag.p, c->smol.p, g,
                       bstart, bend, b, c->blk.p);
    const long long npairs = (long long)nblk * (nblk + 1) / 2;
    const int gb = (int)((npairs + tb - 1) / tb);
    hipLaunchKernelGGL(k_tile_flags, dim3(gb), dim3(tb), 0, st, nblk, npairs, b, cutoff, c->blk.p, c->tileflag.p);
    bytes = c->scratch_sort_bytes;
    PMQ_HIP(rocprim::exclusive_scan(c->scratch_sort.p, bytes, c->tileflag.p, c->tilescan.p, 0, (size_t)npairs,
                                    rocprim::plus<int>(), st));
    hipLaunchKernelGGL(k_tile_write, dim3(gb), dim3(tb), 0, st, nblk, npairs, c->tileflag.p, c->tilescan.p,
                       c->tiles.p, c->ntiles_d.p);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// scratch sizes of the two device-wide primitives for n atoms / nblk blocks
size_t neighbor_scratch_bytes(int n, int nblk)
{
    size_t a = 0, s = 0;
    const long long npairs = (long long)nblk * (nblk + 1) / 2;
    (void)rocprim::radix_sort_keys(nullptr, a, (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                   (size_t)n, 0, 64);
    (void)rocprim::exclusive_scan(nullptr, s, (int *)nullptr, (int *)nullptr, 0, (size_t)npairs,
                                  rocprim::plus<int>());
    return std::max(a, s);
}

} // namespace pmq
