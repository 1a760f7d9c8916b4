// k_pairs.hip -- the per-evaluation tile structure shared by the pair
// kernels (in-cutoff masks, per-block tile lists, deterministic reduction of
// per-tile partial sums) and the dipole-field operator T as a row-owned
// sparse matrix.
//
// Pair kernels over (I,J) tiles do not accumulate into per-atom arrays with
// atomics: every tile stores its i-side and j-side partial sums, and
// k_tile_reduce adds, for every atom, the partials of the tiles of its block
// in a fixed order (the block's tile list), so per-atom results are
// bit-reproducible and no global atomics are issued.
//
// The induced-dipole solve applies T (the Thole-damped, Ewald real-space
// dipole-dipole field tensor, u-scaled for excluded pairs) to a new pair of
// dipole vectors in every PCG iteration while the geometry stays fixed.  So
// T is built once per evaluation and every iteration is a sparse
// matrix-vector product that reads it back.  Each pair tensor is
// R2 r r^T - R1 I, so a stored entry is its two scalars and its column (the
// product rebuilds r from the sorted positions); each pair is stored in both
// rows.
// Every row is owned by one group of lanes, so the product needs no atomics
// and sums each row in a fixed order (bit-reproducible run to run).
//
// Build, all per evaluation on the library's stream:
//   k_tile_masks   per (I,J) tile: 64 row masks (bit j of lane i: pair i,j
//                  within the cutoff) and the transposed column masks, the
//                  scaled-pair mask rows, the tile's pieces (tile_pairs.h)
//                  and the tile count clamped to the buffers' capacity; in
//                  the same launch, per block: the tiles it takes part in,
//                  ordered by the other block's index
//   k_row_counts   per block: each lane's entry offset within its row for
//                  every tile of its list, and the row length
//   (exclusive scan of row lengths -> row pointers)
//   k_pair_fill    per tile: the tensor scalars of every masked pair, written
//                  to the two rows at the offsets above (no atomics); block 0
//                  also checks the entry total against the capacity
//   k_excl_fold    adds (u - 1) lambda rr corrections of excluded pairs to
//                  their two entries
// Apply:
//   k_pol_spmv<G,BK>  G lanes per row, both dipole sets at once (BK: the box
//                     form, for rebuilding r).
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include "pair_matrix.h"
#include "pair_mpole.h"
#include "pmq_internal.h"
#include "tile_pairs.h"

namespace pmq {

struct MaskArgs {
    const int2 *tiles;
    const int *ntiles;         // tiles listed by the neighbor search
    int cap;                   // tiles the buffers hold: the count used is min(ntiles, cap)
    int *ntiles_eff;           // that count, for the later tile kernels
    RunStatus *st;             // the listed count, and overflow bit 1 when it exceeds cap
    const double *sx, *sy, *sz;
    const int *sflag;
    DevBox box;
    double off2;
    unsigned long long *mask;  // 128 per tile: [0,64) row masks, [64,128) column masks
    unsigned long long *mmask; // 64 per tile: row masks of the pairs with an active atom (multipole energy)
    double us2;                // preconditioner cutoff squared (<= off2)
    unsigned long long *nmask; // 128 per tile: the pairs within it (row, column masks), or null
    const double *blk;         // block bounds (PMQ_BLK_STRIDE per block)
    unsigned long long *xmask; // 128 per tile: copies of the in-cutoff and active rows (scaled pairs cleared later)
    int *tpc;                  // per tile: piece code of its in-cutoff pair count (tile_pairs.h)
    int *xlist, *xcnt;         // further pieces of the split tiles (tile | piece << 28), their count
    double *tscal;             // tile-item scalars: zeroed here for the tiles without pairs (no item of theirs runs)
    size_t tstride;
};

// BK 0: no box or a rectangular one (one image per tile where it fits, else
// image()'s per-axis wrapping); BK 1: other cells (image() per pair)
// Transpose of the 64 x 64 bit matrix whose row l is lane l's word (bit c =
// column c): lane c receives column c.  Six butterfly stages; stage j swaps,
// between lanes l and l ^ j, the columns whose index has bit j set in the
// upper lane with those whose index has it clear in the lower (each stage
// exchanges one bit of the row index with the same bit of the column index).
__device__ __forceinline__ unsigned long long bit_transpose64(unsigned long long x, int lane)
{
    const unsigned long long M[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                     0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int j = 32 >> k;
        const unsigned long long o = __shfl_xor(x, j, 64);
        x = (lane & j) ? ((x & ~M[k]) | ((o & ~M[k]) >> j)) : ((x & M[k]) | ((o & M[k]) << j));
    }
    return x;
}

// the masks of tiles wg, wg + nwg, ... (one wave each)
template <int BK>
__device__ __forceinline__ void tile_masks_body(const MaskArgs &a, int wg, int nwg)
{
    const int lane = threadIdx.x;
    const int nlisted = *a.ntiles;
    const int ntiles = min(nlisted, a.cap);
    if (wg == 0 && lane == 0) {
        *a.ntiles_eff = ntiles;
        a.st->ntiles = nlisted;
        if (nlisted > a.cap) a.st->overflow |= 1;
    }
    __shared__ double jx[64], jy[64], jz[64];
    __shared__ int jflag[64];
    const double L[3] = {a.box.xbox, a.box.ybox, a.box.zbox};
    for (int t = wg; t < ntiles; t += nwg) {
        const int2 IJ = a.tiles[t];
        const int si = IJ.x * 64 + lane, sj = IJ.y * 64 + lane;
        // One image for the whole tile where every pair's offset v = xj - xi
        // (bounded by the two blocks' coordinate spans) lies inside one window
        // of image()'s wrapping: |v| < L/2 (left as is), or L/2 < |v| < 3L/2
        // (one edge subtracted toward zero).  The subtraction is then the
        // same operation on the same operand as image()'s, so the vector is
        // the same to the bit.  Other boxes, and tiles that straddle a window
        // edge, take image() per pair.
        double sh[3] = {0.0, 0.0, 0.0};
        bool one = BK == 0;
        if (BK == 0 && a.box.kind == PMQ_BOX_ORTHO) {
            const double *bi = a.blk + PMQ_BLK_STRIDE * IJ.x, *bj = a.blk + PMQ_BLK_STRIDE * IJ.y;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const double dlo = bj[6 + q] - bi[9 + q], dhi = bj[9 + q] - bi[6 + q];
                const double h = 0.5 * L[q], eps = 1e-9 * L[q];
                if (dlo > h + eps && dhi < 3.0 * h - eps) sh[q] = L[q];
                else if (dlo > -3.0 * h + eps && dhi < -h - eps) sh[q] = -L[q];
                else if (!(dlo > -h + eps && dhi < h - eps)) one = false;
            }
        }
        __syncthreads();
        jx[lane] = a.sx[sj]; jy[lane] = a.sy[sj]; jz[lane] = a.sz[sj];
        jflag[lane] = a.sflag[sj];
        const double xi = a.sx[si], yi = a.sy[si], zi = a.sz[si];
        const int fi = a.sflag[si];
        __syncthreads();
        unsigned long long m = 0ull, mm = 0ull, nm = 0ull;
        if (fi & 1) {
            const bool diag = IJ.x == IJ.y;
            if (one) {
#pragma unroll 8
                for (int jj = 0; jj < 64; jj++) {
                    const int fj = jflag[jj];
                    const double xr = (jx[jj] - xi) - sh[0], yr = (jy[jj] - yi) - sh[1], zr = (jz[jj] - zi) - sh[2];
                    const double r2 = xr * xr + yr * yr + zr * zr;
                    const bool in = (fj & 1) && (!diag || jj > lane) && r2 <= a.off2;
                    m |= (unsigned long long)in << jj;
                    mm |= (unsigned long long)(in && ((fi | fj) & 2)) << jj;
                    nm |= (unsigned long long)(in && r2 <= a.us2) << jj;
                }
            } else {
                for (int jj = diag ? lane + 1 : 0; jj < 64; jj++) {
                    const int fj = jflag[jj];
                    if (!(fj & 1)) continue;
                    double xr = jx[jj] - xi, yr = jy[jj] - yi, zr = jz[jj] - zi;
                    if constexpr (BK == 0) {
                        // image() of a rectangular box
                        image_wrap<double>(xr, L[0]);
                        image_wrap<double>(yr, L[1]);
                        image_wrap<double>(zr, L[2]);
                    } else {
                        image<double>(a.box, xr, yr, zr);
                    }
                    const double r2 = xr * xr + yr * yr + zr * zr;
                    if (r2 <= a.off2) {
                        m |= 1ull << jj;
                        if ((fi | fj) & 2) mm |= 1ull << jj;
                        if (r2 <= a.us2) nm |= 1ull << jj;
                    }
                }
            }
        }
        // column mask of lane c: bit l set where lane l's row mask has bit c
        const unsigned long long col = bit_transpose64(m, lane);
        if (a.nmask) {
            a.nmask[(size_t)t * 128 + lane] = nm;
            a.nmask[(size_t)t * 128 + 64 + lane] = bit_transpose64(nm, lane);
        }
        a.mask[(size_t)t * 128 + lane] = m;
        a.mask[(size_t)t * 128 + 64 + lane] = col;
        a.mmask[(size_t)t * 64 + lane] = mm;
        a.xmask[(size_t)t * 128 + lane] = m;
        a.xmask[(size_t)t * 128 + 64 + lane] = mm;
        int npair = __popcll(m);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) npair += __shfl_xor(npair, o, 64);
        if (lane == 0) {
            const int code = tile_piece_code(npair, ntiles <= PMQ_SPLIT_TILES ? PMQ_TILE_PIECES : 1), np = code & 15;
            a.tpc[t] = code;
            if (np == 0)
                for (int q = 0; q < PMQ_TILE_NS; q++) a.tscal[(size_t)q * a.tstride + t] = 0.0;
            if (np > 1) {
                const int k = atomicAdd(a.xcnt, np - 1);
                for (int p = 1; p < np; p++) a.xlist[k + p - 1] = t | (p << 28);
            }
        }
    }
}

// tiles of block I in the order of the other block J (one wave per block);
// tiles beyond the capacity are left out (the evaluation is then redone)
struct AdjArgs {
    int nblk, cap;
    const int *flag, *scan; // kept block pairs and their exclusive count, in (I, J) order
    int *adj, *adjn;
};

__device__ __forceinline__ void block_adj_body(const AdjArgs &q, int I)
{
    const int nblk = q.nblk, cap = q.cap;
    const int *flag = q.flag, *scan = q.scan;
    int *adj = q.adj, *adjn = q.adjn;
    const int lane = threadIdx.x;
    int count = 0;
    for (int base = 0; base < nblk; base += 64) {
        const int J = base + lane;
        int f = 0, slot = 0;
        if (J < nblk) {
            const long long lo = min(I, J), hi = max(I, J);
            const long long tri = lo * nblk - lo * (lo - 1) / 2 + (hi - lo);
            slot = tile_slot((int)lo, (int)hi, nblk, scan[tri]);
            f = flag[tri] && slot < cap;
        }
        const unsigned long long b = __ballot(f != 0);
        const int pos = count + __popcll(b & ((1ull << lane) - 1ull));
        if (f) adj[(size_t)I * nblk + pos] = slot;
        count += __popcll(b);
    }
    if (lane == 0) adjn[I] = count;
}

// one launch: workgroups [0, nblk) the blocks' tile lists, the rest the
// tile masks (independent: both need only the tile list)
template <int BK>
__global__ __launch_bounds__(64) void k_tile_masks(MaskArgs a, AdjArgs q)
{
    if ((int)blockIdx.x < q.nblk) block_adj_body(q, blockIdx.x);
    else tile_masks_body<BK>(a, (int)blockIdx.x - q.nblk, (int)gridDim.x - q.nblk);
}

// the mask of lane `lane` of block I for tile t, in the orientation of I
__device__ __forceinline__ unsigned long long side_mask(const unsigned long long *mask, int t, int2 IJ, int I,
                                                         int lane)
{
    const unsigned long long r = mask[(size_t)t * 128 + lane], c = mask[(size_t)t * 128 + 64 + lane];
    if (IJ.x == IJ.y) return r | c;
    return IJ.x == I ? r : c;
}

// Row layout: with NM a row's entries within the preconditioner cutoff come
// first (nmask; their count in rownear), then the others.  soff and nsoff
// are each tile's offsets within the two parts.  A block's tile list is
// split over RC_W waves: each counts its share first, the counts of the
// waves before it give its starting offsets, and then it writes its tiles'
// offsets.  Tiles are taken U at a time with every load issued
// unconditionally (indices clamped to the last tile of the share), so the
// loads of a group are in flight together.
constexpr int RC_W = 4;

template <bool NM>
__global__ __launch_bounds__(64 * RC_W) void k_row_counts(int nblk, const int2 *__restrict__ tiles,
                                                          const unsigned long long *__restrict__ mask,
                                                          const unsigned long long *__restrict__ nmask,
                                                          const int *__restrict__ adj, const int *__restrict__ adjn,
                                                          int *__restrict__ soff, int *__restrict__ nsoff,
                                                          int *__restrict__ rowcnt, int *__restrict__ rownear)
{
    constexpr int U = 8;
    __shared__ int part[2][RC_W][64];
    const int I = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int na = adjn[I];
    const int *lst = adj + (size_t)I * nblk;
    const int per = (na + RC_W - 1) / RC_W;
    const int b0 = min(na, w * per), b1 = min(na, b0 + per);
    int run = 0, nrun = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            // the offsets where this wave's share starts: the counts of the waves before it
            part[0][w][lane] = run;
            part[1][w][lane] = nrun;
            __syncthreads();
            run = nrun = 0;
            for (int q = 0; q < w; q++) {
                run += part[0][q][lane];
                nrun += part[1][q][lane];
            }
        }
        for (int s0 = b0; s0 < b1; s0 += U) {
            int tt[U];
            int2 ij[U];
            unsigned long long m[U], nm[U];
#pragma unroll
            for (int k = 0; k < U; k++) tt[k] = lst[min(s0 + k, b1 - 1)];
#pragma unroll
            for (int k = 0; k < U; k++) ij[k] = tiles[tt[k]];
#pragma unroll
            for (int k = 0; k < U; k++) {
                m[k] = side_mask(mask, tt[k], ij[k], I, lane);
                nm[k] = NM ? side_mask(nmask, tt[k], ij[k], I, lane) : 0ull;
            }
#pragma unroll
            for (int k = 0; k < U; k++) {
                if (s0 + k >= b1) break;
                if (pass == 1) {
                    const size_t t = (size_t)tt[k];
                    const int2 IJ = ij[k];
                    if (IJ.x == IJ.y) {
                        soff[t * 128 + lane] = run;
                        soff[t * 128 + 64 + lane] = run;
                        if (NM) {
                            nsoff[t * 128 + lane] = nrun;
                            nsoff[t * 128 + 64 + lane] = nrun;
                        }
                    } else {
                        soff[t * 128 + (IJ.x == I ? 0 : 64) + lane] = run;
                        if (NM) nsoff[t * 128 + (IJ.x == I ? 0 : 64) + lane] = nrun;
                    }
                }
                run += __popcll(m[k] & ~nm[k]);
                nrun += __popcll(nm[k]);
            }
        }
    }
    if (w == RC_W - 1) {
        rowcnt[I * 64 + lane] = run + nrun;
        if (NM) rownear[I * 64 + lane] = nrun;
    }
    if (I == 0 && threadIdx.x == 0) rowcnt[nblk * 64] = 0; // the scan's last element becomes the entry total
}

// V: stored type of the tensor (double, or float in the mixed form); T: the
// arithmetic of the pair terms
template <class V>
struct FillArgs {
    const int2 *tiles;
    const int *ntiles;
    TileItems items;
    const double *sx, *sy, *sz, *spol;
    const double *scp; // sorted core charge, valence exponent (HIPPO damping)
    DevBox box;
    double aewald;
    const unsigned long long *mask;
    const int *soff, *rowptr;
    int npad;
    size_t cap;
    TEnt<V> *ent;
    int *col;
    // preconditioner: the pairs within its cutoff (null = none), their
    // offsets, each row's count of them, and their plain tensors
    const unsigned long long *nmask;
    const int *nsoff, *rownear;
    TEnt<V> *ent_s;
    RunStatus *st;  // the entry total (rowptr[npad]) against cap: overflow bit 2
};

template <class V, class T>
__device__ __forceinline__ void store_T(TEnt<V> *ent, size_t e, const T *R)
{
    TEnt<V> r;
    r.r1 = V(R[1]);
    r.r2 = V(R[2]);
    ent[e] = r;
}

// CPD: HIPPO damping (overlap factors of the two valence exponents) instead
// of Thole.  Pairs are taken 64 at a time from the tile's compacted list
// (tile_pairs.h); each pair's two entries sit at its row's base plus the
// number of the row's partners before it in the same part of the row (near
// pairs first).  Near pairs also store the preconditioner's tensor (no
// Ewald) at the same entry of ent_s.
template <bool EW, class T, class V, bool CPD = false>
__global__ __launch_bounds__(64) void k_pair_fill(FillArgs<V> a)
{
    const int lane = threadIdx.x;
    const int np = a.npad;
    const int ntiles = *a.ntiles;
    __shared__ double bx[2][64], by[2][64], bz[2][64];
    __shared__ T bpd[2][64], bth[2][64];
    __shared__ unsigned long long bmask[2][64], bnmask[2][64];
    __shared__ int bbase[2][64], bnbase[2][64];
    __shared__ unsigned short plist[64 * 64];
    if (blockIdx.x == 0 && lane == 0) {
        const int total = a.rowptr[np];
        a.st->nnz = total;
        if ((size_t)total > a.cap) a.st->overflow |= 2;
    }
    for (int v = blockIdx.x, nitems = ntiles + *a.items.xcnt; v < nitems; v += gridDim.x) {
        TileItem it;
        if (!tile_item(v, ntiles, a.items, it)) continue;
        const int t = it.t;
        const int2 IJ = a.tiles[t];
        const bool diag = IJ.x == IJ.y;
        const unsigned long long rm = a.mask[(size_t)t * 128 + lane], cm = a.mask[(size_t)t * 128 + 64 + lane];
        const unsigned long long nr = a.nmask ? a.nmask[(size_t)t * 128 + lane] : 0ull;
        const unsigned long long nc = a.nmask ? a.nmask[(size_t)t * 128 + 64 + lane] : 0ull;
        __syncthreads();
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int s = (side ? IJ.y : IJ.x) * 64 + lane;
            bx[side][lane] = a.sx[s]; by[side][lane] = a.sy[s]; bz[side][lane] = a.sz[s];
            bth[side][lane] = CPD ? 0.0 : a.spol[np + s];
            bpd[side][lane] = CPD ? a.scp[np + s] : a.spol[2 * np + s];
            // this lane's atom on each side: its masks of partners in the tile
            // (all, near) and the entry bases of the two parts of its row
            bmask[side][lane] = diag ? (rm | cm) : (side ? cm : rm);
            bnmask[side][lane] = diag ? (nr | nc) : (side ? nc : nr);
            const int so = (int)((size_t)t * 128 + side * 64 + lane);
            bbase[side][lane] = a.rowptr[s] + (a.nmask ? a.rownear[s] : 0) + a.soff[so];
            bnbase[side][lane] = a.nmask ? a.rowptr[s] + a.nsoff[so] : 0;
        }
        const int npair = min(tile_pair_list(rm, plist, lane), it.b1 * 64);
        __syncthreads();
        for (int base = it.b0 * 64; base < npair; base += 64) {
            if (base + lane >= npair) break;
            const int ent = plist[base + lane];
            const int il = ent >> 8, jl = ent & 63;
            double xd = bx[1][jl] - bx[0][il], yd = by[1][jl] - by[0][il], zd = bz[1][jl] - bz[0][il];
            image<double>(a.box, xd, yd, zd);
            const T xr = T(xd), yr = T(yd), zr = T(zd);
            const T r2 = xr * xr + yr * yr + zr * zr;
            T lam[6], R[6];
            if (CPD) overlap_damp<T>(sqrt(r2), bpd[0][il], bpd[1][jl], lam);
            else thole_lambda<T>(sqrt(r2), bpd[0][il], bpd[1][jl], bth[0][il], bth[1][jl], lam);
            pol_series<EW, T>(r2, T(a.aewald), lam, R);
            const unsigned long long lo_j = (1ull << jl) - 1ull, lo_i = (1ull << il) - 1ull;
            const unsigned long long ni = bnmask[0][il], nj = bnmask[1][jl];
            const bool nearp = (ni >> jl) & 1ull;
            const size_t ei = nearp ? (size_t)bnbase[0][il] + __popcll(ni & lo_j)
                                    : (size_t)bbase[0][il] + __popcll(bmask[0][il] & ~ni & lo_j);
            const size_t ej = nearp ? (size_t)bnbase[1][jl] + __popcll(nj & lo_i)
                                    : (size_t)bbase[1][jl] + __popcll(bmask[1][jl] & ~nj & lo_i);
            if (ei < a.cap) {
                store_T(a.ent, ei, R);
                a.col[ei] = IJ.y * 64 + jl;
            }
            if (ej < a.cap) {
                store_T(a.ent, ej, R);
                a.col[ej] = IJ.x * 64 + il;
            }
            if (nearp) {
                T Rs[6];
                pol_series<false, T>(r2, T(0), lam, Rs);
                if (ei < a.cap) store_T(a.ent_s, ei, Rs);
                if (ej < a.cap) store_T(a.ent_s, ej, Rs);
            }
        }
    }
}

template <class V>
struct FoldArgs {
    const ExclPair *excl;
    int nexcl;
    const int *iperm;
    const double *sx, *sy, *sz, *spol;
    const double *scp;
    int cpd;
    DevBox box;
    int npad;
    const int *rowptr;
    size_t cap;
    TEnt<V> *ent;
    const int *col;
    const int *rownear; // preconditioner: near entries per row (null = none)
    TEnt<V> *ent_s;
};

template <class V>
__device__ __forceinline__ void add_T(TEnt<V> *ent, size_t e, const double *R)
{
    ent[e].r1 = V((double)ent[e].r1 + R[1]);
    ent[e].r2 = V((double)ent[e].r2 + R[2]);
}

// excluded pairs present in the matrix get (u - 1) lambda rr added to their
// two entries; each entry belongs to one excluded pair, so no atomics
template <class V>
__global__ __launch_bounds__(64) void k_excl_fold(FoldArgs<V> a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nexcl) return;
    const ExclPair q = a.excl[t];
    if (q.u == 1.0) return;
    const int np = a.npad;
    const int si = a.iperm[q.i], sk = a.iperm[q.k];
    int ei = -1, ek = -1;
    const int ci = (int)min((size_t)a.rowptr[si + 1], a.cap), ck = (int)min((size_t)a.rowptr[sk + 1], a.cap);
    for (int e = a.rowptr[si]; e < ci; e++)
        if (a.col[e] == sk) { ei = e; break; }
    if (ei < 0) return; // beyond the cutoff: not in the matrix
    for (int e = a.rowptr[sk]; e < ck; e++)
        if (a.col[e] == si) { ek = e; break; }
    double xr = a.sx[sk] - a.sx[si], yr = a.sy[sk] - a.sy[si], zr = a.sz[sk] - a.sz[si];
    image<double>(a.box, xr, yr, zr);
    const double r2 = xr * xr + yr * yr + zr * zr;
    double lam[6], R[6];
    if (a.cpd) overlap_damp<double>(sqrt(r2), a.scp[np + si], a.scp[np + sk], lam);
    else thole_lambda<double>(sqrt(r2), a.spol[2 * np + si], a.spol[2 * np + sk], a.spol[np + si], a.spol[np + sk], lam);
    corr_series<double>(r2, q.u, lam, R);
    add_T(a.ent, (size_t)ei, R);
    if (ek >= 0) add_T(a.ent, (size_t)ek, R);
    // the preconditioner's tensor of a near pair gets the same correction
    if (a.rownear) {
        if (ei < a.rowptr[si] + a.rownear[si]) add_T(a.ent_s, (size_t)ei, R);
        if (ek >= 0 && ek < a.rowptr[sk] + a.rownear[sk]) add_T(a.ent_s, (size_t)ek, R);
    }
}

// dst = T src for both dipole sets, or with EPI the PCG step E1 on T z
// (pair_matrix.h)
template <int G, int BK, class V, bool EPI>
__global__ __launch_bounds__(256) void k_pol_spmv(SpmvArgs a)
{
    if (a.check && a.pcg.iter >= 2 && blockIdx.x == 0) pcg_converged(a.pcg, a.pcg.iter);
    spmv_rows<G, BK, V, EPI>(a, blockIdx.x);
}

// The PCG preconditioner P x = udiag a x + max(a, 1e-8) T_s (a x) over the
// near part of every row, both dipole sets, G lanes per row, applied to
// x = r (the initial residual: z = P r) or x = v (an iteration: the step E2,
// z -= step P v; pcg_steps.h).  The columns' a x come packed with their
// positions (pcg.pkw); the row's own a x is formed here.  With E set: the
// solve's first guess instead, u = a E + max(a, 1e-8) T_s (a E) (zero where
// a = 0), into u and the packed source of the operator (pcg.pks).
template <class V>
struct PrecArgs {
    const int *rowptr, *rownear;
    size_t cap;
    const TEnt<V> *ent_s;
    const int *col;
    RowGeom geo;
    PcgVec pcg;
    const double *E; // the first guess from this field (or null: the preconditioner step)
    // The solve's reciprocal potential grids carried by linearity (PME):
    // the workgroups past nrow fold the grid of this iteration's z (qg, the
    // field gather's input, two sets) into those of the direction and the
    // dipoles, gp = qg + beta gp and gu += step gp (at iteration 0, gu = qg:
    // the grid of the first guess), the same steps E1 and E2 apply to p and
    // u; at the end gu is the potential of the solved dipoles, with no
    // spreading or transform of its own.  qg null: nothing carried.
    const void *qg;
    void *gp, *gu;
    size_t gpts; // points of the two sets together
    int g32;     // float grids
    int nrow;    // workgroups of the rows
};

template <class GT>
__device__ __forceinline__ void carry_grids(const PcgVec &pc, const void *qg_, void *gp_, void *gu_, size_t gpts,
                                            int wg, int nwg)
{
    const GT *qg = (const GT *)qg_;
    GT *gp = (GT *)gp_, *gu = (GT *)gu_;
    const size_t half = gpts / 2;
    const size_t first = (size_t)wg * PCG_TB + threadIdx.x, stride = (size_t)nwg * PCG_TB;
    if (pc.iter == 0) {
        for (size_t i = first; i < gpts; i += stride) gu[i] = qg[i];
        return;
    }
    double step[2];
    pcg_e2_begin(pc, step);
    const double beta[2] = {pc.sc[PCG_BETA], pc.sc[PCG_BETA + 1]};
    for (size_t i = first; i < gpts; i += stride) {
        const int set = i < half ? 0 : 1;
        const GT p = pc.iter == 1 ? qg[i] : GT((double)qg[i] + beta[set] * (double)gp[i]);
        gp[i] = p;
        gu[i] = GT((double)gu[i] + step[set] * (double)p);
    }
}

template <int G, int BK, class V>
__global__ __launch_bounds__(PCG_TB) void k_pcg_prec(PrecArgs<V> a)
{
    const PcgVec &pc = a.pcg;
    if (a.qg && (int)blockIdx.x >= a.nrow) {
        if (pc.iter > 0 && pc.sc[PCG_DONE] != 0.0) return;
        if (a.g32) carry_grids<float>(pc, a.qg, a.gp, a.gu, a.gpts, (int)blockIdx.x - a.nrow, (int)gridDim.x - a.nrow);
        else carry_grids<double>(pc, a.qg, a.gp, a.gu, a.gpts, (int)blockIdx.x - a.nrow, (int)gridDim.x - a.nrow);
        return;
    }
    if (a.E) { // the first guess (no PCG step)
        const int tid = blockIdx.x * blockDim.x + threadIdx.x;
        const int row = tid / G, g = tid % G;
        const int np = pc.npad;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        if (row < np) {
            const int e1 = (int)min((size_t)(a.rowptr[row] + a.rownear[row]), a.cap);
            row_apply<G, PREC_U, BK, V>(a.col, a.ent_s, a.rowptr[row], e1, g, a.geo, a.geo.sx[row], a.geo.sy[row],
                                        a.geo.sz[row], pc.pkw, acc);
        }
#pragma unroll
        for (int c = 0; c < 6; c++)
#pragma unroll
            for (int off = G / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
        if (row < np && g < 6) {
            double v = acc[0];
#pragma unroll
            for (int c = 1; c < 6; c++)
                if (g == c) v = acc[c];
            const size_t i = (size_t)g * np + row;
            const double pol = pc.spol[row];
            const double u = pol != 0.0 ? pol * a.E[i] + fmax(1e-8, pol) * v : 0.0;
            pc.u[i] = u;
            pc.pks[(size_t)PMQ_PK * row + 3 + g] = u;
        }
        return;
    }
    if (pc.iter > 0 && pc.sc[PCG_DONE] != 0.0) return;
    double step[2] = {0.0, 0.0};
    if (pc.iter > 0) pcg_e2_begin(pc, step);
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = tid / G, g = tid % G;
    const int np = pc.npad;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (row < np) {
        const int e1 = (int)min((size_t)(a.rowptr[row] + a.rownear[row]), a.cap);
        row_apply<G, PREC_U, BK, V>(a.col, a.ent_s, a.rowptr[row], e1, g, a.geo, a.geo.sx[row], a.geo.sy[row],
                                    a.geo.sz[row], pc.pkw, acc);
    }
#pragma unroll
    for (int c = 0; c < 6; c++)
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
    double rz[2] = {0, 0}, rr[2] = {0, 0};
    if (row < np && g < 6) {
        double v = acc[0];
#pragma unroll
        for (int c = 1; c < 6; c++)
            if (g == c) v = acc[c];
        const size_t i = (size_t)g * np + row;
        const double pol = pc.spol[row];
        const double x = pc.iter == 0 ? pc.r[i] : pc.v[i];
        const double q = pc.udiag * (pol * x) + fmax(1e-8, pol) * v;
        pcg_e2(pc, row, g, q, step, rz[g / 3], rr[g / 3]);
    }
    pcg_e2_end(pc, rz[0], rz[1], rr[0], rr[1]);
}

// ---------------------------------------------------------------------------
// host side

template <class T, class V>
static int pair_fill(pmq_ctx *c, int tgrid)
{
    const int np = c->npad;
    hipStream_t st = c->stream;
    FillArgs<V> fa;
    fa.tiles = c->tiles.p; fa.ntiles = c->ntiles_eff.p; fa.items = tile_items(c);
    fa.sx = c->sx.p; fa.sy = c->sy.p; fa.sz = c->sz.p; fa.spol = c->spol.p;
    fa.scp = c->scp.p;
    fa.box = c->box;
    fa.aewald = c->prm.use_ewald ? c->prm.aewald : 0.0;
    fa.mask = c->tmask.p; fa.soff = c->soff.p; fa.rowptr = c->rowptr.p;
    fa.npad = np; fa.cap = c->tcap; fa.ent = (TEnt<V> *)c->tval.p; fa.col = c->tcol.p;
    fa.nmask = c->prec_on ? c->nmask.p : nullptr;
    fa.nsoff = c->nsoff.p; fa.rownear = c->rownear.p;
    fa.ent_s = c->prec_on ? (TEnt<V> *)c->tval_s.p : nullptr;
    fa.st = c->status.p;
    {
        ProfScope ps_(c, "pol_pair_fill");
        const bool cpd = c->prm.pentype != PMQ_PEN_NONE;
        if (fa.aewald > 0) {
            if (cpd) hipLaunchKernelGGL((k_pair_fill<true, T, V, true>), dim3(tgrid), dim3(64), 0, st, fa);
            else hipLaunchKernelGGL((k_pair_fill<true, T, V, false>), dim3(tgrid), dim3(64), 0, st, fa);
        } else {
            if (cpd) hipLaunchKernelGGL((k_pair_fill<false, T, V, true>), dim3(tgrid), dim3(64), 0, st, fa);
            else hipLaunchKernelGGL((k_pair_fill<false, T, V, false>), dim3(tgrid), dim3(64), 0, st, fa);
        }
    }
    const int nexcl = (int)c->h_excl.size();
    if (nexcl) {
        FoldArgs<V> fo;
        fo.excl = c->excl.p; fo.nexcl = nexcl; fo.iperm = c->iperm.p;
        fo.sx = c->sx.p; fo.sy = c->sy.p; fo.sz = c->sz.p; fo.spol = c->spol.p;
        fo.scp = c->scp.p;
        fo.cpd = c->prm.pentype != PMQ_PEN_NONE ? 1 : 0;
        fo.box = c->box; fo.npad = np; fo.rowptr = c->rowptr.p; fo.cap = c->tcap;
        fo.ent = (TEnt<V> *)c->tval.p;
        fo.col = c->tcol.p;
        fo.rownear = c->prec_on ? c->rownear.p : nullptr;
        fo.ent_s = c->prec_on ? (TEnt<V> *)c->tval_s.p : nullptr;
        hipLaunchKernelGGL((k_excl_fold<V>), dim3((nexcl + 63) / 64), dim3(64), 0, st, fo);
    }
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

// the scaled pairs cleared from the tile masks (tile_pairs.h), one thread each
__global__ void k_xmask_clear(XclearArgs x) { xmask_clear_one(x, blockIdx.x * blockDim.x + threadIdx.x); }

int tile_structure(pmq_ctx *c, double off2, bool safe, bool clear_scaled)
{
    const int nblk = c->nblk;
    hipStream_t st = c->stream;
    if (safe || c->tile_cap == 0) {
        // size the tile buffers from this evaluation's count (one round trip)
        int ntiles = 0;
        PMQ_HIP(hipMemcpyAsync(&ntiles, c->ntiles_d.p, sizeof(int), hipMemcpyDeviceToHost, st));
        PMQ_HIP(hipStreamSynchronize(st));
        if (ntiles > c->tile_cap || c->tile_cap == 0)
            c->tile_cap = std::min(std::max(1, ntiles + ntiles / 4 + 16), std::max(1, c->ntiles_cap));
    }
    const size_t nt = (size_t)c->tile_cap;
    PMQ_HIP(c->tmask.alloc(128 * nt));
    PMQ_HIP(c->mmask.alloc(64 * nt));
    PMQ_HIP(c->xmask.alloc(128 * nt));
    PMQ_HIP(c->soff.alloc(128 * nt));
    if (c->prec_on) {
        PMQ_HIP(c->nmask.alloc(128 * nt));
        PMQ_HIP(c->nsoff.alloc(128 * nt));
    }
    PMQ_HIP(c->adj.alloc((size_t)nblk * nblk));
    PMQ_HIP(c->adjn.alloc(nblk));
    PMQ_HIP(c->tpart.alloc(PMQ_TILE_PIECES * nt * 2 * PMQ_TILE_NV * 64));
    PMQ_HIP(c->tscal.alloc(PMQ_TILE_PIECES * nt * PMQ_TILE_NS));
    PMQ_HIP(c->tpc.alloc(nt));
    PMQ_HIP(c->xlist.alloc((PMQ_TILE_PIECES - 1) * nt));
    c->tgrid = (int)std::max<size_t>(1, std::min<size_t>(nt, 65536));
    c->igrid = (int)std::max<size_t>(1, std::min<size_t>(nt + nt / 2, 262144));

    MaskArgs ma;
    ma.tiles = c->tiles.p; ma.ntiles = c->ntiles_d.p; ma.cap = c->tile_cap; ma.ntiles_eff = c->ntiles_eff.p;
    ma.st = c->status.p;
    ma.sx = c->sx.p; ma.sy = c->sy.p; ma.sz = c->sz.p; ma.sflag = c->sflag.p;
    ma.box = c->box; ma.off2 = off2; ma.mask = c->tmask.p; ma.mmask = c->mmask.p;
    ma.us2 = c->prec_on ? std::min(c->prm.usolvcut * c->prm.usolvcut, off2) : -1.0;
    ma.nmask = c->prec_on ? c->nmask.p : nullptr;
    ma.blk = c->blk.p;
    ma.xmask = c->xmask.p;
    ma.tpc = c->tpc.p;
    ma.xlist = c->xlist.p;
    ma.xcnt = c->ntiles_d.p + 1;
    ma.tscal = c->tscal.p;
    ma.tstride = item_stride(c);
    const AdjArgs q{nblk, c->tile_cap, c->tileflag.p, c->tilescan.p, c->adj.p, c->adjn.p};
    if (c->box.kind == PMQ_BOX_NONE || c->box.kind == PMQ_BOX_ORTHO)
        hipLaunchKernelGGL(k_tile_masks<0>, dim3(nblk + c->tgrid), dim3(64), 0, st, ma, q);
    else
        hipLaunchKernelGGL(k_tile_masks<1>, dim3(nblk + c->tgrid), dim3(64), 0, st, ma, q);
    const int nexcl = (int)c->h_excl.size();
    if (nexcl && clear_scaled)
        hipLaunchKernelGGL(k_xmask_clear, dim3((nexcl + 255) / 256), dim3(256), 0, st, xclear_args(c));
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int pair_matrix_build(pmq_ctx *c, bool mixed, bool safe)
{
    const int nblk = c->nblk, np = c->npad;
    hipStream_t st = c->stream;
    PMQ_HIP(c->rowcnt.alloc(np + 1));
    PMQ_HIP(c->rowptr.alloc(np + 1));
    PMQ_HIP(c->rownear.alloc(np));
    if (c->prec_on)
        hipLaunchKernelGGL(k_row_counts<true>, dim3(nblk), dim3(64 * RC_W), 0, st, nblk, c->tiles.p, c->tmask.p, c->nmask.p,
                           c->adj.p, c->adjn.p, c->soff.p, c->nsoff.p, c->rowcnt.p, c->rownear.p);
    else
        hipLaunchKernelGGL(k_row_counts<false>, dim3(nblk), dim3(64 * RC_W), 0, st, nblk, c->tiles.p, c->tmask.p, nullptr,
                           c->adj.p, c->adjn.p, c->soff.p, nullptr, c->rowcnt.p, nullptr);
    PMQ_HIP(hipGetLastError());
    size_t bytes = 0;
    PMQ_HIP(rocprim::exclusive_scan(nullptr, bytes, c->rowcnt.p, c->rowptr.p, 0, (size_t)np + 1,
                                    rocprim::plus<int>(), st));
    PMQ_HIP(c->pscan_tmp.alloc(bytes));
    PMQ_HIP(rocprim::exclusive_scan(c->pscan_tmp.p, bytes, c->rowcnt.p, c->rowptr.p, 0, (size_t)np + 1,
                                    rocprim::plus<int>(), st));
    if (safe || c->tcap == 0) {
        int nnz = 0;
        PMQ_HIP(hipMemcpyAsync(&nnz, c->rowptr.p + np, sizeof(int), hipMemcpyDeviceToHost, st));
        PMQ_HIP(hipStreamSynchronize(st));
        if ((size_t)nnz > c->tcap || c->tcap == 0) {
            c->tcap = (size_t)nnz + (size_t)nnz / 4 + 64;
            PMQ_HIP(c->tval.alloc(2 * c->tcap)); // 16 bytes per entry (8 used by the mixed form)
            PMQ_HIP(c->tcol.alloc(c->tcap));
        }
    }
    // the preconditioner's tensors sit at the entries of the near pairs
    if (c->prec_on) PMQ_HIP(c->tval_s.alloc(2 * c->tcap));

    c->tmixed = mixed;
    if (mixed) return pair_fill<float, float>(c, c->igrid);
    return pair_fill<double, double>(c, c->igrid);
}

// ---------------------------------------------------------------------------
// fixed-order sums of the per-tile scalars into the totals: scalar 0 into
// energy component `which` (0 em, 1 ep, 2 er, 3 edsp, 4 ev, 5 ect), 1..9 the
// virial, 10 the inter-molecular energy and 11 the pair count; the first
// pieces of all tiles (a tile without pairs has zeros there from the mask
// kernel), each tile's further pieces after its first.  One workgroup of TB
// threads (red: TB doubles of LDS).
constexpr int SCAL_TB = 1024;

template <int TB>
__device__ __forceinline__ void scalar_sums(int ntiles, const int *__restrict__ tpc, int which, int k0, int k1,
                                            const double *__restrict__ tscal, size_t stride, Totals *tot, double *red)
{
    for (int k = k0; k < k1; k++) {
        // each thread's tiles in order, a tile's pieces together (item p
        // ntiles + t is piece p of tile t)
        double v = 0.0;
        const double *S = tscal + (size_t)k * stride;
#pragma unroll 8
        for (int t = threadIdx.x; t < ntiles; t += TB) {
            const int np = tpc[t] & 15;
            v += S[t];
            for (int p = 1; p < np; p++) v += S[(size_t)p * ntiles + t];
        }
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = TB / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double r = red[0];
            if (k == 0) atomicAdd(&(&tot->em)[which], r);
            else if (k <= 9) atomicAdd(&tot->vir[k - 1], r);
            else if (k == 10) atomicAdd(&tot->einter, r);
            else atomicAdd(&tot->nem, (unsigned long long)__double2ll_rn(r));
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(SCAL_TB) void k_tile_scalars(const int *ntiles_p, const int *__restrict__ tpc, int which,
                                                          int k0, int k1, const double *__restrict__ tscal,
                                                          size_t stride, Totals *tot)
{
    __shared__ double red[SCAL_TB];
    scalar_sums<SCAL_TB>(*ntiles_p, tpc, which, k0, k1, tscal, stride, tot, red);
}

// ---------------------------------------------------------------------------
// deterministic reduction of tile partials

// per-atom sums of NV partials over the tiles of the atom's block; P is the
// partial type (the pair arithmetic), FX sums in 64-bit fixed point.  One
// workgroup of RED_W waves per block: the block's tile list is staged in LDS,
// wave w sums tiles w, w+RED_W, ... (independent loads, several in flight),
// and the RED_W partial sums are added in wave order.
constexpr int RED_MAXT = 1024;                     // LDS entries (tile pieces) per round
constexpr int RED_TPR = RED_MAXT / PMQ_TILE_PIECES; // tiles per round
constexpr int RED_W = 8;                            // waves per block

template <class P, int NV, bool FX>
__global__ __launch_bounds__(64 * RED_W) void k_tile_reduce(TileReduceArgs a)
{
    if ((int)blockIdx.x == a.nblk) { // the scalar sums, when asked for
        __shared__ double sred[64 * RED_W];
        scalar_sums<64 * RED_W>(*a.ntiles, a.tpc, a.scal_which, a.scal_k0, a.scal_k1, a.tscal, a.tstride, a.tot, sred);
        return;
    }
    __shared__ int lt[RED_MAXT];
    __shared__ int nent;
    __shared__ unsigned long long red[RED_W][NV][64]; // fixed-point sums, or the bits of double sums
    const int I = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int si = I * 64 + lane;
    const P *part = (const P *)a.part;
    double acc[NV];
    long long q[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) { acc[c] = 0.0; q[c] = 0; }
    bool want[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) want[c] = a.out[c] != nullptr;
    const int na = a.adjn[I];
    const int ntiles = *a.ntiles;
    for (int base = 0; base < na; base += RED_TPR) {
        const int cnt = min(RED_TPR, na - base);
        __syncthreads();
        // one entry per piece of each tile, tiles in list order and a tile's
        // pieces together (tiles without pairs have none): the partial slot
        // (item * 2 + side; piece p of tile t is item p ntiles + t), bit 30
        // set for a diagonal tile (both sides).  Wave 0 lays them out with a
        // prefix sum of the piece counts.
        if (wv == 0) {
            int m = 0;
            for (int c0 = 0; c0 < cnt; c0 += 64) {
                const int s = c0 + lane;
                int np = 0, slot0 = 0, two = 0;
                if (s < cnt) {
                    const int t = a.adj[(size_t)I * a.nblk + base + s];
                    const int2 IJ = a.tiles[t];
                    np = a.tpc[t] & 15;
                    two = IJ.x == IJ.y;
                    slot0 = t * 2 + (two || IJ.x == I ? 0 : 1);
                }
                int inc = np;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int y = __shfl_up(inc, o, 64);
                    if (lane >= o) inc += y;
                }
                const int off = m + inc - np;
                for (int p = 0; p < np; p++) lt[off + p] = (slot0 + 2 * p * ntiles) | (two << 30);
                m += __shfl(inc, 63, 64);
            }
            if (lane == 0) nent = m;
        }
        __syncthreads();
        const int total = nent;
        // loads are unconditional (a diagonal tile's second side is read from
        // the next slot, any other tile's "second side" is its own slot again
        // and dropped), so the loads of several tiles are in flight together;
        // components nobody wants are read and never stored
        constexpr int RU = sizeof(P) == 4 ? 4 : 2; // tiles in flight per wave
#pragma unroll RU
        for (int s = wv; s < total; s += RED_W) {
            const int e = lt[s];
            const int slot = e & ~(1 << 30);
            const bool two = (e >> 30) & 1;
            const P *pp = part + (size_t)slot * PMQ_TILE_NV * 64 + lane;
            const P *pq = two ? pp + PMQ_TILE_NV * 64 : pp;
            P v[NV], w[NV];
#pragma unroll
            for (int c = 0; c < NV; c++) {
                v[c] = pp[c * 64];
                w[c] = pq[c * 64];
            }
#pragma unroll
            for (int c = 0; c < NV; c++) {
                if (FX) {
                    q[c] += __double2ll_rn((double)v[c] * a.fxscale[c]);
                    if (two) q[c] += __double2ll_rn((double)w[c] * a.fxscale[c]);
                } else {
                    acc[c] += (double)v[c];
                    if (two) acc[c] += (double)w[c];
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NV; c++)
        red[wv][c][lane] = FX ? (unsigned long long)q[c] : (unsigned long long)__double_as_longlong(acc[c]);
    __syncthreads();
    if (wv != 0 || si >= a.n) return;
#pragma unroll
    for (int c = 0; c < NV; c++) {
        if (!want[c]) continue;
        const size_t idx = a.caller_order[c] ? (size_t)a.stride[c] * a.perm[si] + a.off[c] : (size_t)a.stride[c] * si + a.off[c];
        double v, v2;
        if (FX) {
            long long sq = 0;
            for (int w = 0; w < RED_W; w++) sq += (long long)red[w][c][lane];
            long long s1 = sq, s2 = sq;
            if (a.fxq[c]) { s1 += a.fxq[c][idx]; a.fxq[c][idx] = 0; }
            if (a.fxq2[c]) { s2 += a.fxq2[c][idx]; a.fxq2[c][idx] = 0; }
            v = (double)s1 / a.fxscale[c];
            v2 = (double)s2 / a.fxscale[c];
        } else {
            v = 0.0;
            for (int w = 0; w < RED_W; w++) v += __longlong_as_double((long long)red[w][c][lane]);
            v2 = v;
        }
        a.out[c][idx] += v;
        if (a.out2[c]) a.out2[c][idx] += v2;
    }
}

int tile_reduce(pmq_ctx *c, TileReduceArgs &a, int nv, bool float_parts, bool fixed, int scal_which, int k0, int k1)
{
    a.nblk = c->nblk; a.n = c->n;
    a.tiles = c->tiles.p; a.adj = c->adj.p; a.adjn = c->adjn.p; a.perm = c->perm.p;
    a.tpc = c->tpc.p; a.ntiles = c->ntiles_eff.p;
    a.part = c->tpart.p;
    a.scal_which = scal_which; a.scal_k0 = k0; a.scal_k1 = k1;
    a.tscal = c->tscal.p; a.tstride = item_stride(c); a.tot = c->totals.p;
    if (!fixed)
        for (int q = 0; q < 8; q++)
            if (a.fxq[q] || a.fxq2[q]) {
                c->err = "tile_reduce: fixed-point addends need the fixed-point sums";
                return PMQ_ERR_STATE;
            }
    const dim3 g(c->nblk + (k1 > k0 ? 1 : 0)), b(64 * RED_W);
#define PMQ_RED(P, NV, FX) hipLaunchKernelGGL((k_tile_reduce<P, NV, FX>), g, b, 0, c->stream, a)
    if (float_parts) {
        if (nv == 3) { if (fixed) PMQ_RED(float, 3, true); else PMQ_RED(float, 3, false); }
        else if (nv == 6) { if (fixed) PMQ_RED(float, 6, true); else PMQ_RED(float, 6, false); }
        else { if (fixed) PMQ_RED(float, 8, true); else PMQ_RED(float, 8, false); }
    } else {
        if (nv == 3) PMQ_RED(double, 3, false);
        else if (nv == 6) PMQ_RED(double, 6, false);
        else PMQ_RED(double, 8, false);
    }
#undef PMQ_RED
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int tile_scalars(pmq_ctx *c, int which, bool full)
{
    hipLaunchKernelGGL(k_tile_scalars, dim3(1), dim3(SCAL_TB), 0, c->stream, c->ntiles_eff.p, c->tpc.p, which, 0,
                       full ? PMQ_TILE_NS : 1, c->tscal.p, item_stride(c), c->totals.p);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int tile_virial(pmq_ctx *c)
{
    hipLaunchKernelGGL(k_tile_scalars, dim3(1), dim3(SCAL_TB), 0, c->stream, c->ntiles_eff.p, c->tpc.p, 0, 1, 10,
                       c->tscal.p, item_stride(c), c->totals.p);
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

static RowGeom row_geom(const pmq_ctx *c)
{
    RowGeom g;
    g.sx = c->sx.p; g.sy = c->sy.p; g.sz = c->sz.p;
    g.box = c->box;
    const double len[3] = {c->box.xbox, c->box.ybox, c->box.zbox};
    for (int q = 0; q < 3; q++) {
        g.len[q] = len[q];
        g.inv[q] = len[q] > 0.0 ? 1.0 / len[q] : 0.0;
    }
    return g;
}

static int row_box(const pmq_ctx *c)
{
    if (c->box.kind == PMQ_BOX_NONE) return ROW_BOX_NONE;
    if (c->box.kind == PMQ_BOX_ORTHO) return ROW_BOX_ORTHO;
    return ROW_BOX_ANY;
}

SpmvArgs spmv_args(pmq_ctx *c, double *dst, const double *skip, const PcgVec *pcg, int check)
{
    SpmvArgs a;
    a.npad = c->npad; a.rowptr = c->rowptr.p; a.cap = c->tcap; a.ent = c->tval.p;
    a.col = c->tcol.p; a.geo = row_geom(c);
    a.pk = c->pk_src.p; a.dst = dst; a.skip = skip;
    a.pcg = pcg ? *pcg : PcgVec{};
    a.check = check;
    return a;
}

int spmv_box(const pmq_ctx *c) { return row_box(c); }

template <class V>
static int spmv(pmq_ctx *c, double *dst, const double *skip, PcgVec *pcg, int check, bool epi)
{
    constexpr int G = SPMV_G;
    const size_t lanes = (size_t)c->npad * G;
    const dim3 grid((unsigned)((lanes + 255) / 256));
    if (epi) {
        if ((size_t)grid.x * 2 > c->pcg_part1_cap) {
            c->err = "PCG: partial-sum buffer too small for the row product";
            return PMQ_ERR_STATE;
        }
        pcg->n1 = (int)grid.x;
    }
    const SpmvArgs a = spmv_args(c, dst, skip, pcg, check);
    ProfScope ps_(c, "pol_mutual_spmv");
#define PMQ_SPMV(BK, EPI) hipLaunchKernelGGL((k_pol_spmv<G, BK, V, EPI>), grid, dim3(256), 0, c->stream, a)
#define PMQ_SPMV_BOX(EPI)                                                                                      \
    do {                                                                                                       \
        switch (row_box(c)) {                                                                                  \
        case ROW_BOX_NONE: PMQ_SPMV(ROW_BOX_NONE, EPI); break;                                               \
        case ROW_BOX_ORTHO: PMQ_SPMV(ROW_BOX_ORTHO, EPI); break;                                             \
        default: PMQ_SPMV(ROW_BOX_ANY, EPI); break;                                                          \
        }                                                                                                      \
    } while (0)
    if (epi) PMQ_SPMV_BOX(true);
    else PMQ_SPMV_BOX(false);
#undef PMQ_SPMV_BOX
#undef PMQ_SPMV
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

template <class V>
static int precondition(pmq_ctx *c, const PcgVec &pcg, const double *E, bool carry)
{
    constexpr int G = PREC_G;
    PrecArgs<V> a;
    a.rowptr = c->rowptr.p; a.rownear = c->rownear.p; a.cap = c->tcap;
    a.ent_s = (const TEnt<V> *)c->tval_s.p;
    a.col = c->tcol.p; a.geo = row_geom(c);
    a.pcg = pcg;
    a.E = E;
    a.nrow = pair_matrix_prec_blocks(c);
    if (a.nrow != pcg.n2) {
        c->err = "PCG: the preconditioner's partial count";
        return PMQ_ERR_STATE;
    }
    void *qg = nullptr, *gp = nullptr, *gu = nullptr;
    size_t gpts = 0;
    bool g32 = false;
    int ncarry = 0;
    if (carry && !E && pme_carry_grids(c, &qg, &gp, &gu, &gpts, &g32))
        ncarry = (int)std::min<size_t>((gpts + 4 * PCG_TB - 1) / (4 * PCG_TB), 4096);
    a.qg = ncarry ? qg : nullptr;
    a.gp = gp; a.gu = gu; a.gpts = gpts; a.g32 = g32 ? 1 : 0;
    const dim3 grid((unsigned)(a.nrow + ncarry));
    switch (row_box(c)) {
    case ROW_BOX_NONE: hipLaunchKernelGGL((k_pcg_prec<G, ROW_BOX_NONE, V>), grid, dim3(PCG_TB), 0, c->stream, a); break;
    case ROW_BOX_ORTHO: hipLaunchKernelGGL((k_pcg_prec<G, ROW_BOX_ORTHO, V>), grid, dim3(PCG_TB), 0, c->stream, a); break;
    default: hipLaunchKernelGGL((k_pcg_prec<G, ROW_BOX_ANY, V>), grid, dim3(PCG_TB), 0, c->stream, a); break;
    }
    PMQ_HIP(hipGetLastError());
    return PMQ_OK;
}

int pair_matrix_prec_blocks(const pmq_ctx *c) { return (int)(((size_t)c->npad * PREC_G + PCG_TB - 1) / PCG_TB); }

int pair_matrix_precondition(pmq_ctx *c, const PcgVec &pcg, const double *E, bool carry)
{
    if (!c->prec_on) {
        c->err = "preconditioner applied without its matrix";
        return PMQ_ERR_STATE;
    }
    if (c->tmixed) return precondition<float>(c, pcg, E, carry);
    return precondition<double>(c, pcg, E, carry);
}

int pair_matrix_apply(pmq_ctx *c, double *dst, const double *skip, PcgVec *pcg, int mode)
{
    const int check = pcg && mode != PCG_APPLY_NONE;
    const bool epi = pcg && mode == PCG_APPLY_E1;
    if (c->tmixed) return spmv<float>(c, dst, skip, pcg, check, epi);
    return spmv<double>(c, dst, skip, pcg, check, epi);
}

} // namespace pmq
