// tile_pairs.h -- compacted pair schedule of one 64x64 tile (one wavefront).
//
// The tile masks give every lane (an atom of block I) the set of atoms of
// block J it pairs with.  Looping over each lane's own mask leaves lanes idle
// until the fullest row is done (about a fifth of the lanes do work on a
// water box).  Instead the wave hands the tile's pairs out 64 at a time, one
// pair per lane, in rank order: the first pair of every row, then the
// second, ..., and each row hands out its partners starting at its own lane
// position, so the 64 pairs of a batch mostly have distinct i and distinct j
// (few same-address LDS atomics when the batch accumulates into both atoms).
// The order depends only on the masks, and one wave issues every LDS atomic,
// so sums are reproducible run to run.
//
// Two forms of the same schedule.  tile_pair_list writes the whole list
// (up to 4096 entries, 8 KB of LDS) before the first batch.
// TilePairStream produces the pairs one rank at a time (at most 64 entries)
// into a ring of 128 LDS entries and hands out a batch whenever 64 are
// waiting, or at the end: 256 bytes of LDS, so more tile waves share a CU.
// Batches are consecutive runs of the same rank order either way, so the
// sums do not depend on the form.  Measured on a 12,288-atom water box: the
// ring helps the kernels with heavy pair bodies and large LDS footprints
// (multipole energy and polarization gradient tiles, 2-3% each) and slows
// the light ones (pair-matrix fill +23%, permanent field +5%: the list
// generation interleaved with the batches costs more than the occupancy
// gains), so those keep the whole list.
#pragma once

#include <hip/hip_runtime.h>

#include "pmq_internal.h"

namespace pmq {

// rank-ordered pair list of mask rows m (lane = row); entries row << 8 | col;
// returns the number of pairs (<= 4096)
__device__ __forceinline__ int tile_pair_list(unsigned long long m, unsigned short *plist, int lane)
{
    const unsigned long long below = (1ull << lane) - 1ull;
    int base = 0;
    while (true) {
        const bool has = m != 0ull;
        const unsigned long long b = __ballot(has);
        if (b == 0ull) break;
        if (has) {
            // first set bit at or after position `lane`, cyclically
            const unsigned long long rot = lane ? ((m >> lane) | (m << (64 - lane))) : m;
            const int jj = (__ffsll((long long)rot) - 1 + lane) & 63;
            m &= ~(1ull << jj);
            plist[base + __popcll(b & below)] = (unsigned short)((lane << 8) | jj);
        }
        base += __popcll(b);
    }
    return base;
}

constexpr int TILE_RING = 128; // LDS entries of the pair ring (two batches)

struct TilePairStream {
    unsigned long long m; // this lane's partners not yet produced
    int made, taken;      // pairs produced and handed out so far (wave-uniform)
    bool more;            // some row still has partners

    __device__ __forceinline__ explicit TilePairStream(unsigned long long mask)
        : m(mask), made(0), taken(0), more(true) {}

    // The next batch: returns its size (0 when the tile is done; below 64
    // only for the last batch) and sets ent to this lane's entry, row << 8 |
    // col, valid when lane < the size.  Called by the whole wave, the only
    // wave of its workgroup (one-wave tile kernels).
    __device__ __forceinline__ int next(unsigned short *ring, int lane, int &ent)
    {
        const unsigned long long below = (1ull << lane) - 1ull;
        while (more && made - taken < 64) {
            const bool has = m != 0ull;
            const unsigned long long b = __ballot(has);
            if (b == 0ull) {
                more = false;
                break;
            }
            if (has) {
                // first set bit at or after position `lane`, cyclically
                const unsigned long long rot = lane ? ((m >> lane) | (m << (64 - lane))) : m;
                const int jj = (__ffsll((long long)rot) - 1 + lane) & 63;
                m &= ~(1ull << jj);
                ring[(made + __popcll(b & below)) & (TILE_RING - 1)] = (unsigned short)((lane << 8) | jj);
            }
            made += __popcll(b);
        }
        const int cnt = min(64, made - taken);
        __syncthreads();
        ent = ring[(taken + lane) & (TILE_RING - 1)];
        taken += cnt;
        return cnt;
    }
};

// Work items of the pair kernels: pieces of tiles.  A tile's piece count np
// and batches per piece bpp come from its in-cutoff pair count (k_tile_masks:
// np = 0 for a tile without pairs, at most PMQ_TILE_PIECES); piece p takes
// batches [p bpp, (p + 1) bpp) of the tile's rank-ordered schedule, whatever
// subset of those pairs the kernel's own mask keeps.  Items [0, ntiles) are
// the first pieces of the tiles in slot order (diagonal tiles leading), then
// come the further pieces of the heavy tiles, listed by k_tile_masks
// (xlist, *xcnt entries of tile | piece << 28, in no fixed order: each piece
// writes only its own partial slots, so the order of the items does not
// reach the sums).  Piece p of tile t writes partial slot / scalar index
// p ntiles + t.  One wave per tile made the launch wait on the densest tiles
// (a diagonal tile holds about 1,200 pairs, the mean tile about 190), each
// one wave's serial chain of batches while the other SIMDs had drained;
// pieces of at most about six batches bound that chain.  That tail matters
// only while a launch has few tiles per SIMD: above PMQ_SPLIT_TILES tiles
// no tile is split (measured: 4,503 tiles of a 12,288-atom box, the tile
// kernels 20-35% faster split; 36,000 tiles of a 96,000-atom box, no gain,
// and the extra partial slots cost the reduction 10%).
struct TileItems {
    const int *tpc;   // per tile: np | bpp << 4
    const int *xlist; // further pieces
    const int *xcnt;  // their count
};

struct TileItem {
    int t, p;   // tile and piece
    int u;      // partial slot / scalar index: p ntiles + t
    int b0, b1; // the piece's batches of the tile's schedule
};

__device__ __forceinline__ bool tile_item(int u, int ntiles, const TileItems &ti, TileItem &it)
{
    if (u < ntiles) {
        it.t = u;
        it.p = 0;
    } else {
        const int e = ti.xlist[u - ntiles];
        it.t = e & ((1 << 28) - 1);
        it.p = e >> 28;
    }
    const int code = ti.tpc[it.t];
    const int np = code & 15, bpp = code >> 4;
    if (it.p >= np) return false;
    it.u = it.p * ntiles + it.t;
    it.b0 = it.p * bpp;
    it.b1 = it.b0 + bpp;
    return true;
}

// Rows of the tile masks without the scaled pairs (the 1-2..1-5 and
// polarization-group pairs): [0,64) the in-cutoff rows, [64,128) the rows
// with an active atom, written by k_tile_masks and cleared of the scaled
// pairs after it (one thread per scaled pair).  The pair kernels skip those
// pairs and the exclusion kernels evaluate them whole, at their scales, in
// FP64: large intramolecular pair terms are then never formed in one
// precision and cancelled in another.
struct XclearArgs {
    int nexcl;
    const ExclPair *excl;
    const int *iperm;
    const int2 *tiles;
    const int *adj, *adjn; // each block's tiles, ordered by the other block
    int nblk;
    const int *ntiles;
    unsigned long long *xmask;
};

__device__ __forceinline__ void xmask_clear_one(const XclearArgs &x, int q)
{
    if (q >= x.nexcl) return;
    const int s1 = x.iperm[x.excl[q].i], s2 = x.iperm[x.excl[q].k];
    const int a = min(s1, s2), b = max(s1, s2);
    const int I = a >> 6, J = b >> 6;
    // the tile (I, J) in I's list, which is ordered by the other block
    const int *lst = x.adj + (size_t)I * x.nblk;
    int lo = 0, hi = x.adjn[I];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int2 IJ = x.tiles[lst[mid]];
        const int other = IJ.x == I ? IJ.y : IJ.x;
        if (other < J) lo = mid + 1; else hi = mid;
    }
    if (lo >= x.adjn[I]) return;
    const int t = lst[lo];
    const int2 IJ = x.tiles[t];
    if ((IJ.x == I ? IJ.y : IJ.x) != J || t >= *x.ntiles) return; // beyond the cutoff: no tile holds it
    // row: the atom of block IJ.x (the smaller one on a diagonal tile)
    const int row = IJ.x == I ? (a & 63) : (b & 63), bit = IJ.x == I ? (b & 63) : (a & 63);
    const unsigned long long clr = ~(1ull << bit);
    atomicAnd(&x.xmask[(size_t)t * 128 + row], clr);
    atomicAnd(&x.xmask[(size_t)t * 128 + 64 + row], clr);
}

inline XclearArgs xclear_args(const pmq_ctx *c)
{
    return XclearArgs{(int)c->h_excl.size(), c->excl.p, c->iperm.p, c->tiles.p, c->adj.p, c->adjn.p, c->nblk,
                      c->ntiles_eff.p, c->xmask.p};
}

// the item lists of a context's current tile structure
inline TileItems tile_items(const pmq_ctx *c) { return TileItems{c->tpc.p, c->xlist.p, c->ntiles_d.p + 1}; }

// piece code of a tile with npair pairs, at most maxp pieces
__device__ __forceinline__ int tile_piece_code(int npair, int maxp)
{
    const int nb = (npair + 63) / 64;
    if (nb == 0) return 0;
    const int np = min(maxp, (nb + PMQ_PIECE_BATCHES - 1) / PMQ_PIECE_BATCHES);
    return np | (((nb + np - 1) / np) << 4);
}

} // namespace pmq
