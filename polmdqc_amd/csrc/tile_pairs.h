// tile_pairs.h -- compacted pair schedule of one 64x64 tile (one wavefront).
//
// The tile masks give every lane (an atom of block I) the set of atoms of
// block J it pairs with.  Looping over each lane's own mask leaves lanes idle
// until the fullest row is done (about a fifth of the lanes do work on a
// water box).  Instead the wave writes the tile's pairs to LDS once and then
// takes them 64 at a time, one pair per lane.  The list is ordered by rank:
// the first pair of every row, then the second, ..., and each row hands out
// its partners starting at its own lane position, so the 64 pairs of a chunk
// mostly have distinct i and distinct j (few same-address LDS atomics when the
// chunk accumulates into both atoms).  The order depends only on the masks,
// and one wave issues every LDS atomic, so sums are reproducible run to run.
#pragma once

#include <hip/hip_runtime.h>

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

} // namespace pmq
