// xcd.hpp -- XCD-aware numbering of workgroups.
//
// MI355X deals the workgroups of a launch to its 8 XCDs round-robin (block b
// to XCD b % 8), and each XCD has its own L2.  Overlapping PIV windows share
// half their pixels, so numbering the blocks such that those of one XCD take
// one contiguous run of the work keeps the shared pixels in one L2.  Both
// mappings are bijections whatever the grid size and XCD count, so they
// change only where work runs, never what is computed.
#pragma once

#include <hip/hip_runtime.h>

namespace opiv {

constexpr int kXcds = 8;

// block b of a grid of g blocks -> its position in the contiguous order
__device__ __forceinline__ int xcd_block(int b, int g)
{
    const int x = b % kXcds, k = b / kXcds;
    const int q = g / kXcds, r = g % kXcds;      // XCDs x < r take q + 1 blocks, the others q
    return x * q + min(x, r) + k;
}

// Persistent launches: the share [lo, hi) of n work items that block b's XCD
// takes (proportional to its number of blocks), and this block's rank and
// the number of blocks on that XCD, for striding through the share.
struct XcdShare {
    int lo, hi, rank, blocks;
};
__device__ __forceinline__ XcdShare xcd_share(int b, int g, int n)
{
    const int x = b % kXcds;
    const int q = g / kXcds, r = g % kXcds;
    const int start = x * q + min(x, r), count = q + (x < r ? 1 : 0);
    XcdShare s;
    s.lo = (int)((long long)n * start / g);
    s.hi = (int)((long long)n * (start + count) / g);
    s.rank = b / kXcds;
    s.blocks = count;
    return s;
}

} // namespace opiv
