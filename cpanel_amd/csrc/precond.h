// Strip block-Jacobi preconditioner of the batched FP32 inner solve.
//
// The strongest couplings of a panel matrix run around each chordwise strip:
// upper and lower surface panels of the same section, tied together through
// the Kutta folding of the wake into their columns (geometry.cpp:854-872).
// Blocks are therefore whole strips (consecutive spanwise slabs bounded by
// the trailing-edge wake lines, tip caps joining the end strips), merged up to
// a target size.  Their inverses are applied as a block-diagonal product.
// The plain FP64 path keeps the reference's Jacobi preconditioner.
#pragma once

#include <cstdint>
#include <vector>

#include "geometry.h"

namespace cpgpu {

struct BlockPlan {
    std::vector<int> perm;     // body indices grouped by block
    std::vector<int> ptr;      // block q owns perm[ptr[q] .. ptr[q+1])
    std::vector<int64_t> off;  // offset of block q's b*b inverse in the packed array
    int max_b = 0;
    int nblocks() const { return (int)ptr.size() - 1; }
};

// Default strip-block target for n body panels: n / 10 within [768, 3072]
// (larger blocks cut the inner iterations but cost more to apply and, once
// per geometry, to invert); CPGPU_BJ_TARGET (512..3072) overrides it.
int bj_block_target(int n);
constexpr int kBjHardMax = 3072;  // rows of the largest block the inverter takes

// Groups the body panels into strip blocks of at most `target` panels (a
// single strip larger than target stays whole, up to `hard_max`).
void build_strip_blocks(const Geometry& G, int target, int hard_max, BlockPlan& plan);

}  // namespace cpgpu
