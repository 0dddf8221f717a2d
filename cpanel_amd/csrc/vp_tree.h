// K4 host side: the particle octree of the reference, built in one pass and
// flattened for the GPU traversal.
//
// Tree semantics follow octree<particle> / node<particle>
// (octree/include/octree.h:39-195, node.h:30-224): a cubic root centred on the
// bounding box with half width 0.51 * max extent; a leaf splits when it holds
// more than max_members particles, each member going to the octant
// (x > ox) * 4 + (y > oy) * 2 + (z > oz); members keep insertion order.
// Inserting the particles one at a time (the reference) and partitioning the
// whole set top-down (here) produce the same tree: a node splits iff more than
// max_members particles reach it, and stable partitioning keeps each leaf's
// member order.  Pseudo-particles follow particleFMM::computeMultExp
// (particleFMM.cpp:29-58).
//
// Flattened layout: nodes in pre-order (children in octant order), so the
// first child of node i is node i + 1 and skip[i] is the node after i's
// subtree; a traversal is a single forward walk (no stack).  Members are
// stored in leaf order.
#pragma once

#include <cstdint>
#include <vector>

namespace cpgpu {

struct BhNode {           // 80 bytes, read by the traversal kernel
    double ex, ey, ez;    // pseudo-particle position
    double width;         // cell width s = 2 * half (node::isFarField)
    double sx, sy, sz;    // pseudo-particle strength
    double radius;        // pseudo-particle radius |half| * 2 * 1.5
    int32_t skip;         // next node after this subtree
    int32_t first;        // leaf: first member slot
    int32_t count;        // leaf: member count; -1: internal node
    int32_t pad;
};
static_assert(sizeof(BhNode) == 80, "BhNode layout");

struct BhTreeHost {
    std::vector<BhNode> nodes;
    std::vector<int32_t> order;       // member slot -> particle index
    std::vector<int32_t> slot;        // particle index -> member slot
    std::vector<double> mpos, mstr, mrad;  // members in slot order (3, 3, 1 per slot)
    int levels = 0;
    double origin[3] = {0, 0, 0};
    double half = 0;
};

// Builds the tree of n particles (xyz, alpha: 3 per particle, interleaved).
void build_bh_tree(int n, const double* xyz, const double* alpha, const double* radius, int max_members,
                   BhTreeHost& T);

}  // namespace cpgpu
