// Point location in the octree shared by the regulariser's neighbour walk and the uniform volume sampling.
#pragma once
#include "common.hpp"

namespace neat
{

// ActiveNodeIdForGlobalPosition (HyperTree.cu:600-640): from the root down through inactive nodes into the first
// child (in child order) whose closed box contains p (BoxContainsPoint, HyperTree.h:85-95); -1 when p is in no
// active node. The child boxes are not read from the node table: they halve the parent's box as the tree's
// construction does (HyperTree.cpp:94-129, child i takes the upper half on axis k iff bit k of i; exact, every box of
// the construction from the root's is dyadic). Of the children whose closed box contains p, the first in child order
// is the one that takes the lower half on every axis where p lies on the centre plane, so the child index is the
// set of axes with p above the centre. Only the root can fail to contain p; after that p stays inside the box.
__device__ __forceinline__ int32_t active_node_at(const float* pmin, const float* pmax, const uint8_t* flags,
                                                  int32_t first_leaf, const float p[3])
{
    float mn[3] = {pmin[0], pmin[1], pmin[2]};
    float mx[3] = {pmax[0], pmax[1], pmax[2]};
    int32_t n   = 0;
    for (int depth = 0; depth < 10; ++depth)
    {
        if (flags[n] & 1) return n;
        if (n >= first_leaf) return -1;  // a leaf of the table that is not active: no children
        int32_t child = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            if (!(mn[d] <= p[d] && mx[d] >= p[d])) return -1;  // outside the box: no child contains p
            const float c = (mn[d] + mx[d]) * 0.5f;
            if (p[d] > c)
            {
                child |= 1 << d;
                mn[d] = c;
            }
            else
            {
                mx[d] = c;
            }
        }
        n = 8 * n + 1 + child;
    }
    return -1;
}

}  // namespace neat
