// Host-side construction of the search tree (plain C++, no HIP types).
//
// Layout (what the kernels read):
//   nodes: inner nodes, 64 B each, in depth-first order (left inner child = parent + 1).
//          Each node holds the tight bounding boxes of both children plus, per child, a ref
//          and a count: cnt > 0  -> leaf, points [ref, ref + cnt) of the leaf-ordered array;
//                       cnt == 0 -> inner child node `ref`;
//                       cnt < 0  -> empty slot (only in a root that is a single leaf).
//   pts:   n x float4 in leaf order, (x, y, z, bit pattern of the original int32 index).
// The split rule is the median of the longest extent of the node's tight box, with the left
// part rounded up to a multiple of the leaf size so leaves are full. Search results do not
// depend on the split rule: pruning only uses the stored child boxes, and ties are broken
// by the original index.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

struct RmNode {
    float box[2][6];   // per child: lo.x lo.y lo.z hi.x hi.y hi.z
    int32_t ref[2];
    int32_t cnt[2];
};
static_assert(sizeof(RmNode) == 64, "RmNode must be 64 bytes");

struct RmHostTree {
    int64_t n = 0;
    int leaf_max = 0;
    int depth = 0;                 // inner-node levels on the longest root-to-leaf path
    std::vector<RmNode> nodes;
    std::vector<float> pts;        // n x 4, leaf order; w holds the original index bits
    std::vector<int32_t> level;    // depth of each inner node (root = 0), for refits
};

// Build from n x 4 floats (stride 16 B). Returns false with a message on bad input.
bool rm_build_host_tree(const float* xyzw, int64_t n, int leaf_max, RmHostTree& out,
                        std::string& err);

// Shape of the tree for n points and leaf size L: it depends on the counts only (the split
// keeps the first `nl` points of the node in the order of its split axis), so the host lays out
// every inner node before any coordinate is read and the device build fills in the split axes,
// the point order and the boxes. Inner nodes are numbered breadth first.
struct RmShape {
    int64_t n = 0;
    int leaf_max = 0;
    int depth = 0;
    std::vector<RmNode> nodes;       // refs and counts set, boxes left for the device build
    std::vector<int64_t> seg_b;      // per inner node: first position of its segment
    std::vector<int64_t> seg_e;      // one past its last position
    std::vector<int32_t> nl;         // points of its left child
    std::vector<int32_t> level_start;  // inner nodes of level l: [level_start[l], level_start[l+1])
};

bool rm_tree_shape(int64_t n, int leaf_max, RmShape& out, std::string& err);

// Node records of the tree of n >= 1 points with leaf size leaf_max in [1, 64] (its inner nodes,
// or the one root record of a single leaf), without laying the tree out: subtrees of equal
// counts have equal shapes, and the splits of one level produce at most two counts.
int64_t rm_tree_node_count(int64_t n, int leaf_max);
