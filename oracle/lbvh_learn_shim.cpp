// C entry point over the reference's stand-alone LBVH walkthrough
// (learn/03_lbvh.cpp:84-169: Morton keys of the triangle centroids, a stable sort
// of code << 32 | index, Karras 2012's per-node range and split search, and its
// child numbering), compiled in place from /root/reference by `make ref` into
// _ref/liblbvh_learn.so.  Test infrastructure only: it pins the oracle's LBVH
// leaf order, child links and depth.  The walkthrough's own main() is renamed so
// that the file links as a library; nothing else of it is changed or restated.
#define main quarry_learn03_main
#include "03_lbvh.cpp"
#undef main

extern "C" {

// tris: n x 9 floats.  order: n leaf slots -> triangle index.  left/right: n - 1
// internal nodes' child ids (internal 0..n-2, leaves n-1+slot).  Returns the depth
// (the deepest node's level, root 0), or -1 if the walkthrough's check failed.
int learn_lbvh(const float* tris, int n, int* order, int* left, int* right) {
    std::vector<Tri> t(static_cast<size_t>(n));
    std::memcpy(t.data(), tris, sizeof(Tri) * static_cast<size_t>(n));
    std::vector<int> ord;
    std::vector<Node> nodes;
    int depth = 0;
    buildLBVH(t, ord, nodes, depth);
    if (depth < 0) return -1;
    for (int i = 0; i < n; ++i) order[i] = ord[i];
    for (int i = 0; i + 1 < n; ++i) {
        left[i] = nodes[i].left;
        right[i] = nodes[i].right;
    }
    return depth;
}

}  // extern "C"
