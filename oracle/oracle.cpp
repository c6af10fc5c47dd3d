// CPU oracle: a single-precision restatement of the reference's CPU path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker.  The product
// (quarry-gpu_amd/) never links, loads or calls it.
//
// Every arithmetic routine follows the reference's FP32 operation order
// (compiled with -ffp-contract=off, std::fma for the explicit fused ops) so
// that hit flags, t bits and squared-distance bits match the reference CPU
// implementation exactly.  Citations are to ycl-berserk/quarry-gpu.
//
// Exported C entry points (ctypes):
//   orc_ray_brute / orc_closest_brute / orc_knn_brute / orc_radius_brute
//   orc_bvh_new / orc_bvh_free / orc_bvh_ray / orc_bvh_closest / orc_bvh_knn / orc_bvh_radius
//   orc_kd_* (point kd-tree, leaf 8)
//   orc_lbvh_nodes (CPU LBVH with the GPU node layout, for tree-hash parity)
//   orc_ray_hash / orc_closest_hash / orc_list_hash (fingerprints)
//   orc_tri_closest_point, orc_gauss_dist2, orc_tri_intersect (leaf math probes)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

const float INF = std::numeric_limits<float>::infinity();

struct P3 {
    float x, y, z;
};
struct Bx {
    P3 mn, mx;
};
struct T3 {
    P3 a, b, c;
};
struct Q4 {
    float x, y, z, w;
};
struct G10 {
    P3 mean;
    Q4 rot;
    P3 scale;
};
struct R8 {
    P3 o, d;
    float tmin, tmax;
};
static_assert(sizeof(T3) == 36 && sizeof(G10) == 40 && sizeof(R8) == 32, "layouts");

inline P3 sub(const P3& a, const P3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 add(const P3& a, const P3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 mul(const P3& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot3(const P3& a, const P3& b) {
    float r = a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    return r;
}
inline P3 crs(const P3& a, const P3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float at(const P3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// AABB with compare-select extension (src/core/geometry.h:49-81).
inline Bx boxEmpty() { return {{INF, INF, INF}, {-INF, -INF, -INF}}; }
inline void boxAdd(Bx& b, const P3& p) {
    if (p.x < b.mn.x) b.mn.x = p.x;
    if (p.y < b.mn.y) b.mn.y = p.y;
    if (p.z < b.mn.z) b.mn.z = p.z;
    if (p.x > b.mx.x) b.mx.x = p.x;
    if (p.y > b.mx.y) b.mx.y = p.y;
    if (p.z > b.mx.z) b.mx.z = p.z;
}
inline bool boxOk(const Bx& b) { return b.mn.x <= b.mx.x && b.mn.y <= b.mx.y && b.mn.z <= b.mx.z; }
inline void boxUnion(Bx& b, const Bx& o) {
    if (!boxOk(o)) return;
    boxAdd(b, o.mn);
    boxAdd(b, o.mx);
}
inline P3 boxCenter(const Bx& b) { return mul(add(b.mn, b.mx), 0.5f); }

inline Bx triBox(const T3& t) {
    Bx b = boxEmpty();
    boxAdd(b, t.a);
    boxAdd(b, t.b);
    boxAdd(b, t.c);
    return b;
}

// Slab interval of a ray against a box, shared shape of intersectAABB and
// hitParamInsideBox (geometry.h:112-143, 282-316).
bool slab(const R8& r, const Bx& b, float t0, float t1, bool padded, float* outT0, float* outT1) {
    for (int ax = 0; ax < 3; ++ax) {
        const float o = at(r.o, ax), d = at(r.d, ax), lo = at(b.mn, ax), hi = at(b.mx, ax);
        if (std::fabs(d) < 1e-20f) {
            if (o < lo || o > hi) return false;
            continue;
        }
        const float inv = 1.0f / d;
        float ta = (lo - o) * inv;
        float tb = (hi - o) * inv;
        if (ta > tb) std::swap(ta, tb);
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        if (padded) {
            if (t0 > t1 * (1.0f + 1e-6f) + 1e-6f) return false;
        } else if (t0 > t1) {
            return false;
        }
    }
    *outT0 = t0;
    *outT1 = t1;
    return true;
}

bool rayHitsBox(const R8& r, const Bx& b, float tMax, float* tEnter) {
    float t1;
    return slab(r, b, r.tmin, tMax * (1.0f + 1e-6f) + 1e-6f, true, tEnter, &t1);
}

bool tWithinTriBox(const R8& r, const Bx& b, float t) {
    float t0, t1;
    if (!slab(r, b, r.tmin, INF, false, &t0, &t1)) return false;
    const float kPad = 1e-6f;
    return t >= t0 * (1.0f - kPad) - kPad && t <= t1 * (1.0f + kPad) + kPad;
}

// Moller-Trumbore + own-box guard (geometry.h:146-176).
bool rayTriangle(const R8& r, const T3& tri, float tMax, float* tOut) {
    const P3 e1 = sub(tri.b, tri.a);
    const P3 e2 = sub(tri.c, tri.a);
    const P3 p = crs(r.d, e2);
    const float det = dot3(e1, p);
    if (std::fabs(det) < 1e-7f) return false;
    const float inv = 1.0f / det;
    const P3 tv = sub(r.o, tri.a);
    const float u = dot3(tv, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const P3 q = crs(tv, e1);
    const float v = dot3(r.d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = dot3(e2, q) * inv;
    if (t < r.tmin || t > tMax) return false;
    if (!tWithinTriBox(r, triBox(tri), t)) return false;
    *tOut = t;
    return true;
}

inline float d2(const P3& p, const P3& q) {
    const P3 d = sub(p, q);
    return dot3(d, d);
}

// Ericson 5.1.5 (geometry.h:198-259); returns dist2 and writes the region id
// (0 A, 1 B, 2 AB, 3 C, 4 AC, 5 BC, 6 face, 7..9 degenerate A/B/C) and barycentrics.
float triClosest(const P3& p, const T3& t, int* region, float* b1, float* b2) {
    const P3 ab = sub(t.b, t.a), ac = sub(t.c, t.a), ap = sub(p, t.a);
    const float d1 = dot3(ab, ap), d2v = dot3(ac, ap);
    if (d1 <= 0.0f && d2v <= 0.0f) {
        *region = 0; *b1 = 0; *b2 = 0;
        return d2(p, t.a);
    }
    const P3 bp = sub(p, t.b);
    const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        *region = 1; *b1 = 1; *b2 = 0;
        return d2(p, t.b);
    }
    const float vc = d1 * d4 - d3 * d2v;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float s = d1 / (d1 - d3);
        *region = 2; *b1 = s; *b2 = 0;
        return d2(p, add(t.a, mul(ab, s)));
    }
    const P3 cp = sub(p, t.c);
    const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        *region = 3; *b1 = 0; *b2 = 1;
        return d2(p, t.c);
    }
    const float vb = d5 * d2v - d1 * d6;
    if (vb <= 0.0f && d2v >= 0.0f && d6 <= 0.0f) {
        const float s = d2v / (d2v - d6);
        *region = 4; *b1 = 0; *b2 = s;
        return d2(p, add(t.a, mul(ac, s)));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float s = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *region = 5; *b1 = 1.0f - s; *b2 = s;
        return d2(p, add(t.b, mul(sub(t.c, t.b), s)));
    }
    const float denom = va + vb + vc;
    if (!(denom > 1e-20f)) {
        float best = d2(p, t.a);
        *region = 7; *b1 = 0; *b2 = 0;
        const float db = d2(p, t.b), dc = d2(p, t.c);
        if (db < best) { best = db; *region = 8; *b1 = 1; *b2 = 0; }
        if (dc < best) { best = dc; *region = 9; *b1 = 0; *b2 = 1; }
        return best;
    }
    const float s1 = vb / denom, s2 = vc / denom;
    *region = 6; *b1 = s1; *b2 = s2;
    return d2(p, add(add(t.a, mul(ab, s1)), mul(ac, s2)));
}

float triDist2(const P3& p, const T3& t) {
    int r;
    float a, b;
    return triClosest(p, t, &r, &a, &b);
}

// distanceSquaredToAABB (geometry.h:262-277).
float boxDist(const P3& p, const Bx& b) {
    float acc = 0.0f;
    for (int ax = 0; ax < 3; ++ax) {
        const float v = at(p, ax), lo = at(b.mn, ax), hi = at(b.mx, ax);
        float d = 0.0f;
        if (v < lo) {
            d = lo - v;
        } else if (v > hi) {
            d = v - hi;
        } else {
            continue;
        }
        acc += d * d;
    }
    return acc;
}

// gaussian.h:23-41
void axes(const Q4& q, P3& u, P3& v, P3& w) {
    const float n2 = std::fma(q.w, q.w, std::fma(q.z, q.z, std::fma(q.y, q.y, q.x * q.x)));
    const float s = n2 > 0.0f ? 1.0f / std::sqrt(n2) : 1.0f;
    const float x = q.x * s, y = q.y * s, z = q.z * s, ww = q.w * s;
    const float yz2 = std::fma(y, y, z * z);
    const float xz2 = std::fma(x, x, z * z);
    const float xy2 = std::fma(x, x, y * y);
    u = {std::fma(-2.0f, yz2, 1.0f), 2.0f * std::fma(x, y, z * ww), 2.0f * std::fma(x, z, -(y * ww))};
    v = {2.0f * std::fma(x, y, -(z * ww)), std::fma(-2.0f, xz2, 1.0f), 2.0f * std::fma(y, z, x * ww)};
    w = {2.0f * std::fma(x, z, y * ww), 2.0f * std::fma(y, z, -(x * ww)), std::fma(-2.0f, xy2, 1.0f)};
}

// gaussian.h:46-64
Bx gaussBox(const G10& g) {
    P3 u, v, w;
    axes(g.rot, u, v, w);
    const float a2 = g.scale.x * g.scale.x, b2 = g.scale.y * g.scale.y, c2 = g.scale.z * g.scale.z;
    float lo[3] = {g.mean.x, g.mean.y, g.mean.z};
    float hi[3] = {g.mean.x, g.mean.y, g.mean.z};
    for (int ax = 0; ax < 3; ++ax) {
        const float ui = at(u, ax), vi = at(v, ax), wi = at(w, ax);
        const float h = std::sqrt(std::fma(c2, wi * wi, std::fma(b2, vi * vi, a2 * ui * ui)));
        const float pad = h + 1e-6f * (1.0f + std::fabs(at(g.mean, ax)) + h);
        lo[ax] -= pad;
        hi[ax] += pad;
    }
    return {{lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
}

// gaussian.h:80-131: returns dist2, root t and whether the point is inside.
float gaussDist(const P3& p, const G10& g, float* rootT, int* inside) {
    P3 u, v, w;
    axes(g.rot, u, v, w);
    const P3 d = sub(p, g.mean);
    const float y0 = dot3(d, u), y1 = dot3(d, v), y2 = dot3(d, w);
    const float a = g.scale.x, b = g.scale.y, c = g.scale.z;
    const float aa = a * a, bb = b * b, cc = c * c;
    const float g0 = a * y0, g1 = b * y1, g2 = c * y2;
    const float r0 = y0 / a, r1 = y1 / b, r2 = y2 / c;
    if (std::fma(r2, r2, std::fma(r1, r1, r0 * r0)) <= 1.0f) {
        *rootT = 0.0f;
        *inside = 1;
        return 0.0f;
    }
    *inside = 0;
    const float len = std::sqrt(std::fma(y2, y2, std::fma(y1, y1, y0 * y0)));
    const float amax = a > b ? (a > c ? a : c) : (b > c ? b : c);
    float lo = 0.0f, hi = 2.0f * amax * len, t = 0.0f;
    if (len > amax) t = amax * (len - amax);
    for (int it = 0; it < 14; ++it) {
        const float i0 = 1.0f / (t + aa), i1 = 1.0f / (t + bb), i2 = 1.0f / (t + cc);
        const float s0 = g0 * i0, s1 = g1 * i1, s2 = g2 * i2;
        const float f = std::fma(s2, s2, std::fma(s1, s1, s0 * s0)) - 1.0f;
        if (f > 0.0f) lo = t; else hi = t;
        const float fp = -2.0f * (s0 * s0 * i0 + s1 * s1 * i1 + s2 * s2 * i2);
        float nxt = 0.5f * (lo + hi);
        if (fp < 0.0f) {
            const float nt = t - f / fp;
            if (nt > lo && nt < hi) nxt = nt;
        }
        t = nxt;
    }
    *rootT = t;
    const float i0 = 1.0f / (t + aa), i1 = 1.0f / (t + bb), i2 = 1.0f / (t + cc);
    const float e0 = y0 - g0 * i0 * a, e1 = y1 - g1 * i1 * b, e2 = y2 - g2 * i2 * c;
    return std::fma(e2, e2, std::fma(e1, e1, e0 * e0));
}

// ---- leaf kinds ------------------------------------------------------------------
enum { KTRI = 0, KPT = 1, KGS = 2 };

size_t isz(int kind) { return kind == KTRI ? 36 : (kind == KPT ? 12 : 40); }

Bx itemBox(int kind, const void* base, int i) {
    const char* p = static_cast<const char*>(base) + static_cast<size_t>(i) * isz(kind);
    if (kind == KTRI) return triBox(*reinterpret_cast<const T3*>(p));
    if (kind == KPT) {
        Bx b = boxEmpty();
        boxAdd(b, *reinterpret_cast<const P3*>(p));
        return b;
    }
    return gaussBox(*reinterpret_cast<const G10*>(p));
}

float itemDist(int kind, const void* base, int i, const P3& q) {
    const char* p = static_cast<const char*>(base) + static_cast<size_t>(i) * isz(kind);
    if (kind == KTRI) return triDist2(q, *reinterpret_cast<const T3*>(p));
    if (kind == KPT) return d2(q, *reinterpret_cast<const P3*>(p));
    float t;
    int in;
    return gaussDist(q, *reinterpret_cast<const G10*>(p), &t, &in);
}

// Split centroid of the reference CPU BVH: triangle box centre
// (src/cpu/bvh.cpp:94) or Gaussian mean (src/cpu/gauss_bvh.cpp:84).
P3 splitCentroid(int kind, const void* base, int i) {
    const char* p = static_cast<const char*>(base) + static_cast<size_t>(i) * isz(kind);
    if (kind == KTRI) return boxCenter(triBox(*reinterpret_cast<const T3*>(p)));
    if (kind == KPT) return *reinterpret_cast<const P3*>(p);
    return reinterpret_cast<const G10*>(p)->mean;
}

struct Hit {
    float d;
    int id;
};
inline bool better(const Hit& a, const Hit& b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

uint64_t packKey(const Hit& h) {
    uint32_t bits;
    std::memcpy(&bits, &h.d, 4);
    return (static_cast<uint64_t>(bits) << 32) | static_cast<uint32_t>(h.id);
}

template <typename F>
void parallelFor(int n, int threads, F&& f) {
    if (threads <= 1 || n < 64) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    const int chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&, b, e] {
            for (int i = b; i < e; ++i) f(i);
        });
    }
    for (auto& th : pool) th.join();
}

// ---- median-split CPU BVH (src/cpu/bvh.cpp, src/cpu/gauss_bvh.cpp) ----------------
struct CNode {
    Bx box;
    int start, count, right;
};

struct CBvh {
    int kind;
    std::vector<CNode> nodes;
    std::vector<char> items;  // leaf order
    std::vector<int> ids;
};

int buildRange(CBvh& out, const void* src, std::vector<int>& order, const std::vector<Bx>& boxes,
               std::vector<P3>& cen, int start, int end) {
    const int idx = static_cast<int>(out.nodes.size());
    out.nodes.push_back(CNode{boxEmpty(), 0, 0, -1});
    Bx box = boxEmpty(), cb = boxEmpty();
    for (int i = start; i < end; ++i) {
        boxUnion(box, boxes[order[i]]);
        boxAdd(cb, cen[order[i]]);
    }
    out.nodes[idx].box = box;
    const int count = end - start;
    if (count <= 4) {
        out.nodes[idx].start = start;
        out.nodes[idx].count = count;
        return idx;
    }
    const P3 ext = sub(cb.mx, cb.mn);
    const int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
    int mid = start + count / 2;
    std::nth_element(order.begin() + start, order.begin() + mid, order.begin() + end,
                     [&](int a, int b) { return at(cen[a], axis) < at(cen[b], axis); });
    buildRange(out, src, order, boxes, cen, start, mid);
    const int r = buildRange(out, src, order, boxes, cen, mid, end);
    out.nodes[idx].right = r;
    (void)src;
    return idx;
}

float leafDist(const CBvh& b, int slot, const P3& q) { return itemDist(b.kind, b.items.data(), slot, q); }

Hit bvhClosest(const CBvh& b, const P3& p) {
    // Best-first with a min-heap on the box lower bound (src/cpu/bvh.cpp:222-259).
    Hit best{std::numeric_limits<float>::max(), -1};
    bool found = false;
    thread_local std::vector<std::pair<float, int>> heap;
    heap.clear();
    auto cmp = [](const std::pair<float, int>& x, const std::pair<float, int>& y) { return x.first > y.first; };
    heap.push_back({boxDist(p, b.nodes[0].box), 0});
    while (!heap.empty()) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        const auto cur = heap.back();
        heap.pop_back();
        if (found && cur.first > best.d) break;
        const CNode& nd = b.nodes[cur.second];
        if (nd.count > 0) {
            for (int i = 0; i < nd.count; ++i) {
                const int s = nd.start + i;
                const Hit h{leafDist(b, s, p), b.ids[s]};
                if (!found || better(h, best)) {
                    best = h;
                    found = true;
                }
            }
            continue;
        }
        const int kids[2] = {cur.second + 1, nd.right};
        for (int c : kids) {
            const float bd = boxDist(p, b.nodes[c].box);
            if (!found || bd <= best.d) {
                heap.push_back({bd, c});
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
        }
    }
    if (!found) best.id = -1;
    return best;
}

void bvhKnn(const CBvh& b, const P3& p, int k, std::vector<Hit>& out) {
    // Max-heap of the k best (src/cpu/bvh.cpp:261-313).
    out.clear();
    std::vector<Hit> heap;
    auto worse = [](const Hit& x, const Hit& y) { return better(x, y); };
    thread_local std::vector<std::pair<float, int>> fr;
    fr.clear();
    auto cmp = [](const std::pair<float, int>& x, const std::pair<float, int>& y) { return x.first > y.first; };
    fr.push_back({boxDist(p, b.nodes[0].box), 0});
    while (!fr.empty()) {
        std::pop_heap(fr.begin(), fr.end(), cmp);
        const auto cur = fr.back();
        fr.pop_back();
        const bool full = static_cast<int>(heap.size()) >= k;
        if (full && cur.first > heap.front().d) break;
        const CNode& nd = b.nodes[cur.second];
        if (nd.count > 0) {
            for (int i = 0; i < nd.count; ++i) {
                const int s = nd.start + i;
                const Hit h{leafDist(b, s, p), b.ids[s]};
                if (static_cast<int>(heap.size()) >= k) {
                    if (!better(h, heap.front())) continue;
                    std::pop_heap(heap.begin(), heap.end(), worse);
                    heap.pop_back();
                }
                heap.push_back(h);
                std::push_heap(heap.begin(), heap.end(), worse);
            }
            continue;
        }
        const int kids[2] = {cur.second + 1, nd.right};
        for (int c : kids) {
            const float bd = boxDist(p, b.nodes[c].box);
            if (static_cast<int>(heap.size()) < k || bd <= heap.front().d) {
                fr.push_back({bd, c});
                std::push_heap(fr.begin(), fr.end(), cmp);
            }
        }
    }
    out.assign(heap.begin(), heap.end());
    std::sort(out.begin(), out.end(), better);
}

void bvhRadius(const CBvh& b, const P3& p, float r, std::vector<Hit>& out) {
    out.clear();
    if (r < 0.0f) return;
    const float r2 = r * r;
    thread_local std::vector<int> st;
    st.clear();
    st.push_back(0);
    while (!st.empty()) {
        const int i = st.back();
        st.pop_back();
        const CNode& nd = b.nodes[i];
        if (boxDist(p, nd.box) > r2) continue;
        if (nd.count > 0) {
            for (int j = 0; j < nd.count; ++j) {
                const int s = nd.start + j;
                const float d = leafDist(b, s, p);
                if (d > r2) continue;
                out.push_back({d, b.ids[s]});
            }
        } else {
            st.push_back(nd.right);
            st.push_back(i + 1);
        }
    }
    std::sort(out.begin(), out.end(), better);
}

// Closest hit: DFS, left child popped first, t re-tested per pop
// (src/cpu/bvh.cpp:138-178).
uint64_t bvhRay(const CBvh& b, const R8& r) {
    float bestT = r.tmax;
    bool hit = false;
    float hitT = 0.0f;
    int st[128];
    int sp = 0;
    st[sp++] = 0;
    while (sp > 0) {
        const int i = st[--sp];
        const CNode& nd = b.nodes[i];
        float te;
        if (!rayHitsBox(r, nd.box, bestT, &te)) continue;
        if (nd.count > 0) {
            for (int j = 0; j < nd.count; ++j) {
                const T3& tri = *reinterpret_cast<const T3*>(b.items.data() + static_cast<size_t>(nd.start + j) * 36);
                float t;
                if (rayTriangle(r, tri, bestT, &t)) {
                    bestT = t;
                    hit = true;
                    hitT = t;
                }
            }
        } else {
            st[sp++] = nd.right;
            st[sp++] = i + 1;
        }
    }
    if (!hit) return 0ull;
    uint32_t bits;
    std::memcpy(&bits, &hitT, 4);
    return (1ull << 32) | bits;
}

// ---- point kd-tree (src/cpu/point_kd.cpp) -----------------------------------------
struct KNode {
    Bx box;
    int start, count, left, right;
};
struct Kd {
    std::vector<KNode> nodes;
    std::vector<P3> pts;
    std::vector<int> ids;
};

int kdBuild(Kd& t, std::vector<std::pair<P3, int>>& s, int start, int count) {
    const int id = static_cast<int>(t.nodes.size());
    t.nodes.push_back(KNode{boxEmpty(), start, count, -1, -1});
    Bx box = boxEmpty();
    for (int i = start; i < start + count; ++i) boxAdd(box, s[i].first);
    t.nodes[id].box = box;
    if (count <= 8) return id;
    const P3 ext = sub(box.mx, box.mn);
    const int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
    const int mid = start + count / 2;
    std::nth_element(s.begin() + start, s.begin() + mid, s.begin() + start + count,
                     [axis](const std::pair<P3, int>& a, const std::pair<P3, int>& b) {
                         return at(a.first, axis) < at(b.first, axis);
                     });
    t.nodes[id].count = 0;
    const int l = kdBuild(t, s, start, mid - start);
    const int r = kdBuild(t, s, mid, start + count - mid);
    t.nodes[id].left = l;
    t.nodes[id].right = r;
    return id;
}

void kdClosestRec(const Kd& t, int id, const P3& p, Hit& best, bool& found) {
    const KNode& nd = t.nodes[id];
    if (found && boxDist(p, nd.box) > best.d) return;
    if (nd.count > 0) {
        for (int i = nd.start; i < nd.start + nd.count; ++i) {
            const Hit h{d2(p, t.pts[i]), t.ids[i]};
            if (!found || better(h, best)) {
                best = h;
                found = true;
            }
        }
        return;
    }
    const float dl = boxDist(p, t.nodes[nd.left].box), dr = boxDist(p, t.nodes[nd.right].box);
    const int first = dl <= dr ? nd.left : nd.right, second = dl <= dr ? nd.right : nd.left;
    kdClosestRec(t, first, p, best, found);
    kdClosestRec(t, second, p, best, found);
}

void kdKnnRec(const Kd& t, int id, const P3& p, int k, std::vector<Hit>& best) {
    const KNode& nd = t.nodes[id];
    if (static_cast<int>(best.size()) == k && boxDist(p, nd.box) > best.back().d) return;
    if (nd.count > 0) {
        for (int i = nd.start; i < nd.start + nd.count; ++i) {
            const Hit h{d2(p, t.pts[i]), t.ids[i]};
            if (static_cast<int>(best.size()) == k && !better(h, best.back())) continue;
            auto pos = std::upper_bound(best.begin(), best.end(), h, better);
            best.insert(pos, h);
            if (static_cast<int>(best.size()) > k) best.pop_back();
        }
        return;
    }
    const float dl = boxDist(p, t.nodes[nd.left].box), dr = boxDist(p, t.nodes[nd.right].box);
    const int first = dl <= dr ? nd.left : nd.right, second = dl <= dr ? nd.right : nd.left;
    kdKnnRec(t, first, p, k, best);
    kdKnnRec(t, second, p, k, best);
}

void kdRadiusRec(const Kd& t, int id, const P3& p, float r2, std::vector<Hit>& out) {
    const KNode& nd = t.nodes[id];
    if (boxDist(p, nd.box) > r2) return;
    if (nd.count > 0) {
        for (int i = nd.start; i < nd.start + nd.count; ++i) {
            const float d = d2(p, t.pts[i]);
            if (d <= r2) out.push_back({d, t.ids[i]});
        }
        return;
    }
    kdRadiusRec(t, nd.left, p, r2, out);
    kdRadiusRec(t, nd.right, p, r2, out);
}

// ---- CPU LBVH with the GPU node layout ------------------------------------------------
// Morton keys exactly as src/gpu/lbvh.cu:277-315 (centroid bounds with fmin/fmax,
// quantize, 30-bit interleave, key = code << 32 | index), a full 64-bit sort, the
// radix tree by recursive splitting at the highest differing bit (the same tree
// Karras 2012 emits), leaf boxes and parent boxes merged with the left child as
// the first compare-select operand (src/gpu/lbvh.cu:345-354, 506-512).
struct GNode {
    float lo[3], hi[3];
    int left, right;
};
static_assert(sizeof(GNode) == 32, "gpu node layout");

P3 lbvhCentroid(int kind, const void* base, int i) {
    const char* p = static_cast<const char*>(base) + static_cast<size_t>(i) * isz(kind);
    if (kind == KTRI) {
        const T3& t = *reinterpret_cast<const T3*>(p);
        return {(t.a.x + t.b.x + t.c.x) / 3.0f, (t.a.y + t.b.y + t.c.y) / 3.0f,
                (t.a.z + t.b.z + t.c.z) / 3.0f};
    }
    if (kind == KPT) return *reinterpret_cast<const P3*>(p);
    return reinterpret_cast<const G10*>(p)->mean;
}

uint32_t spread(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

uint32_t q1023(float v, float lo, float ext) {
    if (ext <= 0.0f) return 0;
    float q = (v - lo) / ext;
    q = q * 1023.0f;
    q = q + 0.5f;
    q = std::fmax(0.0f, q);
    q = std::fmin(1023.0f, q);
    return static_cast<uint32_t>(static_cast<int>(q));
}

int highestDiff(uint64_t a, uint64_t b) { return a == b ? 64 : __builtin_clzll(a ^ b); }

}  // namespace

extern "C" {

// ---- leaf math probes -------------------------------------------------------------
int orc_tri_intersect(const float* ray, const float* tri, float* t) {
    const R8& r = *reinterpret_cast<const R8*>(ray);
    return rayTriangle(r, *reinterpret_cast<const T3*>(tri), r.tmax, t) ? 1 : 0;
}
float orc_tri_closest(const float* p, const float* tri, int* region, float* b1, float* b2) {
    return triClosest(*reinterpret_cast<const P3*>(p), *reinterpret_cast<const T3*>(tri), region, b1, b2);
}
float orc_gauss_dist2(const float* p, const float* g, float* rootT, int* inside) {
    return gaussDist(*reinterpret_cast<const P3*>(p), *reinterpret_cast<const G10*>(g), rootT, inside);
}
void orc_gauss_bounds(const float* g, float* box6) {
    const Bx b = gaussBox(*reinterpret_cast<const G10*>(g));
    const float v[6] = {b.mn.x, b.mn.y, b.mn.z, b.mx.x, b.mx.y, b.mx.z};
    std::memcpy(box6, v, sizeof(v));
}
float orc_box_dist2(const float* p, const float* box6) {
    const Bx b{{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
    return boxDist(*reinterpret_cast<const P3*>(p), b);
}
int orc_ray_box(const float* ray, const float* box6, float tmax, float* tEnter) {
    const Bx b{{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
    return rayHitsBox(*reinterpret_cast<const R8*>(ray), b, tmax, tEnter) ? 1 : 0;
}

// ---- brute force (src/cpu/brute.cpp) ----------------------------------------------------
void orc_ray_brute(const float* tris, int n, const float* rays, int nr, uint64_t* out, int threads) {
    const T3* T = reinterpret_cast<const T3*>(tris);
    const R8* R = reinterpret_cast<const R8*>(rays);
    parallelFor(nr, threads, [&](int q) {
        float best = R[q].tmax;
        bool hit = false;
        for (int i = 0; i < n; ++i) {
            float t;
            if (rayTriangle(R[q], T[i], best, &t)) {
                best = t;
                hit = true;
            }
        }
        uint32_t bits;
        std::memcpy(&bits, &best, 4);
        out[q] = hit ? ((1ull << 32) | bits) : 0ull;
    });
}

// k == 0: closest (one record per query).  Records are {u32 dist2 bits, i32 id}.
void orc_knn_brute(int kind, const void* items, int n, const float* pts, int nq, int k,
                   uint32_t* out, int threads) {
    const P3* Q = reinterpret_cast<const P3*>(pts);
    const int kk = k == 0 ? 1 : k;
    parallelFor(nq, threads, [&](int q) {
        std::vector<Hit> all(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) all[i] = {itemDist(kind, items, i, Q[q]), i};
        const int m = std::min(n, kk);
        std::partial_sort(all.begin(), all.begin() + m, all.end(), better);
        for (int j = 0; j < kk; ++j) {
            uint32_t* rec = out + (static_cast<size_t>(q) * kk + j) * 2;
            if (j < m) {
                std::memcpy(&rec[0], &all[j].d, 4);
                rec[1] = static_cast<uint32_t>(all[j].id);
            } else {
                rec[0] = 0u;
                rec[1] = 0xffffffffu;
            }
        }
    });
}

// Radius: fills counts[nq]; if keys != NULL also writes the sorted u64 keys at
// offsets (exclusive scan of counts, computed by the caller).
void orc_radius_brute(int kind, const void* items, int n, const float* pts, int nq, float r,
                      int* counts, const int* offsets, uint64_t* keys, int threads) {
    const P3* Q = reinterpret_cast<const P3*>(pts);
    const float r2 = r * r;
    parallelFor(nq, threads, [&](int q) {
        std::vector<Hit> hits;
        if (r >= 0.0f) {
            for (int i = 0; i < n; ++i) {
                const float d = itemDist(kind, items, i, Q[q]);
                if (d <= r2) hits.push_back({d, i});
            }
        }
        std::sort(hits.begin(), hits.end(), better);
        counts[q] = static_cast<int>(hits.size());
        if (keys)
            for (size_t j = 0; j < hits.size(); ++j) keys[offsets[q] + j] = packKey(hits[j]);
    });
}

// ---- median-split BVH ------------------------------------------------------------------
void* orc_bvh_new(int kind, const void* items, int n) {
    if (n <= 0) return nullptr;
    auto* b = new CBvh();
    b->kind = kind;
    std::vector<int> order(static_cast<size_t>(n));
    std::vector<Bx> boxes(static_cast<size_t>(n));
    std::vector<P3> cen(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        order[i] = i;
        boxes[i] = itemBox(kind, items, i);
        cen[i] = kind == KTRI ? boxCenter(boxes[i]) : splitCentroid(kind, items, i);
    }
    b->nodes.reserve(static_cast<size_t>(n) / 2 + 16);
    buildRange(*b, items, order, boxes, cen, 0, n);
    const size_t s = isz(kind);
    b->items.resize(static_cast<size_t>(n) * s);
    for (int i = 0; i < n; ++i)
        std::memcpy(&b->items[static_cast<size_t>(i) * s], static_cast<const char*>(items) + static_cast<size_t>(order[i]) * s, s);
    b->ids = order;
    return b;
}

void orc_bvh_free(void* h) { delete static_cast<CBvh*>(h); }

void orc_bvh_ray(void* h, const float* rays, int nr, uint64_t* out, int threads) {
    const CBvh& b = *static_cast<CBvh*>(h);
    const R8* R = reinterpret_cast<const R8*>(rays);
    parallelFor(nr, threads, [&](int q) { out[q] = bvhRay(b, R[q]); });
}

void orc_bvh_knn(void* h, const float* pts, int nq, int k, uint32_t* out, int threads) {
    const CBvh& b = *static_cast<CBvh*>(h);
    const P3* Q = reinterpret_cast<const P3*>(pts);
    const int kk = k == 0 ? 1 : k;
    parallelFor(nq, threads, [&](int q) {
        std::vector<Hit> res;
        if (k == 0) {
            const Hit h1 = bvhClosest(b, Q[q]);
            if (h1.id >= 0) res.push_back(h1);
        } else {
            bvhKnn(b, Q[q], k, res);
        }
        for (int j = 0; j < kk; ++j) {
            uint32_t* rec = out + (static_cast<size_t>(q) * kk + j) * 2;
            if (j < static_cast<int>(res.size())) {
                std::memcpy(&rec[0], &res[j].d, 4);
                rec[1] = static_cast<uint32_t>(res[j].id);
            } else {
                rec[0] = 0u;
                rec[1] = 0xffffffffu;
            }
        }
    });
}

void orc_bvh_radius(void* h, const float* pts, int nq, float r, int* counts, const int* offsets,
                    uint64_t* keys, int threads) {
    const CBvh& b = *static_cast<CBvh*>(h);
    const P3* Q = reinterpret_cast<const P3*>(pts);
    parallelFor(nq, threads, [&](int q) {
        std::vector<Hit> res;
        bvhRadius(b, Q[q], r, res);
        counts[q] = static_cast<int>(res.size());
        if (keys)
            for (size_t j = 0; j < res.size(); ++j) keys[offsets[q] + j] = packKey(res[j]);
    });
}

// ---- kd-tree ------------------------------------------------------------------------
void* orc_kd_new(const float* pts, int n) {
    if (n <= 0) return nullptr;
    auto* t = new Kd();
    std::vector<std::pair<P3, int>> s(static_cast<size_t>(n));
    const P3* P = reinterpret_cast<const P3*>(pts);
    for (int i = 0; i < n; ++i) s[i] = {P[i], i};
    t->nodes.reserve(static_cast<size_t>(n) / 4 + 16);
    kdBuild(*t, s, 0, n);
    for (auto& e : s) {
        t->pts.push_back(e.first);
        t->ids.push_back(e.second);
    }
    return t;
}
void orc_kd_free(void* h) { delete static_cast<Kd*>(h); }
void orc_kd_knn(void* h, const float* pts, int nq, int k, uint32_t* out, int threads) {
    const Kd& t = *static_cast<Kd*>(h);
    const P3* Q = reinterpret_cast<const P3*>(pts);
    const int kk = k == 0 ? 1 : k;
    parallelFor(nq, threads, [&](int q) {
        std::vector<Hit> res;
        if (k == 0) {
            Hit best{0, -1};
            bool found = false;
            kdClosestRec(t, 0, Q[q], best, found);
            if (found) res.push_back(best);
        } else {
            kdKnnRec(t, 0, Q[q], k, res);
        }
        for (int j = 0; j < kk; ++j) {
            uint32_t* rec = out + (static_cast<size_t>(q) * kk + j) * 2;
            if (j < static_cast<int>(res.size())) {
                std::memcpy(&rec[0], &res[j].d, 4);
                rec[1] = static_cast<uint32_t>(res[j].id);
            } else {
                rec[0] = 0u;
                rec[1] = 0xffffffffu;
            }
        }
    });
}
void orc_kd_radius(void* h, const float* pts, int nq, float r, int* counts, const int* offsets,
                   uint64_t* keys, int threads) {
    const Kd& t = *static_cast<Kd*>(h);
    const P3* Q = reinterpret_cast<const P3*>(pts);
    const float r2 = r * r;
    parallelFor(nq, threads, [&](int q) {
        std::vector<Hit> res;
        if (r >= 0.0f) kdRadiusRec(t, 0, Q[q], r2, res);
        std::sort(res.begin(), res.end(), better);
        counts[q] = static_cast<int>(res.size());
        if (keys)
            for (size_t j = 0; j < res.size(); ++j) keys[offsets[q] + j] = packKey(res[j]);
    });
}

// ---- CPU LBVH ------------------------------------------------------------------------------
// nodes_out: (2n-1) x 32 B; prim_ids_out: n; returns the depth (max leaf depth).
int orc_lbvh_nodes(int kind, const void* items, int n, void* nodesOut, int* primIdsOut) {
    if (n < 2) return -1;
    float lo[3] = {INF, INF, INF}, hi[3] = {-INF, -INF, -INF};
    std::vector<P3> cen(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        cen[i] = lbvhCentroid(kind, items, i);
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], at(cen[i], a));
            hi[a] = std::fmax(hi[a], at(cen[i], a));
        }
    }
    const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    std::vector<uint64_t> keys(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const uint32_t code = spread(q1023(cen[i].x, lo[0], ext[0])) |
                              (spread(q1023(cen[i].y, lo[1], ext[1])) << 1) |
                              (spread(q1023(cen[i].z, lo[2], ext[2])) << 2);
        keys[i] = (static_cast<uint64_t>(code) << 32) | static_cast<uint32_t>(i);
    }
    std::sort(keys.begin(), keys.end());
    GNode* nodes = static_cast<GNode*>(nodesOut);
    std::vector<int> depthOf(static_cast<size_t>(2 * n - 1), 0);
    // Radix tree: the node over sorted range [first, last] (first < last) splits at
    // the last position whose key agrees with keys[first] above the highest bit in
    // which keys[first] and keys[last] differ.  Its id is Karras's: the range end
    // that is not shared with... equivalently, a range that is the left child of
    // its parent gets id `last`, a right child gets id `first`, the root gets 0.
    struct Task {
        int first, last, id, depth;
    };
    std::vector<Task> stack;
    stack.push_back({0, n - 1, 0, 0});
    int maxDepth = 0;
    while (!stack.empty()) {
        const Task tk = stack.back();
        stack.pop_back();
        const int bit = highestDiff(keys[tk.first], keys[tk.last]);
        // binary search for the split: largest s in [first, last) with prefix > bit
        int a = tk.first, b = tk.last;
        while (b - a > 1) {
            const int m = a + (b - a) / 2;
            if (highestDiff(keys[tk.first], keys[m]) > bit) a = m; else b = m;
        }
        const int split = a;
        const int leftId = split == tk.first ? n - 1 + split : split;
        const int rightId = split + 1 == tk.last ? n - 1 + tk.last : split + 1;
        nodes[tk.id].left = leftId;
        nodes[tk.id].right = rightId;
        depthOf[tk.id] = tk.depth;
        depthOf[leftId] = tk.depth + 1;
        depthOf[rightId] = tk.depth + 1;
        if (tk.depth + 1 > maxDepth) maxDepth = tk.depth + 1;
        if (leftId < n - 1) stack.push_back({tk.first, split, leftId, tk.depth + 1});
        if (rightId < n - 1) stack.push_back({split + 1, tk.last, rightId, tk.depth + 1});
    }
    // Leaves and boxes.
    const size_t s = isz(kind);
    std::vector<Bx> box(static_cast<size_t>(2 * n - 1));
    for (int k = 0; k < n; ++k) {
        const int src = static_cast<int>(keys[k] & 0xffffffffu);
        primIdsOut[k] = src;
        box[n - 1 + k] = itemBox(kind, items, src);
        GNode& lf = nodes[n - 1 + k];
        lf.left = -(k + 1);
        lf.right = 1;
    }
    (void)s;
    // Internal boxes deepest first: process nodes by decreasing depth.
    std::vector<int> ord(static_cast<size_t>(n - 1));
    for (int i = 0; i < n - 1; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return depthOf[x] > depthOf[y]; });
    for (int i : ord) {
        const Bx& L = box[nodes[i].left];
        const Bx& R = box[nodes[i].right];
        Bx m;
        m.mn.x = L.mn.x < R.mn.x ? L.mn.x : R.mn.x;
        m.mn.y = L.mn.y < R.mn.y ? L.mn.y : R.mn.y;
        m.mn.z = L.mn.z < R.mn.z ? L.mn.z : R.mn.z;
        m.mx.x = L.mx.x > R.mx.x ? L.mx.x : R.mx.x;
        m.mx.y = L.mx.y > R.mx.y ? L.mx.y : R.mx.y;
        m.mx.z = L.mx.z > R.mx.z ? L.mx.z : R.mx.z;
        box[i] = m;
    }
    for (int i = 0; i < 2 * n - 1; ++i) {
        nodes[i].lo[0] = box[i].mn.x;
        nodes[i].lo[1] = box[i].mn.y;
        nodes[i].lo[2] = box[i].mn.z;
        nodes[i].hi[0] = box[i].mx.x;
        nodes[i].hi[1] = box[i].mx.y;
        nodes[i].hi[2] = box[i].mx.z;
    }
    return maxDepth;
}

// ---- fingerprints (src/core/fingerprint.h:16-94) ---------------------------------------------
uint64_t orc_fnv_bytes(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= c[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

// RayStats: mix (hit << 32 | t bits) per ray; also hits and sum of t.
uint64_t orc_ray_hash(const uint64_t* v, int n, uint64_t* hits, double* sumT) {
    uint64_t h = 0xcbf29ce484222325ull, hc = 0;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        if (v[i] >> 32) {
            ++hc;
            float t;
            const uint32_t bits = static_cast<uint32_t>(v[i]);
            std::memcpy(&t, &bits, 4);
            s += static_cast<double>(t);
        }
        h ^= v[i];
        h *= 0x100000001b3ull;
    }
    if (hits) *hits = hc;
    if (sumT) *sumT = s;
    return h;
}

// QueryStats single recipe: mix (found << 32 | d2 bits), then (id + 1 or 0).
uint64_t orc_closest_hash(const uint32_t* rec, int n, double* sumD2) {
    uint64_t h = 0xcbf29ce484222325ull;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const uint32_t bits = rec[2 * i];
        const int id = static_cast<int>(rec[2 * i + 1]);
        const bool found = id >= 0;
        const uint64_t a = (static_cast<uint64_t>(found ? 1u : 0u) << 32) | (found ? bits : 0u);
        h ^= a;
        h *= 0x100000001b3ull;
        const uint64_t b = found ? static_cast<uint64_t>(id) + 1u : 0u;
        h ^= b;
        h *= 0x100000001b3ull;
        if (found) {
            float d;
            std::memcpy(&d, &bits, 4);
            s += static_cast<double>(d);
        }
    }
    if (sumD2) *sumD2 = s;
    return h;
}

// QueryStats list recipe over kNN records (stride k, padding id -1 skipped):
// mix the list length, then per hit d2 bits and id + 1.
uint64_t orc_knn_hash(const uint32_t* rec, int n, int k, double* sumD2) {
    uint64_t h = 0xcbf29ce484222325ull;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        int len = 0;
        while (len < k && static_cast<int>(rec[2 * (static_cast<size_t>(i) * k + len) + 1]) >= 0) ++len;
        h ^= static_cast<uint64_t>(len);
        h *= 0x100000001b3ull;
        for (int j = 0; j < len; ++j) {
            const uint32_t* r = rec + 2 * (static_cast<size_t>(i) * k + j);
            h ^= static_cast<uint64_t>(r[0]);
            h *= 0x100000001b3ull;
            h ^= static_cast<uint64_t>(static_cast<int>(r[1])) + 1u;
            h *= 0x100000001b3ull;
            float d;
            std::memcpy(&d, &r[0], 4);
            s += static_cast<double>(d);
        }
    }
    if (sumD2) *sumD2 = s;
    return h;
}

// Same list recipe over a radius CSR.
uint64_t orc_csr_hash(const int* offsets, const uint64_t* keys, int n, double* sumD2) {
    uint64_t h = 0xcbf29ce484222325ull;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const int b = offsets[i], e = offsets[i + 1];
        h ^= static_cast<uint64_t>(e - b);
        h *= 0x100000001b3ull;
        for (int j = b; j < e; ++j) {
            const uint32_t bits = static_cast<uint32_t>(keys[j] >> 32);
            const uint32_t id = static_cast<uint32_t>(keys[j]);
            h ^= static_cast<uint64_t>(bits);
            h *= 0x100000001b3ull;
            h ^= static_cast<uint64_t>(static_cast<int>(id)) + 1u;
            h *= 0x100000001b3ull;
            float d;
            std::memcpy(&d, &bits, 4);
            s += static_cast<double>(d);
        }
    }
    if (sumD2) *sumD2 = s;
    return h;
}

}  // extern "C"
