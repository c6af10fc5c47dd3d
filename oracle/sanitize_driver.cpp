// Sanitizer run of the oracle (test infrastructure): `make -C oracle asan` links this
// driver with oracle.cpp under AddressSanitizer and UndefinedBehaviorSanitizer (host
// code only), and tests/test_oracle_sanitized.py runs it.  Every public oracle entry
// point is exercised on seeded random scenes -- triangles, points and Gaussians,
// with duplicates, empty result sets, k above the item count, a negative radius --
// and each tree's results are compared with brute force, so an out-of-bounds access,
// a use after free, an overflow or a wrong result ends the run with a non-zero status.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
int orc_tri_intersect(const float* ray, const float* tri, float* t);
float orc_tri_closest(const float* p, const float* tri, int* region, float* b1, float* b2);
float orc_gauss_dist2(const float* p, const float* g, float* rootT, int* inside);
void orc_gauss_bounds(const float* g, float* box6);
float orc_box_dist2(const float* p, const float* box6);
int orc_ray_box(const float* ray, const float* box6, float tmax, float* tEnter);
void orc_ray_brute(const float* tris, int n, const float* rays, int nr, uint64_t* out, int threads);
void orc_knn_brute(int kind, const void* items, int n, const float* pts, int nq, int k, uint32_t* out,
                   int threads);
void orc_radius_brute(int kind, const void* items, int n, const float* pts, int nq, float r, int* counts,
                      const int* offsets, uint64_t* keys, int threads);
void* orc_bvh_new(int kind, const void* items, int n);
void orc_bvh_free(void* h);
void orc_bvh_ray(void* h, const float* rays, int nr, uint64_t* out, int threads);
void orc_bvh_knn(void* h, const float* pts, int nq, int k, uint32_t* out, int threads);
void orc_bvh_radius(void* h, const float* pts, int nq, float r, int* counts, const int* offsets,
                    uint64_t* keys, int threads);
void* orc_kd_new(const float* pts, int n);
void orc_kd_free(void* h);
void orc_kd_knn(void* h, const float* pts, int nq, int k, uint32_t* out, int threads);
void orc_kd_radius(void* h, const float* pts, int nq, float r, int* counts, const int* offsets,
                   uint64_t* keys, int threads);
int orc_lbvh_nodes(int kind, const void* items, int n, void* nodesOut, int* primIdsOut);
uint64_t orc_fnv_bytes(const void* p, size_t n);
uint64_t orc_ray_hash(const uint64_t* v, int n, uint64_t* hits, double* sumT);
uint64_t orc_closest_hash(const uint32_t* rec, int n, double* sumD2);
uint64_t orc_knn_hash(const uint32_t* rec, int n, int k, double* sumD2);
uint64_t orc_csr_hash(const int* offsets, const uint64_t* keys, int n, double* sumD2);
}

namespace {

int gFailures = 0;

void expect(bool ok, const char* what, int kind) {
    if (!ok) {
        std::fprintf(stderr, "MISMATCH: %s (kind %d)\n", what, kind);
        ++gFailures;
    }
}

int floatsPer(int kind) { return kind == 0 ? 9 : kind == 1 ? 3 : 10; }

std::vector<float> scene(int kind, int n, std::mt19937& rng) {
    std::uniform_real_distribution<float> pos(-20.0f, 20.0f), off(-1.5f, 1.5f), q(-1.0f, 1.0f),
        sc(0.05f, 1.2f);
    std::vector<float> v(static_cast<size_t>(n) * floatsPer(kind));
    for (int i = 0; i < n; ++i) {
        float* e = &v[static_cast<size_t>(i) * floatsPer(kind)];
        const float x = pos(rng), y = pos(rng), z = pos(rng);
        if (kind == 0) {
            const float t[9] = {x, y, z, x + off(rng), y + off(rng), z + off(rng), x + off(rng), y + off(rng),
                                z + off(rng)};
            std::memcpy(e, t, sizeof(t));
        } else if (kind == 1) {
            e[0] = x, e[1] = y, e[2] = z;
        } else {
            const float g[10] = {x, y, z, q(rng), q(rng), q(rng), q(rng) + 1.5f, sc(rng), sc(rng), sc(rng)};
            std::memcpy(e, g, sizeof(g));
        }
    }
    // duplicates: equal distances, ordered by index
    for (int i = 0; i < n / 20; ++i)
        std::memcpy(&v[static_cast<size_t>(n - 1 - i) * floatsPer(kind)], &v[0], sizeof(float) * floatsPer(kind));
    return v;
}

std::vector<int> offsetsOf(const std::vector<int>& counts) {
    std::vector<int> o(counts.size() + 1, 0);
    for (size_t i = 0; i < counts.size(); ++i) o[i + 1] = o[i] + counts[i];
    return o;
}

void checkKind(int kind, int n, int nq, std::mt19937& rng) {
    const std::vector<float> items = scene(kind, n, rng);
    std::uniform_real_distribution<float> qd(-26.0f, 26.0f);
    std::vector<float> pts(static_cast<size_t>(nq) * 3);
    for (float& x : pts) x = qd(rng);

    void* bvh = orc_bvh_new(kind, items.data(), n);
    for (int k : {0, 1, 7, 32, n + 3 > 40 ? 40 : n + 3}) {
        const int kk = k == 0 ? 1 : k;
        std::vector<uint32_t> a(static_cast<size_t>(nq) * kk * 2), b(a.size());
        orc_knn_brute(kind, items.data(), n, pts.data(), nq, k, a.data(), 2);
        orc_bvh_knn(bvh, pts.data(), nq, k, b.data(), 2);
        expect(a == b, "bvh knn vs brute", kind);
        double s = 0.0;
        if (k == 0) orc_closest_hash(a.data(), nq, &s); else orc_knn_hash(a.data(), nq, kk, &s);
    }
    for (float r : {-1.0f, 0.0f, 2.5f, 9.0f}) {
        std::vector<int> ca(nq), cb(nq);
        orc_radius_brute(kind, items.data(), n, pts.data(), nq, r, ca.data(), nullptr, nullptr, 2);
        orc_bvh_radius(bvh, pts.data(), nq, r, cb.data(), nullptr, nullptr, 2);
        expect(ca == cb, "bvh radius counts vs brute", kind);
        const std::vector<int> off = offsetsOf(ca);
        std::vector<uint64_t> ka(static_cast<size_t>(off.back()) + 1), kb(ka.size());
        orc_radius_brute(kind, items.data(), n, pts.data(), nq, r, ca.data(), off.data(), ka.data(), 2);
        orc_bvh_radius(bvh, pts.data(), nq, r, cb.data(), off.data(), kb.data(), 2);
        expect(ka == kb, "bvh radius keys vs brute", kind);
        double s = 0.0;
        orc_csr_hash(off.data(), ka.data(), nq, &s);
    }
    orc_bvh_free(bvh);

    if (kind == 1) {
        void* kd = orc_kd_new(items.data(), n);
        std::vector<uint32_t> a(static_cast<size_t>(nq) * 8 * 2), b(a.size());
        orc_knn_brute(kind, items.data(), n, pts.data(), nq, 8, a.data(), 1);
        orc_kd_knn(kd, pts.data(), nq, 8, b.data(), 1);
        expect(a == b, "kd knn vs brute", kind);
        std::vector<int> ca(nq), cb(nq);
        orc_radius_brute(kind, items.data(), n, pts.data(), nq, 3.0f, ca.data(), nullptr, nullptr, 1);
        orc_kd_radius(kd, pts.data(), nq, 3.0f, cb.data(), nullptr, nullptr, 1);
        expect(ca == cb, "kd radius counts vs brute", kind);
        orc_kd_free(kd);
    }

    std::vector<unsigned char> nodes(static_cast<size_t>(2 * n - 1) * 32);
    std::vector<int> ids(n);
    const int depth = orc_lbvh_nodes(kind, items.data(), n, nodes.data(), ids.data());
    expect(depth > 0 && depth < 64, "lbvh depth", kind);
    std::vector<int> seen(n, 0);
    for (int id : ids) {
        if (id >= 0 && id < n) ++seen[id];
    }
    bool perm = true;
    for (int c : seen) perm &= c == 1;
    expect(perm, "lbvh leaf order is a permutation", kind);
    orc_fnv_bytes(nodes.data(), nodes.size());

    // leaf math probes on the scene's items
    for (int i = 0; i < n; i += 97) {
        const float* it = &items[static_cast<size_t>(i) * floatsPer(kind)];
        const float* p = &pts[static_cast<size_t>(i % nq) * 3];
        if (kind == 0) {
            int region = 0;
            float b1 = 0, b2 = 0, t = 0;
            expect(orc_tri_closest(p, it, &region, &b1, &b2) >= 0.0f, "triangle distance", kind);
            const float ray[8] = {p[0], p[1], p[2], it[0] - p[0], it[1] - p[1], it[2] - p[2], 1e-6f, INFINITY};
            orc_tri_intersect(ray, it, &t);
        } else if (kind == 2) {
            float box[6], rootT = 0;
            int inside = 0;
            orc_gauss_bounds(it, box);
            const float d = orc_gauss_dist2(p, it, &rootT, &inside);
            expect(d >= orc_box_dist2(p, box), "Gaussian distance bounded below by its box", kind);
            float tEnter = 0;
            const float ray[8] = {p[0], p[1], p[2], 0.0f, 0.0f, 1.0f, 1e-6f, INFINITY};
            orc_ray_box(ray, box, INFINITY, &tEnter);
        }
    }
}

void checkRays(std::mt19937& rng) {
    const int n = 3000, nr = 2000;
    const std::vector<float> tris = scene(0, n, rng);
    std::uniform_real_distribution<float> o(-30.0f, 30.0f), d(-1.0f, 1.0f);
    std::vector<float> rays(static_cast<size_t>(nr) * 8);
    for (int i = 0; i < nr; ++i) {
        float* r = &rays[static_cast<size_t>(i) * 8];
        r[0] = o(rng), r[1] = o(rng), r[2] = o(rng), r[3] = d(rng), r[4] = d(rng), r[5] = d(rng);
        if (i % 50 == 0) r[3] = r[4] = 0.0f;  // axes with |d| < 1e-20
        r[6] = 1e-6f;
        r[7] = i % 7 == 0 ? 5.0f : INFINITY;
    }
    std::vector<uint64_t> a(nr), b(nr);
    orc_ray_brute(tris.data(), n, rays.data(), nr, a.data(), 2);
    void* bvh = orc_bvh_new(0, tris.data(), n);
    orc_bvh_ray(bvh, rays.data(), nr, b.data(), 2);
    orc_bvh_free(bvh);
    expect(a == b, "bvh rays vs brute", 0);
    uint64_t hits = 0;
    double sumT = 0.0;
    orc_ray_hash(a.data(), nr, &hits, &sumT);
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    for (int kind = 0; kind < 3; ++kind) {
        checkKind(kind, 2, 50, rng);     // the smallest tree
        checkKind(kind, 37, 200, rng);   // k above the item count
        checkKind(kind, 4000, 600, rng);
    }
    checkRays(rng);
    std::printf("sanitized oracle run: %d mismatches\n", gFailures);
    return gFailures == 0 ? 0 : 1;
}
