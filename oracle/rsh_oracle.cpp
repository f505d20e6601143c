// rsh_oracle.cpp -- CPU restatement of the ray-space-hierarchy (CRSH) pipeline.
//
// TEST INFRASTRUCTURE ONLY.  This is the correctness oracle: a host C++
// restatement of the kernels in projects/CUDA-Engine/source/RayTracer.cu and
// of the frame driver in projects/CUDA-Engine/source/main.cpp.  Only tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
//
// Citations: RT = projects/CUDA-Engine/source/RayTracer.cu,
//            MAIN = projects/CUDA-Engine/source/main.cpp.
//
// How each part is pinned (checks live in tests/test_oracle_*.py; each check
// is an independent float64 / closed-form / brute-force model, and the
// shading, surface and ray-creation checks also run mutated models to show
// that a plausible mistake would fail them):
//   keys            closed forms of the atan2(v, v) bins + numpy float32
//                   evaluation of the same expressions; octahedral keys against
//                   a numpy restatement and round-trip decode bounds; camera-ray
//                   keys against an explicit bit interleave and the 4x2-leaf /
//                   8x8-node grouping they promise
//   trim            == stable copy_if of live slots
//   compress/sort/  == run-length encoding / std::stable_sort /
//   decompress         stable sort of (key, value) by key
//   hierarchy       containment invariants (every ray origin inside its leaf
//                   sphere, every direction inside its leaf cone, child inside
//                   parent), special cases (single ray: radius 0, spread 0)
//   traversal       closed-form candidate sets == literal flag/scan
//                   transcription; brute force on tiny scenes
//   intersection    Moller-Trumbore known answers (centroid, edges, parallel,
//                   behind origin); brute-force any/closest hit over all
//                   triangles agrees where culling is conservative
//   ray creation    shadow / reflection wrappers and the frame's camera,
//                   shadow, reflection and bounce rays against float64
//                   (tests/test_oracle_shading.py)
//   surface, shading G-buffer against a float64 brute-force surface stage;
//                   RgbToInt known answers (clamp, truncation, NaN, channel
//                   order); the Blinn-Phong term against float64 and closed
//                   forms (N.L = 1 with the eye at the light: att =
//                   1/((0.75 + 0.0005 d) + 0.00005 d^2), specular 1); frame
//                   pixels, including the packed-integer add x Ns of two
//                   bounces with bounce shadows, against float64
//   object update   UpdateVertex / UpdateBoundingSphere against numpy 4x4
//                   products on random transforms
#include "oracle_api.h"
#include "oracle_math.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace oracle;

namespace {

const float EPS = 0.01f;                       // RT:34
const float HALF_PI_F = 3.1415927f / 2.0f;     // Constants.h:2-3
const float PI_F = 3.1415927f;
const uint32_t MISS32 = 0xFFFFFFFFu;
const uint64_t MISS64 = 0xFFFFFFFFFFFFFFFFull;

inline f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
inline f4 ld4(const float* p) { return f4{p[0], p[1], p[2], p[3]}; }
inline void st3(float* p, f3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
inline void st4(float* p, f4 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w; }

// Per-stage wall time of the frame being rendered (std::chrono::steady_clock,
// the counterpart of the reference's per-stage timers, TestManager.cpp:
// 284-309): create, trim, compress, sort, decompress, hierarchy, traversal,
// shading.  One frame renders at a time (the oracle is not re-entrant).
enum { SC_CREATE, SC_TRIM, SC_COMPRESS, SC_SORT, SC_DECOMPRESS, SC_HIERARCHY, SC_TRAVERSE, SC_SHADE, SC_STAGES };
double g_stage_ms[SC_STAGES];
struct StageClock {
    int stage;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit StageClock(int s) : stage(s) {}
    void lap(int next) {
        const auto t = std::chrono::steady_clock::now();
        g_stage_ms[stage] += std::chrono::duration<double, std::milli>(t - t0).count();
        t0 = t;
        stage = next;
    }
    ~StageClock() { lap(stage); }
};

// ------------------------------------------------------------------ keys ---
// RT:147-168 CreateShadowRayIndex
uint32_t shadow_key_ref(f3 o, f3 d, uint32_t l) {
    (void)o;
    uint32_t index = (l + 1u) << 30;
    index += clampu(f2u_sat((atan2_diag(d.x) + HALF_PI_F) * 5.0f), 0u, 31u) << 15;
    index += clampu(f2u_sat(acos_f(d.z) * 5.0f), 0u, 31u);
    index++;
    return index;
}

// RT:170-198 CreateReflectionRayIndex
uint32_t reflection_key_ref(f3 o, f3 d) {
    float distance = length(o);
    uint32_t index = clampu(f2u_sat((atan2_diag(o.x) + HALF_PI_F) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u_sat(acos_f(o.z / distance) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u_sat((atan2_diag(d.x) + HALF_PI_F) * 5.0f), 0u, 30u);
    index = (index << 5) | clampu(f2u_sat(acos_f(d.z) * 5.0f), 0u, 30u);
    index++;
    return index;
}

// RT:200-227 CreateRefractionRayIndex: the origin's distance (7-bit field
// shifted by 26 in a 32-bit word, so its top bits are lost), then the four
// fields of the reflection key.
uint32_t refraction_key_ref(f3 o, f3 d) {
    float distance = length(o);
    uint32_t index = clampu(f2u_sat(distance), 0u, 128u);
    index = (index << 7) | clampu(f2u_sat((atan2_diag(o.x) + HALF_PI_F) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u_sat(acos_f(o.z / distance) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u_sat((atan2_diag(d.x) + HALF_PI_F) * 5.0f), 0u, 30u);
    index = (index << 5) | clampu(f2u_sat(acos_f(d.z) * 5.0f), 0u, 30u);
    index++;
    return index;
}

// helper_math.h:1422-1431 refract, pow(x, 2) as x * x; zero on total internal
// reflection.
f3 refract(f3 i, f3 n, float ior) {
    const float c = dot(i, n);
    const float k = 1.0f - (ior * ior) * (1.0f - c * c);
    if (k < 0.0f) return f3{0.0f, 0.0f, 0.0f};
    return i * ior - n * (ior * c + std::sqrt(k));
}

// Octahedral key (configuration hash; no counterpart in the reference):
//   [type:2][cell z|y|x : 3*cb][u|v : 2*db]   (layout 0)
//   [type:2][z y x u v bits interleaved, most significant level first] (layout 2)
//   [type:2][u|v : 2*db][cell z|y|x : 3*cb]   (layout 1)
// cell = min(sat((o - scene_min) * cell_scale), 2^cb - 1) per axis;
// (u, v) = octahedral map of d onto [0, 2^db)^2.
uint32_t key_oct(const oracle_config& c, f3 o, f3 d, uint32_t type) {
    const uint32_t cb = c.cell_bits, db = c.dir_bits;
    const uint32_t cmax = (1u << cb) - 1u, dmax = (1u << db) - 1u;
    uint32_t cx = std::min(f2u_sat((o.x - c.scene_min[0]) * c.cell_scale[0]), cmax);
    uint32_t cy = std::min(f2u_sat((o.y - c.scene_min[1]) * c.cell_scale[1]), cmax);
    uint32_t cz = std::min(f2u_sat((o.z - c.scene_min[2]) * c.cell_scale[2]), cmax);
    float s = (std::fabs(d.x) + std::fabs(d.y)) + std::fabs(d.z);
    float px = d.x / s, py = d.y / s;
    if (d.z < 0.0f) {
        float ax = std::fabs(px), ay = std::fabs(py);
        float nx = (1.0f - ay) * (px >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - ax) * (py >= 0.0f ? 1.0f : -1.0f);
        px = nx; py = ny;
    }
    const float fs = (float)(1u << db);
    uint32_t u = std::min(f2u_sat((px * 0.5f + 0.5f) * fs), dmax);
    uint32_t v = std::min(f2u_sat((py * 0.5f + 0.5f) * fs), dmax);
    uint32_t cell = (cz << (2 * cb)) | (cy << cb) | cx;
    uint32_t dir = (u << db) | v;
    uint32_t body;
    if (c.key_layout == 2) {
        // bit-interleaved, most significant level first: z y x u v at each level
        body = 0;
        const uint32_t top = cb > db ? cb : db;
        for (uint32_t i = top; i-- > 0;) {
            if (i < cb) body = (body << 3) | (((cz >> i) & 1u) << 2) | (((cy >> i) & 1u) << 1) | ((cx >> i) & 1u);
            if (i < db) body = (body << 2) | (((u >> i) & 1u) << 1) | ((v >> i) & 1u);
        }
    } else {
        body = c.key_layout == 0 ? ((cell << (2 * db)) | dir) : ((dir << (3 * cb)) | cell);
    }
    return (type << (3 * cb + 2 * db)) | body;
}

// ---------------------------------------------------------- geometry -------
// RT:257-280 RayTriangleIntersection (Moller-Trumbore).  The reference takes
// the reciprocal with __fdividef; here it is the correctly rounded 1.0f/det.
float ray_triangle(f3 o, f3 d, f3 v0, f3 e1, f3 e2) {
    f3 tvec = o - v0;
    f3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    det = 1.0f / det;
    float u = dot(tvec, pvec) * det;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    f3 qvec = cross(tvec, e1);
    float v = dot(d, qvec) * det;
    if (v < 0.0f || (u + v) > 1.0f) return -1.0f;
    return dot(e2, qvec) * det;
}

// RT:283-298 SphereNodeIntersection
bool sphere_node(f4 sphere, f4 cone, f4 tri, float cosine, float tangent) {
    if (cone.w == HALF_PI_F) return true;
    f3 cd = xyz(cone), sc = xyz(sphere), tc = xyz(tri);
    f3 s2t = tc - sc;
    f3 proj = dot(s2t, cd) * cd;
    return (dot(s2t, cd) * tangent + (sphere.w + tri.w) / cosine) >= length(s2t - proj);
}

// RT:301-376 CreateTriangleBoundingSphere
f4 triangle_sphere(f3 a, f3 b, f3 c) {
    float dotABAB = dot(b - a, b - a);
    float dotABAC = dot(b - a, c - a);
    float dotACAC = dot(c - a, c - a);
    float d = 2.0f * (dotABAB * dotACAC - dotABAC * dotABAC);
    f3 ref = a;
    f3 center;
    float radius;
    if (std::fabs(d) <= EPS) {
        float dAB = length(a - b), dBC = length(b - c), dCA = length(c - a);
        if (dAB > dBC) {
            if (dAB > dCA) { center = (a + b) * 0.5f; radius = dAB * 0.5f; }
            else { center = (a + c) * 0.5f; radius = dCA * 0.5f; }
        } else {
            if (dBC > dCA) { center = (b + c) * 0.5f; radius = dBC * 0.5f; }
            else { center = (a + c) * 0.5f; radius = dCA * 0.5f; }
        }
        return mk4(center, radius);
    }
    float s = (dotABAB * dotACAC - dotACAC * dotABAC) / d;
    float t = (dotACAC * dotABAB - dotABAB * dotABAC) / d;
    if (s <= 0.0f) center = (a + c) * 0.5f;
    else if (t <= 0.0f) center = (a + b) * 0.5f;
    else if (s + t >= 1.0f) { center = (b + c) * 0.5f; ref = b; }
    else center = (a + s * (b - a)) + t * (c - a);
    radius = length(center - ref);
    return mk4(center, radius);
}

// RT:379-406 CreateHierarchyCone0 (mode 0), and the guarded variant (mode 1):
// unit ray direction, dot clamped to [-1, 1], parallel and antipodal cases
// resolved without NaN, spread kept <= pi/2.
f4 cone0(f4 cone, f4 ray, uint32_t mode) {
    if (mode == 0) {
        f3 cd = xyz(cone), rd = xyz(ray);
        float spread = acos_f(dot(cd, rd));
        if (cone.w > spread) return cone;
        f3 q = normalize(dot(cd, rd) * cd - rd);
        f3 e = cd * cos_f(cone.w) + q * sin_f(cone.w);
        f3 nd = normalize(rd + e);
        float ns = acos_f(dot(nd, rd));
        return mk4(nd, ns);
    }
    f3 cd = xyz(cone), rd = normalize(xyz(ray));
    float c = clampf(dot(cd, rd), -1.0f, 1.0f);
    float spread = acos_f(c);
    if (cone.w >= spread) return cone;
    if (spread + cone.w >= PI_F) return mk4(cd, HALF_PI_F);
    f3 qv = c * cd - rd;
    float qq = dot(qv, qv);
    if (!(qq > 0.0f)) return mk4(cd, HALF_PI_F);
    f3 q = qv * (1.0f / sqrtf(qq));
    f3 e = cd * cos_f(cone.w) + q * sin_f(cone.w);
    f3 sum = rd + e;
    float ss = dot(sum, sum);
    if (!(ss > 0.0f)) return mk4(cd, HALF_PI_F);
    f3 nd = sum * (1.0f / sqrtf(ss));
    float ns = acos_f(clampf(dot(nd, rd), -1.0f, 1.0f));
    return mk4(nd, fmin_(ns, HALF_PI_F));
}

// RT:408-425 CreateHierarchyConeN (mode 1 guards the antipodal case).
f4 coneN(f4 c1, f4 c2, uint32_t mode) {
    f3 d1 = xyz(c1), d2 = xyz(c2);
    if (mode == 0) {
        float spread = acos_f(dot(d1, d2)) * 0.5f;
        f3 d = normalize(d1 + d2);
        float s = clampf(spread + fmax_(c1.w, c2.w), 0.0f, HALF_PI_F);
        return mk4(d, s);
    }
    float spread = acos_f(clampf(dot(d1, d2), -1.0f, 1.0f)) * 0.5f;
    f3 sum = d1 + d2;
    float ss = dot(sum, sum);
    if (!(ss > 0.0f)) return mk4(d1, HALF_PI_F);
    f3 d = sum * (1.0f / sqrtf(ss));
    float s = clampf(spread + fmax_(c1.w, c2.w), 0.0f, HALF_PI_F);
    return mk4(d, s);
}

// RT:427-446 CreateHierarchySphere
f4 sphere_union(f4 s1, f4 s2) {
    f3 c1 = xyz(s1), c2 = xyz(s2);
    float dist = length(c2 - c1);
    if (dist + s2.w <= s1.w) return s1;
    if (dist + s1.w <= s2.w) return s2;
    f3 c = (c1 + c2) * 0.5f;
    float r = dist * 0.5f + fmax_(s1.w, s2.w);
    return mk4(c, r);
}

// RT:100-107 RgbToInt
uint32_t rgb_to_int(float r, float g, float b) {
    r = clampf(r, 0.0f, 255.0f);
    g = clampf(g, 0.0f, 255.0f);
    b = clampf(b, 0.0f, 255.0f);
    return (uint32_t)((int)r | ((int)g << 8) | ((int)b << 16));
}

// RT:1770 reference closest-hit packing (mode 1) / exact packing (mode 0).
inline uint64_t pack_hit(float t, uint32_t tri, uint32_t mode) {
    if (mode == 0) return ((uint64_t)fbits(t) << 32) | (uint64_t)tri;
    return (uint64_t)((f2u_sat(t * 10.0f + 1.0f) << 20) + tri);
}
inline bool is_miss(uint64_t rec, uint32_t mode) { return mode == 0 ? rec == MISS64 : rec == MISS32; }
inline uint32_t hit_tri(uint64_t rec, uint32_t mode) {
    return mode == 0 ? (uint32_t)(rec & 0xFFFFFFFFull) : (uint32_t)(rec & 0x000FFFFFull);
}

inline void atomic_min64(uint64_t* p, uint64_t v) {
    uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

// ------------------------------------------------------------ scene -------
struct Scene {
    const oracle_scene* s;
    std::vector<f4> tri_sphere;      // per triangle (A12 evaluated once)
    std::vector<f3> v0, e1, e2;      // per triangle
    explicit Scene(const oracle_scene* sc) : s(sc) {
        uint32_t T = sc->triangle_total;
        tri_sphere.resize(T); v0.resize(T); e1.resize(T); e2.resize(T);
        for (uint32_t t = 0; t < T; t++) {
            f3 a = ld3(sc->tri_pos + 12 * t), b = ld3(sc->tri_pos + 12 * t + 4), c = ld3(sc->tri_pos + 12 * t + 8);
            tri_sphere[t] = triangle_sphere(a, b, c);
            v0[t] = a; e1[t] = b - a; e2[t] = c - a;
        }
    }
    f4 obj_sphere(uint32_t k) const {
        const float* p = s->object_spheres + 6 * k;
        return f4{p[0], p[1], p[2], p[3]};
    }
    uint32_t obj_first(uint32_t k) const { return (uint32_t)s->object_spheres[6 * k + 4]; }
    uint32_t obj_last(uint32_t k) const { return (uint32_t)s->object_spheres[6 * k + 5]; }
};

struct GBuf {
    uint32_t W = 0, H = 0;
    std::vector<f4> pos, nrm, dif, spec;
    void init(uint32_t w, uint32_t h) {
        W = w; H = h;
        pos.assign((size_t)w * h, f4{0, 0, 0, 0}); nrm = pos; dif = pos; spec = pos;
    }
};

struct Batch {                 // slot-indexed ray batch (RT Ray layout, T1/T2)
    uint32_t S = 0;
    std::vector<f3> rays;        // 2*S: origin at 2i, direction at 2i+1
    std::vector<uint32_t> keys, values, flags;
    void init(uint32_t s) {
        S = s;
        rays.assign(2 * (size_t)s, f3{0, 0, 0});
        keys.assign(s, 0); values.assign(s, 0); flags.assign(s, 1);
    }
};

struct Hier {
    uint32_t depth = 0;
    std::vector<uint32_t> count, offset;
    std::vector<f4> nodes;         // T3: sphere at 2g, cone at 2g+1
    std::vector<float> cosv, tanv; // per node, cos/tan of the spread
};

struct PassStats { uint32_t N = 0, C = 0; uint64_t top = 0, leaf = 0, tests = 0; };

// --------------------------------------------------------- stages ---------
// RT:844-884 + RT:2302-2319: stable compaction of flag==0 slots.
void trim(const Batch& b, std::vector<uint32_t>& tk, std::vector<uint32_t>& tv) {
    tk.clear(); tv.clear();
    for (uint32_t i = 0; i < b.S; i++)
        if (b.flags[i] == 0) { tk.push_back(b.keys[i]); tv.push_back(b.values[i]); }
}

// RT:886-991 (compress), RT:2401 (stable radix sort), RT:993-1059 (decompress).
void compress_sort_decompress(const std::vector<uint32_t>& tk, const std::vector<uint32_t>& tv,
                              std::vector<uint32_t>& sk, std::vector<uint32_t>& sv, uint32_t* chunks,
                              uint32_t* out_bases = nullptr, uint32_t* out_ckeys = nullptr) {
    const uint32_t N = (uint32_t)tk.size();
    StageClock clk(SC_COMPRESS);
    std::vector<uint32_t> bases, sizes, ckeys, cvals;
    for (uint32_t x = 0; x < N; x++)
        if (x == 0 || tk[x] != tk[x - 1]) { bases.push_back(x); ckeys.push_back(tk[x]); }
    const uint32_t C = (uint32_t)bases.size();
    sizes.resize(C);
    for (uint32_t c = 0; c < C; c++) { sizes[c] = (c + 1 < C ? bases[c + 1] : N) - bases[c]; cvals.push_back(c); }
    if (out_bases) std::copy(bases.begin(), bases.end(), out_bases);
    if (out_ckeys) std::copy(ckeys.begin(), ckeys.end(), out_ckeys);
    clk.lap(SC_SORT);
    std::vector<uint32_t> perm(C);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return ckeys[a] < ckeys[b]; });
    clk.lap(SC_DECOMPRESS);
    sk.resize(N); sv.resize(N);
    uint32_t pos = 0;
    for (uint32_t i = 0; i < C; i++) {
        uint32_t c = cvals[perm[i]];
        for (uint32_t j = 0; j < sizes[c]; j++) { sk[pos] = tk[bases[c] + j]; sv[pos] = tv[bases[c] + j]; pos++; }
    }
    if (chunks) *chunks = C;
}

// RT:1061-1150 + RT:2469-2535.  Generalised leaf / branch / depth.  The leaf
// loop stops at the last live ray (the reference reads one slot past it,
// RT:1090 -- a documented deviation).
void build_hierarchy(const Batch& b, const std::vector<uint32_t>& sv, uint32_t N, uint32_t leaf,
                     uint32_t branch, uint32_t depth, float r0, float s0, uint32_t mode, Hier& h) {
    h.depth = depth;
    h.count.assign(depth, 0); h.offset.assign(depth, 0);
    uint32_t n = (N + leaf - 1) / leaf, total = 0;
    for (uint32_t k = 0; k < depth; k++) {
        h.count[k] = n; h.offset[k] = total; total += n;
        n = (n + branch - 1) / branch;
    }
    h.nodes.assign(2 * (size_t)total, f4{0, 0, 0, 0});
    for (uint32_t j = 0; j < h.count[0]; j++) {
        uint32_t slot = sv[j * leaf];
        f4 sphere = mk4(b.rays[2 * (size_t)slot], r0);
        f3 d0 = b.rays[2 * (size_t)slot + 1];
        f4 cone = mk4(mode == 0 ? d0 : normalize(d0), s0);
        for (uint32_t i = 1; i < leaf; i++) {
            uint32_t p = j * leaf + i;
            if (p >= N) break;
            uint32_t sl = sv[p];
            sphere = sphere_union(sphere, mk4(b.rays[2 * (size_t)sl], r0));
            cone = cone0(cone, mk4(b.rays[2 * (size_t)sl + 1], s0), mode);
        }
        h.nodes[2 * (size_t)j] = sphere;
        h.nodes[2 * (size_t)j + 1] = cone;
    }
    for (uint32_t k = 1; k < depth; k++) {
        uint32_t ro = h.offset[k - 1], wo = h.offset[k];
        for (uint32_t j = 0; j < h.count[k]; j++) {
            f4 sphere = h.nodes[2 * (size_t)(ro + j * branch)];
            f4 cone = h.nodes[2 * (size_t)(ro + j * branch) + 1];
            for (uint32_t i = 1; i < branch; i++) {
                uint32_t c = j * branch + i;
                if (c >= h.count[k - 1]) break;
                sphere = sphere_union(sphere, h.nodes[2 * (size_t)(ro + c)]);
                cone = coneN(cone, h.nodes[2 * (size_t)(ro + c) + 1], mode);
            }
            h.nodes[2 * (size_t)(wo + j)] = sphere;
            h.nodes[2 * (size_t)(wo + j) + 1] = cone;
        }
    }
    h.cosv.resize(total); h.tanv.resize(total);
    for (uint32_t g = 0; g < total; g++) {
        h.cosv[g] = cos_f(h.nodes[2 * (size_t)g + 1].w);
        h.tanv[g] = tan_f(h.nodes[2 * (size_t)g + 1].w);
    }
}

// Object culling at the top level (RT:1152-1313, CRSH only), restated in
// closed form: triangle t of object o is a candidate of top node n iff the
// node intersects o's sphere, except for the triangles of a batch that o
// covers from before the batch and whose last index equals the batch's upper
// bound -- RT:1192-1244 has no branch for that case (lastTri == upper), so
// those pairs are never produced.  Objects tile [0, T) in order.
void excluded_range(const Scene& sc, uint32_t k, uint32_t Tb, uint32_t* lo, uint32_t* hi) {
    uint32_t y = sc.obj_first(k), z = sc.obj_last(k);
    *lo = *hi = 0;
    if (Tb == 0 || z == 0 || z % Tb != 0) return;
    uint32_t lower = z - Tb;
    if (y < lower) { *lo = lower; *hi = z; }
}

struct Traverser {
    const Scene& sc;
    const Hier& h;
    uint32_t algorithm, branch, Tb;
    // callback per surviving (leaf, triangle)
    template <class F>
    void descend(uint32_t level, uint32_t parent, uint32_t t, F& f, PassStats& st) const {
        const f4 tri = sc.tri_sphere[t];
        for (uint32_t i = 0; i < branch; i++) {
            uint32_t c = parent * branch + i;
            if (c >= h.count[level]) break;
            uint32_t g = h.offset[level] + c;
            if (!sphere_node(h.nodes[2 * (size_t)g], h.nodes[2 * (size_t)g + 1], tri, h.cosv[g], h.tanv[g])) continue;
            if (level == 0) { st.leaf++; f(c, t); }
            else descend(level - 1, c, t, f, st);
        }
    }
    template <class F>
    void top_node(uint32_t j, F& f, PassStats& st) const {
        const uint32_t top = h.depth - 1;
        const uint32_t g = h.offset[top] + j;
        const f4 sphere = h.nodes[2 * (size_t)g], cone = h.nodes[2 * (size_t)g + 1];
        const float cs = h.cosv[g], tn = h.tanv[g];
        auto tri_test = [&](uint32_t t) {
            st.top++;
            if (!sphere_node(sphere, cone, sc.tri_sphere[t], cs, tn)) return;
            if (top == 0) { st.leaf++; f(j, t); }
            else descend(top - 1, j, t, f, st);
        };
        if (algorithm == 1) {
            for (uint32_t t = 0; t < sc.s->triangle_total; t++) tri_test(t);
            return;
        }
        for (uint32_t k = 0; k < sc.s->object_total; k++) {
            if (!sphere_node(sphere, cone, sc.obj_sphere(k), cs, tn)) continue;
            uint32_t lo, hi;
            excluded_range(sc, k, Tb, &lo, &hi);
            uint32_t y = sc.obj_first(k), z = sc.obj_last(k);
            for (uint32_t t = y; t <= z; t++) {
                if (t >= lo && t < hi) continue;
                tri_test(t);
            }
        }
    }
};

// Build one batch's hierarchy and run traversal + intersection.
// kind 0: closest hit (RT:1711-1774), kind 1: shadow any-hit (RT:1498-1614).
// tmax (any hit, optional): the caller's per-slot limit instead of |o - P| - 2 eps.
// Where a pass's keys, flags, sorted order, packets and hierarchy nodes go
// (oracle_frame_out's pass_keys ... pass_nodes; NULL: not wanted).
struct PassDump {
    uint32_t* keys = nullptr;
    uint32_t* flags = nullptr;
    uint32_t* order = nullptr;
    uint32_t* bases = nullptr;
    uint32_t* ckeys = nullptr;
    float* nodes = nullptr;
};

void trace(const Scene& sc, const oracle_config& cfg, const Batch& b, uint32_t kind,
           const GBuf* shadow_g, uint32_t Lmax, std::vector<uint64_t>& result, PassStats& st,
           int threads, const float* tmax = nullptr, const PassDump* dump = nullptr) {
    (void)threads;
    std::vector<uint32_t> tk, tv, sk, sv;
    if (dump && dump->keys) std::copy(b.keys.begin(), b.keys.end(), dump->keys);
    if (dump && dump->flags) std::copy(b.flags.begin(), b.flags.end(), dump->flags);
    {
        StageClock clk(SC_TRIM);
        trim(b, tk, tv);
        st.N = (uint32_t)tk.size();
        result.assign(b.S, kind == 0 ? (cfg.hit_mode == 0 ? MISS64 : (uint64_t)MISS32) : 0ull);
    }
    if (st.N == 0) return;
    if (cfg.algorithm == 0)
        compress_sort_decompress(tk, tv, sk, sv, &st.C, dump ? dump->bases : nullptr, dump ? dump->ckeys : nullptr);
    else { sk = tk; sv = tv; st.C = st.N; }
    Hier h;
    {
        StageClock clk(SC_HIERARCHY);
        build_hierarchy(b, sv, st.N, cfg.leaf, cfg.branch, cfg.depth, cfg.initial_radius, cfg.initial_spread,
                        cfg.cone_mode, h);
    }
    if (dump && dump->order) std::copy(sv.begin(), sv.end(), dump->order);
    if (dump && dump->nodes)
        for (size_t i = 0; i < h.nodes.size(); i++) st4(dump->nodes + 4 * i, h.nodes[i]);
    StageClock clk(SC_TRAVERSE);
    Traverser tr{sc, h, cfg.algorithm, cfg.branch, cfg.triangle_batch};
    const uint32_t N = st.N, leaf = cfg.leaf, W = cfg.width;
    uint64_t* res = result.data();
    const uint32_t ntop = h.count[cfg.depth - 1];
    uint64_t top = 0, lf = 0, tests = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : top, lf, tests)
    for (uint32_t j = 0; j < ntop; j++) {
        PassStats local;
        uint64_t ltests = 0;
        auto hitf = [&](uint32_t c, uint32_t t) {
            const f3 v0 = sc.v0[t], e1 = sc.e1[t], e2 = sc.e2[t];
            for (uint32_t i = 0; i < leaf; i++) {
                uint32_t p = c * leaf + i;
                if (p >= N) return;
                uint32_t slot = sv[p];
                f3 o = b.rays[2 * (size_t)slot], d = b.rays[2 * (size_t)slot + 1];
                float th = ray_triangle(o + d * EPS, d, v0, e1, e2);
                ltests++;
                if (kind == 1) {
                    float lim;
                    if (tmax) {
                        lim = tmax[slot];
                    } else {
                        const uint32_t i = cfg.soft_shadows ? slot >> 4 : slot;
                        uint32_t x = i % W, y = i / (W * Lmax);
                        f3 P = xyz(shadow_g->pos[(size_t)x + (size_t)y * W]);
                        lim = length(o - P) - EPS * 2.0f;
                    }
                    if (th > EPS && th < lim)
                        __atomic_store_n(&res[slot], (uint64_t)MISS32, __ATOMIC_RELAXED);
                } else if (th > EPS) {
                    atomic_min64(&res[slot], pack_hit(th, t, cfg.hit_mode));
                }
            }
        };
        tr.top_node(j, hitf, local);
        top += local.top; lf += local.leaf; tests += ltests;
    }
    st.top = top; st.leaf = lf; st.tests = tests;
}

// ---------------------------------------------------------- views --------
// Views of a batch laid out as one image of P = sum(w*h) pixels (view i from
// pixel off[i]); a single frame is one view of W x H.  The view id goes into
// the octahedral key at cfg.segment_shift when that is set.
struct ViewSet {
    std::vector<uint32_t> off, w, h;
    std::vector<f3> eye, D, R, U;
    uint32_t view_of(uint32_t p) const {
        uint32_t a = 0, b = (uint32_t)off.size();
        while (b - a > 1) {
            uint32_t m = (a + b) >> 1;
            if (off[m] <= p) a = m; else b = m;
        }
        return a;
    }
};

uint32_t view_key(const oracle_config& cfg, f3 o, f3 d, uint32_t type, uint32_t view) {
    uint32_t k = key_oct(cfg, o, d, type);
    if (cfg.segment_shift) k |= view << cfg.segment_shift;
    return k;
}

// ---------------------------------------------------- ray creation --------
// Soft shadows (RT:1550-1603, compiled out by Constants.h:32): grid point c
// (i = c % 4, j = c / 4) about the light at offsets (a - 2, b - 2) / 8 of two
// unnormalised plane axes, a = i + (i >= 2), b = j + (j >= 2).  The first
// axis crosses the coordinate axis picked by RT:1556-1565's test against
// min(|x|, max(|y|, |z|)) with the shadow direction d; the second is d x it.
void soft_axes(f3 d, f3& a0, f3& a1) {
    const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    const float m = fmin_(ax, fmax_(ay, az));
    if (ax == m) a0 = cross(f3{1.0f, 0.0f, 0.0f}, d);
    else if (ay == m) a0 = cross(f3{0.0f, 1.0f, 0.0f}, d);
    else a0 = cross(f3{0.0f, 0.0f, 1.0f}, d);
    a1 = cross(d, a0);
}

f3 soft_point(f3 o, f3 a0, f3 a1, uint32_t c) {
    const int i = (int)(c & 3u), j = (int)(c >> 2);
    const int a = i + (i >= 2 ? 1 : 0), bb = j + (j >= 2 ? 1 : 0);
    return (o + (a0 * (float)(a - 2)) * 0.125f) + (a1 * (float)(bb - 2)) * 0.125f;
}

// RT:1680-1692: log2(1 + 2 s), s summing, over the open cells in cell order,
// 0.55/16 (cells 5, 6, 9, 10), 0.5375/16 (cells 0, 3, 12, 15) or 0.525/16.
float soft_factor(uint32_t occluded) {
    float s = 0.0f;
    for (uint32_t c = 0; c < 16; c++) {
        if ((occluded >> c) & 1u) continue;
        if (c == 5 || c == 6 || c == 9 || c == 10) s += 0.55f / 16.0f;
        else if (c == 0 || c == 3 || c == 12 || c == 15) s += 0.5375f / 16.0f;
        else s += 0.525f / 16.0f;
    }
    return log2_f(1.0f + s * 2.0f);
}

// RT:682-740 CreateShadowRays over slots i = x + W*(l + L*y).  With soft
// shadows every live slot i holds 16 rays, 16 i + c from grid point c towards
// the slot's surface point (RT:1568-1590's interpolated rays).
void create_shadow(const Scene& sc, const oracle_config& cfg, const GBuf& g, uint32_t Lmax, const ViewSet& vs,
                   Batch& b) {
    StageClock clk_(SC_CREATE);
    const uint32_t W = g.W, H = g.H;
    const uint32_t sh = cfg.soft_shadows ? 4u : 0u;
    b.init((W * H * Lmax) << sh);
#pragma omp parallel for schedule(static)
    for (int64_t rr = 0; rr < (int64_t)b.S; rr++) {
        const uint32_t r = (uint32_t)rr, i = r >> sh;
        uint32_t x = i % W, y = i / (W * Lmax), l = (i / W) % Lmax;
        f3 P = xyz(g.pos[(size_t)x + (size_t)y * W]);
        f3 N = normalize(xyz(g.nrm[(size_t)x + (size_t)y * W]));
        b.flags[r] = 1;
        if (length(P) != 0.0f) {
            f3 o = ld3(sc.s->light_pos + 4 * l);
            f3 d = normalize(P - o);
            float diffuse = fmax_(dot(d, -N), 0.0f);
            if (diffuse > 0.0f) {
                if (sh) {
                    f3 a0, a1;
                    soft_axes(d, a0, a1);
                    o = soft_point(o, a0, a1, r & 15u);
                    d = normalize(P - o);
                }
                b.rays[2 * (size_t)r] = o;
                b.rays[2 * (size_t)r + 1] = d;
                b.keys[r] = cfg.key_mode == 0 ? shadow_key_ref(o, d, l)
                                              : view_key(cfg, o, d, 1, vs.view_of(x + y * W));
                b.values[r] = r;
                b.flags[r] = 0;
            }
        }
    }
}

// RT:742-791 CreateReflectionRays, slot x + y*W*L; gap slots (L > 1) are
// marked empty.  With reflect_gate only materials whose diffuse.w > 0 reflect.
void create_reflection(const oracle_config& cfg, const GBuf& g, const ViewSet& vs, uint32_t Lmax, Batch& b) {
    StageClock clk_(SC_CREATE);
    const uint32_t W = g.W, H = g.H;
    b.init(W * H * Lmax);
#pragma omp parallel for schedule(static)
    for (int64_t pp = 0; pp < (int64_t)W * H; pp++) {
        uint32_t x = (uint32_t)(pp % W), y = (uint32_t)(pp / W);
        uint32_t slot = x + y * W * Lmax;
        f3 P = xyz(g.pos[pp]);
        f3 N = normalize(xyz(g.nrm[pp]));
        bool live = length(P) != 0.0f && (!cfg.reflect_gate || g.dif[pp].w > 0.0f);
        if (live) {
            const uint32_t v = vs.view_of((uint32_t)pp);
            f3 d = reflect(normalize(P - vs.eye[v]), normalize(N));
            b.rays[2 * (size_t)slot] = P;
            b.rays[2 * (size_t)slot + 1] = d;
            b.keys[slot] = cfg.key_mode == 0 ? reflection_key_ref(P, d) : view_key(cfg, P, d, 2, v);
            b.values[slot] = slot;
            b.flags[slot] = 0;
        }
    }
}

// Next bounce (RT:1995-2010): origin P, direction reflect(d, N) with the
// barycentric (unnormalised) normal; gated by material reflectivity.
void create_bounce(const oracle_config& cfg, const GBuf& g, const Batch& prev,
                   const std::vector<uint64_t>& rec, uint32_t Lmax, const ViewSet& vs, Batch& b) {
    StageClock clk_(SC_CREATE);
    const uint32_t W = g.W, H = g.H;
    b.init(W * H * Lmax);
#pragma omp parallel for schedule(static)
    for (int64_t pp = 0; pp < (int64_t)W * H; pp++) {
        uint32_t x = (uint32_t)(pp % W), y = (uint32_t)(pp / W);
        uint32_t slot = x + y * W * Lmax;
        if (is_miss(rec[slot], cfg.hit_mode)) continue;
        if (cfg.reflect_gate && !(g.dif[pp].w > 0.0f)) continue;
        f3 P = xyz(g.pos[pp]);
        f3 N = xyz(g.nrm[pp]);
        f3 d = reflect(prev.rays[2 * (size_t)slot + 1], N);
        b.rays[2 * (size_t)slot] = P;
        b.rays[2 * (size_t)slot + 1] = d;
        b.keys[slot] = cfg.key_mode == 0 ? reflection_key_ref(P, d) : view_key(cfg, P, d, 2, vs.view_of((uint32_t)pp));
        b.values[slot] = slot;
        b.flags[slot] = 0;
    }
}

// Primary rays are keyed by screen position, not by the hash (the reference
// rasterises primary visibility, so this stage has no reference order): 8x8
// pixel tiles in row-major order over the view, Morton order inside a tile
// (x bit 0, y bit 0, x bit 1, ...), the view id above segment_shift.  Sorted,
// a leaf of 8 rays is a 4x2 pixel block and a 64-ray node one tile.
uint32_t primary_key(uint32_t x, uint32_t y, uint32_t w, uint32_t view, uint32_t segment_shift) {
    const uint32_t lx = x & 7u, ly = y & 7u;
    const uint32_t m = (lx & 1u) | ((ly & 1u) << 1) | ((lx & 2u) << 1) | ((ly & 2u) << 2) | ((lx & 4u) << 2) |
                       ((ly & 4u) << 3);
    uint32_t k = (((y >> 3) * ((w + 7u) >> 3) + (x >> 3)) << 6) | m;
    if (segment_shift) k |= view << segment_shift;
    return k;
}

// Primary rays (configuration stage; formula of RT:667), per view: pixel p of
// view v is (x, y) within it, dir = D + R*(x/w - 0.5) + U*(y/h - 0.5).
void create_primary(const oracle_config& cfg, const ViewSet& vs, Batch& b) {
    StageClock clk_(SC_CREATE);
    const uint32_t P = cfg.width * cfg.height;
    b.init(P);
#pragma omp parallel for schedule(static)
    for (int64_t pp = 0; pp < (int64_t)P; pp++) {
        const uint32_t v = vs.view_of((uint32_t)pp);
        const uint32_t q = (uint32_t)pp - vs.off[v];
        uint32_t x = q % vs.w[v], y = q / vs.w[v];
        float a = (float)x / (float)vs.w[v] - 0.5f;
        float c = (float)y / (float)vs.h[v] - 0.5f;
        const f3 eye = vs.eye[v];
        f3 d = normalize((vs.D[v] + vs.R[v] * a) + vs.U[v] * c);
        b.rays[2 * pp] = eye;
        b.rays[2 * pp + 1] = d;
        b.keys[pp] = primary_key(x, y, vs.w[v], v, cfg.segment_shift);
        b.values[pp] = (uint32_t)pp;
        // Optional row band (a bounded sample of the frame for timing).
        bool in_band = cfg.row_end == 0 || (y >= cfg.row_begin && y < cfg.row_end);
        b.flags[pp] = in_band ? 0 : 1;
    }
}

// Hit -> surface record (RT:1807-1843): hit point o + d*(eps + t), area-ratio
// barycentric normal, material of the triangle's first vertex.  Primary
// G-buffers store the normalised normal (blinn_phong_fragment_shader.glsl:72).
void surface_from_hits(const Scene& sc, const oracle_config& cfg, const Batch& b,
                       const std::vector<uint64_t>& rec, uint32_t slot_stride, bool normalize_n, GBuf& g) {
    StageClock clk_(SC_SHADE);
    const uint32_t W = cfg.width, H = cfg.height;
    g.init(W, H);
#pragma omp parallel for schedule(static)
    for (int64_t pp = 0; pp < (int64_t)W * H; pp++) {
        uint32_t x = (uint32_t)(pp % W), y = (uint32_t)(pp / W);
        uint32_t slot = x + y * W * slot_stride;
        uint64_t r = rec[slot];
        if (is_miss(r, cfg.hit_mode)) continue;
        uint32_t t = hit_tri(r, cfg.hit_mode);
        const float* tp = sc.s->tri_pos + 12 * (size_t)t;
        const float* tn = sc.s->tri_nrm + 12 * (size_t)t;
        f3 v0 = ld3(tp), v1 = ld3(tp + 4), v2 = ld3(tp + 8);
        f3 o = b.rays[2 * (size_t)slot], d = b.rays[2 * (size_t)slot + 1];
        float th = ray_triangle(o + d * EPS, d, v0, v1 - v0, v2 - v0);
        f3 P = o + d * (EPS + th);
        float aABC = length(cross(v1 - v0, v2 - v0));
        float aPBC = length(cross(v1 - P, v2 - P));
        float aPCA = length(cross(v0 - P, v2 - P));
        f3 N = ((aPBC / aABC) * ld3(tn) + (aPCA / aABC) * ld3(tn + 4)) +
               ((1.0f - (aPBC / aABC)) - (aPCA / aABC)) * ld3(tn + 8);
        if (normalize_n) N = normalize(N);
        int32_t m = sc.s->tri_mat[3 * (size_t)t];
        g.pos[pp] = mk4(P, 1.0f);
        g.nrm[pp] = mk4(N, 0.0f);
        g.dif[pp] = ld4(sc.s->mat_diffuse + 4 * (size_t)m);
        g.spec[pp] = ld4(sc.s->mat_specular + 4 * (size_t)m);
    }
}

// Blinn-Phong term of one light (RT:1655-1702 and RT:1847-1876); soft: both
// components scaled last by `factor` (RT:1693-1696).
inline f3 light_term(const oracle_scene* s, uint32_t l, f3 P, f3 N, f3 eye, f4 Kd, f4 Ks, float scale, bool scaled,
                     bool soft = false, float factor = 1.0f) {
    f3 color{0, 0, 0};
    f3 Ld = ld3(s->light_pos + 4 * l) - P;
    float dist = length(Ld);
    Ld = normalize(Ld);
    f3 lc = ld3(s->light_col + 4 * l);
    float Ix = s->light_int[2 * l], Iy = s->light_int[2 * l + 1];
    float att = 1.0f / ((0.75f + dist * 0.0005f) + (dist * dist) * 0.00005f);
    float diffuse = clampf(fmax_(dot(Ld, N), 0.0f), 0.0f, 1.0f);
    if (diffuse > 0.0f) {
        f3 Hv = normalize(Ld - normalize(P - eye));
        float specular = clampf(pow_f(fmax_(dot(Hv, N), 0.0f), Ks.w), 0.0f, 1.0f);
        f3 cd, cs;
        if (scaled) {
            cd = ((((xyz(Kd) * lc) * diffuse) * Ix) * att) * scale;
            cs = ((((xyz(Ks) * lc) * specular) * Iy) * att) * scale;
        } else {
            cd = (((xyz(Kd) * lc) * diffuse) * Ix) * att;
            cs = (((xyz(Ks) * lc) * specular) * Iy) * att;
        }
        if (soft) {
            cd = cd * factor;
            cs = cs * factor;
        }
        color = color + cd;
        color = color + cs;
    }
    return color;
}

// RT:1616-1709 ColorPrimaryShadowRay
// Soft shadows: the 16 flags of slot i as a mask of occluded cells.
uint32_t soft_mask(const std::vector<uint64_t>& sflags, size_t i) {
    uint32_t m = 0;
    for (uint32_t c = 0; c < 16; c++)
        if (sflags[16 * i + c] == MISS32) m |= 1u << c;
    return m;
}

void shade_primary(const Scene& sc, const GBuf& g, const std::vector<uint64_t>& sflags, uint32_t Lmax,
                   uint32_t Lt, const ViewSet& vs, bool soft, uint32_t* pixels) {
    StageClock clk_(SC_SHADE);
    const uint32_t W = g.W, H = g.H;
#pragma omp parallel for schedule(static)
    for (int64_t pp = 0; pp < (int64_t)W * H; pp++) {
        uint32_t x = (uint32_t)(pp % W), y = (uint32_t)(pp / W);
        f3 color{0, 0, 0};
        f3 P = xyz(g.pos[pp]), N = xyz(g.nrm[pp]);
        if (length(P) != 0.0f) {
            const f3 eye = vs.eye[vs.view_of((uint32_t)pp)];
            for (uint32_t l = 0; l < Lt; l++) {
                const size_t slot = (size_t)(x + y * W * Lmax) + (size_t)W * l;
                if (soft) {
                    // RT:1644-1647 / 1680-1697: no light when every cell is occluded
                    const uint32_t m = soft_mask(sflags, slot);
                    if (m != 0xFFFFu)
                        color = color + light_term(sc.s, l, P, N, eye, g.dif[pp], g.spec[pp], 1.0f, false, true,
                                                   soft_factor(m));
                } else if (sflags[slot] != MISS32) {
                    color = color + light_term(sc.s, l, P, N, eye, g.dif[pp], g.spec[pp], 1.0f, false);
                }
            }
        }
        pixels[pp] = rgb_to_int(color.x * 255.0f, color.y * 255.0f, color.z * 255.0f);
    }
}

// RT:1776-1882 ColorReflectionRay, shading from the surface record; with
// bounce shadows a light whose shadow flag is set is skipped.
void shade_reflection(const Scene& sc, const oracle_config& cfg, const GBuf& g,
                      const std::vector<uint64_t>& rec, const std::vector<uint64_t>* sflags,
                      uint32_t Lmax, uint32_t Lt, const ViewSet& vs, uint32_t* pixels) {
    StageClock clk_(SC_SHADE);
    const uint32_t W = g.W, H = g.H;
#pragma omp parallel for schedule(static)
    for (int64_t pp = 0; pp < (int64_t)W * H; pp++) {
        uint32_t x = (uint32_t)(pp % W), y = (uint32_t)(pp / W);
        uint32_t slot = x + y * W * Lmax;
        if (is_miss(rec[slot], cfg.hit_mode)) continue;
        const f3 eye = vs.eye[vs.view_of((uint32_t)pp)];
        f3 P = xyz(g.pos[pp]), N = xyz(g.nrm[pp]);
        f4 Kd = g.dif[pp], Ks = g.spec[pp];
        f3 color{0, 0, 0};
        for (uint32_t l = 0; l < Lt; l++) {
            const size_t sl = (size_t)slot + (size_t)W * l;
            if (sflags && cfg.soft_shadows) {
                const uint32_t m = soft_mask(*sflags, sl);
                if (m != 0xFFFFu)
                    color = color + light_term(sc.s, l, P, N, eye, Kd, Ks, 0.75f, true, true, soft_factor(m));
                continue;
            }
            if (sflags && (*sflags)[sl] == MISS32) continue;
            color = color + light_term(sc.s, l, P, N, eye, Kd, Ks, 0.75f, true);
        }
        pixels[pp] += rgb_to_int(color.x * Ks.w, color.y * Ks.w, color.z * Ks.w);
    }
}

// ----------------------------------------------------- literal traversal ---
// RT:2537-2852 with the flag / inclusive-scan mechanics kept as written, over
// triangle batches of size Tb (MAIN:685-692).  Emits (leaf << 32 | tri).
void leaf_pairs_literal(const Scene& sc, const Hier& h, uint32_t branch, uint32_t algorithm,
                        uint32_t Tb, std::vector<uint64_t>& out) {
    const uint32_t T = sc.s->triangle_total, K = sc.s->object_total;
    const uint32_t top = h.depth - 1, ntop = h.count[top];
    for (uint32_t offset = 0; offset < T; offset += Tb) {
        const uint32_t tb = std::min(Tb, T - offset);
        const uint32_t lower = offset, upper = offset + tb;
        const uint64_t hitMax = (uint64_t)ntop * tb;
        std::vector<uint64_t> cand;     // (node, tri-in-batch)
        if (algorithm == 0) {
            std::vector<uint32_t> flags(hitMax, 0);
            for (uint32_t n = 0; n < ntop; n++) {
                uint32_t g = h.offset[top] + n;
                f4 sphere = h.nodes[2 * (size_t)g], cone = h.nodes[2 * (size_t)g + 1];
                for (uint32_t k = 0; k < K; k++) {
                    const float* p = sc.s->object_spheres + 6 * k;
                    float ry = p[4], rz = p[5];
                    bool res = sphere_node(sphere, cone, f4{p[0], p[1], p[2], p[3]}, cos_f(cone.w), tan_f(cone.w));
                    uint32_t* f = flags.data() + (size_t)n * tb;
                    if (res) {
                        if (ry >= lower && ry < upper && rz >= lower && rz < upper)
                            f[(uint32_t)ry - lower] += (uint32_t)rz - (uint32_t)ry + 1;
                        else if (ry < lower && rz >= lower && rz < upper)
                            f[0] += (uint32_t)rz - lower + 1;
                        else if (ry >= lower && ry < upper && rz >= upper)
                            f[(uint32_t)ry - lower] += tb;
                        else if (ry < lower && rz > upper)
                            f[0] += (uint32_t)rz - (uint32_t)ry + 1 - ((uint32_t)rz - upper + 1);
                    } else {
                        if (ry >= lower && ry < upper && rz >= lower && rz + 1 < upper)
                            f[(uint32_t)rz - lower + 1] += (uint32_t)rz - (uint32_t)ry + 1;
                        else if (ry < offset && rz >= offset && (rz + 1) < (offset + tb))
                            f[(uint32_t)rz - lower + 1] += (uint32_t)rz - lower + 1;
                    }
                }
            }
            std::vector<uint32_t> scan(hitMax);
            uint32_t acc = 0;
            for (uint64_t x = 0; x < hitMax; x++) { acc += flags[x]; scan[x] = acc; }
            for (uint64_t x = 0; x < hitMax; x++) {
                uint32_t n = (uint32_t)(x / tb), t = (uint32_t)(x % tb);
                int off = (int)scan[x];
                if (n > 0) off -= (int)scan[(size_t)n * tb - 1];
                if ((uint32_t)off > t) cand.push_back(((uint64_t)n << 32) | t);
            }
        } else {
            for (uint64_t x = 0; x < hitMax; x++) cand.push_back(((x / tb) << 32) | (x % tb));
        }
        // Top-level triangle-sphere test, then descend level by level.
        std::vector<uint64_t> cur;
        for (uint64_t c : cand) {
            uint32_t n = (uint32_t)(c >> 32), t = (uint32_t)c;
            uint32_t g = h.offset[top] + n;
            f4 cone = h.nodes[2 * (size_t)g + 1];
            if (sphere_node(h.nodes[2 * (size_t)g], cone, sc.tri_sphere[offset + t], cos_f(cone.w), tan_f(cone.w)))
                cur.push_back(c);
        }
        for (int level = (int)h.depth - 2; level >= 0; level--) {
            std::vector<uint64_t> next;
            for (uint64_t c : cur) {
                uint32_t n = (uint32_t)(c >> 32), t = (uint32_t)c;
                for (uint32_t i = 0; i < branch; i++) {
                    uint32_t ch = n * branch + i;
                    if (ch >= h.count[level]) continue;
                    uint32_t g = h.offset[level] + ch;
                    f4 cone = h.nodes[2 * (size_t)g + 1];
                    if (sphere_node(h.nodes[2 * (size_t)g], cone, sc.tri_sphere[offset + t], cos_f(cone.w), tan_f(cone.w)))
                        next.push_back(((uint64_t)ch << 32) | t);
                }
            }
            cur.swap(next);
        }
        for (uint64_t c : cur) out.push_back((c & 0xFFFFFFFF00000000ull) | ((c & 0xFFFFFFFFull) + offset));
    }
}

void hier_from_arrays(const float* hierarchy, const uint32_t* counts, uint32_t depth, Hier& h) {
    h.depth = depth;
    h.count.assign(counts, counts + depth);
    h.offset.assign(depth, 0);
    uint32_t total = 0;
    for (uint32_t k = 0; k < depth; k++) { h.offset[k] = total; total += counts[k]; }
    h.nodes.resize(2 * (size_t)total);
    for (size_t i = 0; i < 2 * (size_t)total; i++) h.nodes[i] = ld4(hierarchy + 4 * i);
    h.cosv.resize(total); h.tanv.resize(total);
    for (uint32_t g = 0; g < total; g++) {
        h.cosv[g] = cos_f(h.nodes[2 * (size_t)g + 1].w);
        h.tanv[g] = tan_f(h.nodes[2 * (size_t)g + 1].w);
    }
}

// The dump targets of the pass record_pass will record next.
PassDump dump_for(const oracle_frame_out* out) {
    PassDump d;
    if (!out || out->pass_count >= 16) return d;
    const uint32_t i = out->pass_count;
    d.keys = out->pass_keys[i]; d.flags = out->pass_flags[i]; d.order = out->pass_order[i];
    d.bases = out->pass_bases[i]; d.ckeys = out->pass_ckeys[i]; d.nodes = out->pass_nodes[i];
    return d;
}

void record_pass(oracle_frame_out* out, uint32_t kind, uint32_t S, const PassStats& st,
                 const std::vector<uint64_t>& res) {
    if (!out || out->pass_count >= 16) return;
    uint32_t i = out->pass_count++;
    out->pass_kind[i] = kind; out->pass_slots[i] = S; out->pass_rays[i] = st.N;
    out->pass_chunks[i] = st.C; out->pass_pairs_top[i] = st.top; out->pass_pairs_leaf[i] = st.leaf;
    out->pass_ray_tests[i] = st.tests;
    if (out->pass_hits[i]) std::copy(res.begin(), res.end(), out->pass_hits[i]);
}

// --------------------------------------------------- wrapper mirrors state --
struct Bound {
    uint32_t light_maximum = 1, depth = 2, subdivision = 8, algorithm = 0;
    const float *g_dif = nullptr, *g_spec = nullptr, *g_pos = nullptr, *g_nrm = nullptr;
    oracle_scene scene{};
    const float *orig_pos = nullptr, *orig_nrm = nullptr, *orig_spheres = nullptr;
    const int32_t* orig_obj = nullptr;
} B;

}  // namespace

// =========================================================== C interface ===
extern "C" {

float oracle_acos(float x) { return acos_f(x); }
float oracle_cos(float x) { return cos_f(x); }
float oracle_sin(float x) { return sin_f(x); }
float oracle_tan(float x) { return tan_f(x); }
float oracle_pow(float x, float y) { return pow_f(x, y); }
uint32_t oracle_f2u_sat(float x) { return f2u_sat(x); }
uint32_t oracle_shadow_key_ref(const float* o, const float* d, uint32_t l) { return shadow_key_ref(ld3(o), ld3(d), l); }
uint32_t oracle_reflection_key_ref(const float* o, const float* d) { return reflection_key_ref(ld3(o), ld3(d)); }
uint32_t oracle_refraction_key_ref(const float* o, const float* d) { return refraction_key_ref(ld3(o), ld3(d)); }
void oracle_refract(const float* i, const float* n, float ior, float* out3) {
    const f3 r = refract(ld3(i), ld3(n), ior);
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
uint32_t oracle_key_oct(const oracle_config* c, const float* o, const float* d, uint32_t type) {
    return key_oct(*c, ld3(o), ld3(d), type);
}
float oracle_ray_triangle(const float* o, const float* d, const float* v0, const float* e1, const float* e2) {
    return ray_triangle(ld3(o), ld3(d), ld3(v0), ld3(e1), ld3(e2));
}
void oracle_soft_grid(const float* light, const float* d, float* out48) {
    f3 a0, a1;
    soft_axes(ld3(d), a0, a1);
    for (uint32_t c = 0; c < 16; c++) {
        const f3 p = soft_point(ld3(light), a0, a1, c);
        out48[3 * c] = p.x; out48[3 * c + 1] = p.y; out48[3 * c + 2] = p.z;
    }
}
float oracle_soft_factor(uint32_t occluded) { return soft_factor(occluded); }
int oracle_sphere_node(const float* sphere, const float* cone, const float* tri) {
    f4 c = ld4(cone);
    return sphere_node(ld4(sphere), c, ld4(tri), cos_f(c.w), tan_f(c.w)) ? 1 : 0;
}
void oracle_triangle_sphere(const float* v0, const float* v1, const float* v2, float* out4) {
    st4(out4, triangle_sphere(ld3(v0), ld3(v1), ld3(v2)));
}
void oracle_cone0(const float* cone, const float* ray, uint32_t mode, float* out4) {
    st4(out4, cone0(ld4(cone), ld4(ray), mode));
}
void oracle_coneN(const float* c1, const float* c2, float* out4) { st4(out4, coneN(ld4(c1), ld4(c2), 0)); }
void oracle_sphere_union(const float* s1, const float* s2, float* out4) { st4(out4, sphere_union(ld4(s1), ld4(s2))); }
uint32_t oracle_rgb_to_int(float r, float g, float b) { return rgb_to_int(r, g, b); }
uint32_t oracle_primary_key(uint32_t x, uint32_t y, uint32_t w, uint32_t view, uint32_t segment_shift) {
    return primary_key(x, y, w, view, segment_shift);
}
void oracle_light_term(const float* light_pos4, const float* light_col4, const float* light_int2, const float* P,
                       const float* N, const float* eye, const float* Kd4, const float* Ks4, float scale, int scaled,
                       float* out3) {
    oracle_scene one{};
    one.light_total = 1;
    one.light_pos = light_pos4;
    one.light_col = light_col4;
    one.light_int = light_int2;
    st3(out3, light_term(&one, 0, ld3(P), ld3(N), ld3(eye), ld4(Kd4), ld4(Ks4), scale, scaled != 0));
}

}  // extern "C"

namespace {

// The frame driver (MAIN:630-1054, 1234-1303) over a view set; cfg.width /
// height give the image (a batch: P x 1).
int render_views(const oracle_scene* scene, const ViewSet& vs, const oracle_config& cfg, const float* gbuffer_in,
                 oracle_frame_out* out, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    if (cfg.leaf == 0 || cfg.branch == 0 || cfg.depth == 0 || cfg.width == 0 || cfg.height == 0) return -1;
    Scene sc(scene);
    const uint32_t W = cfg.width, H = cfg.height, Lt = scene->light_total, Lmax = Lt;
    if (out) out->pass_count = 0;
    std::fill(g_stage_ms, g_stage_ms + SC_STAGES, 0.0);
    std::vector<uint32_t> pixels((size_t)W * H, 0);
    GBuf g0;
    if (cfg.primary_rays) {
        Batch b;
        create_primary(cfg, vs, b);
        std::vector<uint64_t> rec;
        PassStats st;
        const PassDump dp = dump_for(out);
        trace(sc, cfg, b, 0, nullptr, 1, rec, st, threads, nullptr, &dp);
        record_pass(out, 0, b.S, st, rec);
        surface_from_hits(sc, cfg, b, rec, 1, true, g0);
    } else {
        g0.init(W, H);
        const size_t n = (size_t)W * H;
        for (size_t i = 0; i < n; i++) {
            g0.dif[i] = ld4(gbuffer_in + 4 * i);
            g0.spec[i] = ld4(gbuffer_in + 4 * (n + i));
            g0.pos[i] = ld4(gbuffer_in + 4 * (2 * n + i));
            g0.nrm[i] = ld4(gbuffer_in + 4 * (3 * n + i));
        }
    }
    if (out && out->gbuffer_pos) {
        for (size_t i = 0; i < (size_t)W * H; i++) {
            st4(out->gbuffer_pos + 4 * i, g0.pos[i]); st4(out->gbuffer_nrm + 4 * i, g0.nrm[i]);
            st4(out->gbuffer_dif + 4 * i, g0.dif[i]); st4(out->gbuffer_spec + 4 * i, g0.spec[i]);
        }
    }
    {
        Batch b;
        create_shadow(sc, cfg, g0, Lmax, vs, b);
        std::vector<uint64_t> flags;
        PassStats st;
        const PassDump dp = dump_for(out);
        trace(sc, cfg, b, 1, &g0, Lmax, flags, st, threads, nullptr, &dp);
        record_pass(out, 1, b.S, st, flags);
        if (st.N > 0) shade_primary(sc, g0, flags, Lmax, Lt, vs, cfg.soft_shadows != 0, pixels.data());
    }
    Batch prev;
    GBuf gprev;
    std::vector<uint64_t> prev_rec;
    for (uint32_t bounce = 1; bounce <= cfg.bounces; bounce++) {
        Batch b;
        if (bounce == 1) create_reflection(cfg, g0, vs, Lmax, b);
        else create_bounce(cfg, gprev, prev, prev_rec, Lmax, vs, b);
        std::vector<uint64_t> rec;
        PassStats st;
        const PassDump dp = dump_for(out);
        trace(sc, cfg, b, 0, nullptr, Lmax, rec, st, threads, nullptr, &dp);
        record_pass(out, 2, b.S, st, rec);
        if (st.N == 0) break;
        GBuf gb;
        surface_from_hits(sc, cfg, b, rec, Lmax, false, gb);
        std::vector<uint64_t> sflags;
        bool have_flags = false;
        if (cfg.bounce_shadows) {
            Batch sbt;
            create_shadow(sc, cfg, gb, Lmax, vs, sbt);
            PassStats sst;
            const PassDump dps = dump_for(out);
            trace(sc, cfg, sbt, 1, &gb, Lmax, sflags, sst, threads, nullptr, &dps);
            record_pass(out, 1, sbt.S, sst, sflags);
            have_flags = true;
        }
        shade_reflection(sc, cfg, gb, rec, have_flags ? &sflags : nullptr, Lmax, Lt, vs, pixels.data());
        prev = std::move(b);
        gprev = std::move(gb);
        prev_rec = std::move(rec);
    }
    if (out && out->pixels) std::copy(pixels.begin(), pixels.end(), out->pixels);
    if (out) std::copy(g_stage_ms, g_stage_ms + SC_STAGES, out->stage_ms);
    return 0;
}

void add_view(ViewSet& vs, const oracle_camera& c, uint32_t off, uint32_t w, uint32_t h) {
    vs.off.push_back(off); vs.w.push_back(w); vs.h.push_back(h);
    vs.eye.push_back(ld3(c.eye)); vs.D.push_back(ld3(c.direction));
    vs.R.push_back(ld3(c.right)); vs.U.push_back(ld3(c.up));
}

}  // namespace

extern "C" {

int oracle_render(const oracle_scene* scene, const oracle_camera* camera, const oracle_config* config,
                  const float* gbuffer_in, oracle_frame_out* out, int threads) {
    ViewSet vs;
    add_view(vs, *camera, 0, config->width, config->height);
    return render_views(scene, vs, *config, gbuffer_in, out, threads);
}

int oracle_render_views(const oracle_scene* scene, const oracle_view* views, uint32_t n,
                        const oracle_config* config, oracle_frame_out* out, int threads) {
    if (n == 0) return -1;
    ViewSet vs;
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; i++) {
        add_view(vs, views[i].camera, (uint32_t)off, views[i].width, views[i].height);
        off += (uint64_t)views[i].width * views[i].height;
    }
    if (off == 0 || off > 0xFFFFFFFFull) return -1;
    oracle_config cfg = *config;
    cfg.width = n == 1 ? views[0].width : (uint32_t)off;
    cfg.height = n == 1 ? views[0].height : 1u;
    cfg.row_begin = cfg.row_end = 0;
    return render_views(scene, vs, cfg, nullptr, out, threads);
}

// Ray query (rsh_engine_trace): n rays float[6n], keyed as frame rays of type
// kind (0 closest -> type 0, 1 any -> type 1) with segs at segment_shift.
int oracle_trace_rays(const oracle_scene* scene, const float* rays, const float* tmax, const uint32_t* segs,
                      uint32_t n, uint32_t kind, const oracle_config* config, uint64_t* out, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    const oracle_config& cfg = *config;
    if (n == 0 || kind > 1 || (kind == 1 && !tmax) || cfg.leaf == 0 || cfg.branch == 0 || cfg.depth == 0) return -1;
    Scene sc(scene);
    Batch b;
    b.init(n);
    for (uint32_t i = 0; i < n; i++) {
        const f3 o = ld3(rays + 6 * (size_t)i), d = ld3(rays + 6 * (size_t)i + 3);
        b.rays[2 * (size_t)i] = o;
        b.rays[2 * (size_t)i + 1] = d;
        if (cfg.key_mode == 0) b.keys[i] = kind == 1 ? shadow_key_ref(o, d, 0) : reflection_key_ref(o, d);
        else b.keys[i] = view_key(cfg, o, d, kind == 1 ? 1u : 0u, segs ? segs[i] : 0u);
        b.values[i] = i;
        b.flags[i] = 0;
    }
    std::vector<uint64_t> res;
    PassStats st;
    oracle_config c1 = cfg;
    c1.width = n; c1.height = 1;
    trace(sc, c1, b, kind, nullptr, 1, res, st, threads, tmax);
    std::copy(res.begin(), res.end(), out);
    return 0;
}

uint64_t oracle_leaf_pairs_literal(const float* hierarchy, const uint32_t* level_counts, uint32_t depth,
                                   uint32_t branch, const oracle_scene* s, uint32_t algorithm,
                                   uint32_t triangle_batch, uint64_t* out, uint64_t cap) {
    Scene sc(s);
    Hier h;
    hier_from_arrays(hierarchy, level_counts, depth, h);
    std::vector<uint64_t> pairs;
    leaf_pairs_literal(sc, h, branch, algorithm, triangle_batch, pairs);
    std::sort(pairs.begin(), pairs.end());
    uint64_t n = std::min<uint64_t>(cap, pairs.size());
    std::copy(pairs.begin(), pairs.begin() + n, out);
    return pairs.size();
}

uint64_t oracle_leaf_pairs_closed(const float* hierarchy, const uint32_t* level_counts, uint32_t depth,
                                  uint32_t branch, const oracle_scene* s, uint32_t algorithm,
                                  uint32_t triangle_batch, uint64_t* out, uint64_t cap) {
    Scene sc(s);
    Hier h;
    hier_from_arrays(hierarchy, level_counts, depth, h);
    Traverser tr{sc, h, algorithm, branch, triangle_batch};
    std::vector<uint64_t> pairs;
    PassStats st;
    auto f = [&](uint32_t c, uint32_t t) { pairs.push_back(((uint64_t)c << 32) | t); };
    for (uint32_t j = 0; j < h.count[depth - 1]; j++) tr.top_node(j, f, st);
    std::sort(pairs.begin(), pairs.end());
    uint64_t n = std::min<uint64_t>(cap, pairs.size());
    std::copy(pairs.begin(), pairs.begin() + n, out);
    return pairs.size();
}

// ----------------------------------------------------- wrapper mirrors -----
void oracle_set_constants(uint32_t light_maximum, uint32_t depth, uint32_t subdivision, uint32_t algorithm) {
    B.light_maximum = light_maximum; B.depth = depth; B.subdivision = subdivision; B.algorithm = algorithm;
}
void oracle_bind_gbuffer(const float* dif, const float* spec, const float* pos, const float* nrm) {
    B.g_dif = dif; B.g_spec = spec; B.g_pos = pos; B.g_nrm = nrm;
}
void oracle_bind_scene(const oracle_scene* s) { B.scene = *s; }
void oracle_bind_original(const float* pos, const float* nrm, const int32_t* obj, const float* spheres4) {
    B.orig_pos = pos; B.orig_nrm = nrm; B.orig_obj = obj; B.orig_spheres = spheres4;
}

// RT:448-511 UpdateVertex via RT:2046-2069
void oracle_TriangleUpdateWrapper(const float* model, const float* normal, uint32_t T, float* pos_out, float* nrm_out) {
    for (uint32_t x = 0; x < 3 * T; x++) {
        uint32_t m = (uint32_t)B.orig_obj[x];
        const float* M = model + 16 * (size_t)m;
        f4 v = ld4(B.orig_pos + 4 * (size_t)x);
        float u[4];
        for (int i = 0; i < 4; i++) {
            u[i] = 0.0f;
            u[i] += M[i * 4 + 0] * v.x; u[i] += M[i * 4 + 1] * v.y;
            u[i] += M[i * 4 + 2] * v.z; u[i] += M[i * 4 + 3] * v.w;
        }
        st4(pos_out + 4 * (size_t)x, f4{u[0], u[1], u[2], (float)m});
        const float* Nm = normal + 16 * (size_t)m;
        f4 n = ld4(B.orig_nrm + 4 * (size_t)x);
        float w[4];
        for (int i = 0; i < 4; i++) {
            w[i] = 0.0f;
            w[i] += Nm[i * 4 + 0] * n.x; w[i] += Nm[i * 4 + 1] * n.y;
            w[i] += Nm[i * 4 + 2] * n.z; w[i] += Nm[i * 4 + 3] * n.w;
        }
        st4(nrm_out + 4 * (size_t)x, mk4(normalize(f3{w[0], w[1], w[2]}), 0.0f));
    }
}

// RT:513-570 UpdateBoundingSphere via RT:2071-2091
void oracle_BoundingSphereUpdateWrapper(const float* translation, const float* scale, uint32_t K, float* out3) {
    for (uint32_t x = 0; x < K; x++) {
        const float* Tm = translation + 16 * (size_t)x;
        const float* Sm = scale + 16 * (size_t)x;
        f4 c = ld4(B.orig_spheres + 8 * (size_t)x);
        f4 rb = ld4(B.orig_spheres + 8 * (size_t)x + 4);
        float uc[4];
        for (int i = 0; i < 4; i++) {
            uc[i] = 0.0f;
            uc[i] += Tm[i * 4 + 0] * c.x; uc[i] += Tm[i * 4 + 1] * c.y;
            uc[i] += Tm[i * 4 + 2] * c.z; uc[i] += Tm[i * 4 + 3] * 1.0f;
        }
        float ms = 1.17549435e-38f;   // FLT_MIN
        for (int i = 0; i < 3; i++)
            if (std::fabs(Sm[i * 4 + i]) > ms) ms = std::fabs(Sm[i * 4 + i]);
        st3(out3 + 6 * (size_t)x, f3{uc[0], uc[1], uc[2]});
        st3(out3 + 6 * (size_t)x + 3, f3{rb.x * ms, rb.y, rb.z});
    }
}

// RT:682-740 via RT:2210-2241 (reference keys).
void oracle_ShadowRayCreationWrapper(uint32_t W, uint32_t H, uint32_t L, float* rays, uint32_t* flags,
                                     uint32_t* keys, uint32_t* values) {
    for (uint32_t i = 0; i < W * H * L; i++) {
        uint32_t x = i % W, y = i / (W * L), l = (i / W) % L;
        f3 P = ld3(B.g_pos + 4 * (size_t)(x + y * W));
        f3 N = normalize(ld3(B.g_nrm + 4 * (size_t)(x + y * W)));
        if (length(P) != 0.0f) {
            f3 o = ld3(B.scene.light_pos + 4 * l);
            f3 d = normalize(P - o);
            float diffuse = fmax_(dot(d, -N), 0.0f);
            if (diffuse > 0.0f) {
                st3(rays + 6 * (size_t)i, o); st3(rays + 6 * (size_t)i + 3, d);
                keys[i] = shadow_key_ref(o, d, l);
                values[i] = i;
                flags[i] = 0;
                continue;
            }
        }
        flags[i] = 1;
    }
}

// RT:742-791 via RT:2243-2274 (only covered slots are written, as in the reference).
void oracle_ReflectionRayCreationWrapper(uint32_t W, uint32_t H, const float* eyep, float* rays, uint32_t* flags,
                                         uint32_t* keys, uint32_t* values) {
    f3 eye = ld3(eyep);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            uint32_t r = x + y * W * B.light_maximum;
            f3 P = ld3(B.g_pos + 4 * (size_t)(x + y * W));
            f3 N = normalize(ld3(B.g_nrm + 4 * (size_t)(x + y * W)));
            if (length(P) != 0.0f) {
                f3 d = reflect(normalize(P - eye), normalize(N));
                st3(rays + 6 * (size_t)r, P); st3(rays + 6 * (size_t)r + 3, d);
                keys[r] = reflection_key_ref(P, d);
                values[r] = r;
                flags[r] = 0;
            } else {
                flags[r] = 1;
            }
        }
}

// RT:793-842 CreateRefractionRays (never launched by the reference), with the
// arguments of the reflection wrapper.
void oracle_RefractionRayCreationWrapper(uint32_t W, uint32_t H, const float* eyep, float* rays, uint32_t* flags,
                                         uint32_t* keys, uint32_t* values) {
    f3 eye = ld3(eyep);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            uint32_t r = x + y * W * B.light_maximum;
            f3 P = ld3(B.g_pos + 4 * (size_t)(x + y * W));
            f3 N = normalize(ld3(B.g_nrm + 4 * (size_t)(x + y * W)));
            if (length(P) != 0.0f) {
                f3 d = refract(normalize(P - eye), normalize(N), 1.0f / 1.52f);
                st3(rays + 6 * (size_t)r, P); st3(rays + 6 * (size_t)r + 3, d);
                keys[r] = refraction_key_ref(P, d);
                values[r] = r;
                flags[r] = 0;
            } else {
                flags[r] = 1;
            }
        }
}

// RT:2276-2324 (inclusive scan of the flags, CreateTrimmedRays RT:844-884).
void oracle_RayTrimmingWrapper(const uint32_t* keys, const uint32_t* values, uint32_t W, uint32_t H, uint32_t L,
                               const uint32_t* flags, uint32_t* scan, uint32_t* tk, uint32_t* tv, uint32_t* total) {
    uint32_t S = W * H * L, acc = 0;
    for (uint32_t i = 0; i < S; i++) { acc += flags[i]; scan[i] = acc; }
    for (uint32_t x = 0; x < S; x++) {
        if (x == 0) {
            if (scan[0] == 0) { tk[0] = keys[0]; tv[0] = values[0]; }
        } else if (scan[x] == scan[x - 1]) {
            tk[x - scan[x]] = keys[x]; tv[x - scan[x]] = values[x];
        }
    }
    *total = S - scan[S - 1];
}

// RT:2326-2383
void oracle_RayCompressionWrapper(const uint32_t* tk, const uint32_t* tv, uint32_t N, uint32_t* flags,
                                  uint32_t* scan, uint32_t* bases, uint32_t* sizes, uint32_t* ck,
                                  uint32_t* cv, uint32_t* total) {
    (void)tv;
    for (uint32_t x = 0; x < N; x++) flags[x] = (x == 0 || tk[x] != tk[x - 1]) ? 1u : 0u;
    uint32_t acc = 0;
    for (uint32_t x = 0; x < N; x++) { acc += flags[x]; scan[x] = acc; }
    uint32_t C = N ? scan[N - 1] : 0;
    for (uint32_t x = 0; x < N; x++) {
        if (flags[x] == 0) continue;
        uint32_t p = scan[x] - 1;
        bases[p] = x; ck[p] = tk[x]; cv[p] = p;
    }
    for (uint32_t x = 0; x < C; x++) sizes[x] = (x == C - 1 ? N : bases[x + 1]) - bases[x];
    *total = C;
}

// RT:2385-2409 (CUB stable LSD radix sort over 32 bits)
void oracle_RaySortingWrapper(const uint32_t* ck, const uint32_t* cv, uint32_t C, uint32_t* sk, uint32_t* sv) {
    std::vector<uint32_t> perm(C);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return ck[a] < ck[b]; });
    for (uint32_t i = 0; i < C; i++) { sk[i] = ck[perm[i]]; sv[i] = cv[perm[i]]; }
}

// RT:2411-2467
void oracle_RayDecompressionWrapper(const uint32_t* bases, const uint32_t* sizes, const uint32_t* skeys,
                                    const uint32_t* svalues, const uint32_t* tk, const uint32_t* tv,
                                    uint32_t* skeleton, uint32_t* scan, uint32_t C, uint32_t* rk, uint32_t* rv) {
    (void)skeys;
    for (uint32_t x = 0; x < C; x++) skeleton[x] = sizes[svalues[x]];
    uint32_t acc = 0;
    for (uint32_t x = 0; x < C; x++) { scan[x] = acc; acc += skeleton[x]; }
    for (uint32_t x = 0; x < C; x++) {
        uint32_t c = svalues[x], base = bases[c], size = sizes[c], start = scan[x];
        for (uint32_t j = 0; j < size; j++) { rk[start + j] = tk[base + j]; rv[start + j] = tv[base + j]; }
    }
}

// RT:2469-2535 (reference cone union, fixed-size leaves of HIERARCHY_SUBDIVISION)
void oracle_HierarchyCreationWrapper(const float* rays, const uint32_t* skeys, const uint32_t* svalues, uint32_t N,
                                     float r0, float s0, float* hierarchy) {
    (void)skeys;
    uint32_t maxslot = 0;
    for (uint32_t i = 0; i < N; i++) maxslot = std::max(maxslot, svalues[i]);
    Batch b;
    b.S = maxslot + 1;
    b.rays.resize(2 * (size_t)b.S);
    for (size_t i = 0; i < 2 * (size_t)b.S; i++) b.rays[i] = ld3(rays + 3 * i);
    std::vector<uint32_t> sv(svalues, svalues + N);
    Hier h;
    build_hierarchy(b, sv, N, B.subdivision, B.subdivision, B.depth, r0, s0, 0, h);
    for (size_t i = 0; i < h.nodes.size(); i++) st4(hierarchy + 4 * i, h.nodes[i]);
}

static void mirror_levels(uint32_t N, std::vector<uint32_t>& count, std::vector<uint32_t>& offset) {
    count.assign(B.depth, 0); offset.assign(B.depth, 0);
    uint32_t n = N / B.subdivision + (N % B.subdivision != 0 ? 1 : 0), total = 0;
    for (uint32_t k = 0; k < B.depth; k++) {
        count[k] = n; offset[k] = total; total += n;
        n = n / B.subdivision + (n % B.subdivision != 0 ? 1 : 0);
    }
}

static inline uint32_t create_hit(uint32_t node, uint32_t tri) { return ((node << 14) & 0xFFFFC000u) + (tri & 0x3FFFu); }
static inline uint32_t hit_node(uint32_t hit) { return (hit & 0xFFFFC000u) >> 14; }
static inline uint32_t hit_triangle(uint32_t hit) { return hit & 0x3FFFu; }

static inline bool node_vs_triangle(const float* hierarchy, uint32_t g, const float* tri_pos, uint32_t t) {
    f4 s = ld4(hierarchy + 8 * (size_t)g), c = ld4(hierarchy + 8 * (size_t)g + 4);
    f4 ts = triangle_sphere(ld3(tri_pos + 12 * (size_t)t), ld3(tri_pos + 12 * (size_t)t + 4), ld3(tri_pos + 12 * (size_t)t + 8));
    return sphere_node(s, c, ts, cos_f(c.w), tan_f(c.w));
}

// RT:2537-2739
void oracle_HierarchyTraversalWarmUpWrapper(const float* hierarchy, const float* spheres3, const float* tri_pos,
                                            uint32_t K, uint32_t N, uint32_t Tb, uint32_t offset, uint32_t* flags,
                                            uint32_t* scan, uint32_t* hits, uint32_t* thits, uint32_t* hit_total) {
    std::vector<uint32_t> count, off;
    mirror_levels(N, count, off);
    const uint32_t top = B.depth - 1, ntop = count[top], nodeOffset = off[top];
    const uint32_t hitMax = ntop * Tb;
    uint32_t hitTotal;
    if (B.algorithm == 0) {
        std::fill(flags, flags + hitMax, 0u);
        const uint32_t lower = offset, upper = offset + Tb;
        for (uint32_t n = 0; n < ntop; n++) {
            f4 sphere = ld4(hierarchy + 8 * (size_t)(nodeOffset + n)), cone = ld4(hierarchy + 8 * (size_t)(nodeOffset + n) + 4);
            for (uint32_t k = 0; k < K; k++) {
                const float* p = spheres3 + 6 * k;
                float ry = p[4], rz = p[5];
                bool res = sphere_node(sphere, cone, f4{p[0], p[1], p[2], p[3]}, cos_f(cone.w), tan_f(cone.w));
                uint32_t* f = flags + (size_t)n * Tb;
                if (res) {
                    if (ry >= lower && ry < upper && rz >= lower && rz < upper) f[(uint32_t)ry - lower] += (uint32_t)rz - (uint32_t)ry + 1;
                    else if (ry < lower && rz >= lower && rz < upper) f[0] += (uint32_t)rz - lower + 1;
                    else if (ry >= lower && ry < upper && rz >= upper) f[(uint32_t)ry - lower] += Tb;
                    else if (ry < lower && rz > upper) f[0] += (uint32_t)rz - (uint32_t)ry + 1 - ((uint32_t)rz - upper + 1);
                } else {
                    if (ry >= lower && ry < upper && rz >= lower && rz + 1 < upper) f[(uint32_t)rz - lower + 1] += (uint32_t)rz - (uint32_t)ry + 1;
                    else if (ry < offset && rz >= offset && (rz + 1) < (offset + Tb)) f[(uint32_t)rz - lower + 1] += (uint32_t)rz - lower + 1;
                }
            }
        }
        uint32_t acc = 0;
        for (uint32_t x = 0; x < hitMax; x++) { acc += flags[x]; scan[x] = acc; }
        for (uint32_t x = 0; x < hitMax; x++) {
            uint32_t n = x / Tb, t = x % Tb;
            int o = (int)scan[x];
            if (n > 0) o -= (int)scan[n * Tb - 1];
            hits[x] = ((uint32_t)o > t) ? create_hit(n, t) : 0;
            flags[x] = ((uint32_t)o > t) ? 0 : 1;
        }
        acc = 0;
        for (uint32_t x = 0; x < hitMax; x++) { acc += flags[x]; scan[x] = acc; }
        for (uint32_t x = 0; x < hitMax; x++) {
            if (x == 0) { if (scan[0] == 0) thits[0] = hits[0]; }
            else if (scan[x] == scan[x - 1]) thits[x - scan[x]] = hits[x];
        }
        hitTotal = hitMax - scan[hitMax - 1];
    } else {
        hitTotal = hitMax;
    }
    *hit_total = hitTotal;
    if (hitTotal == 0) return;
    for (uint32_t x = 0; x < hitTotal; x++) {
        uint32_t hit = B.algorithm == 0 ? thits[x] : create_hit(x / Tb, x % Tb);
        uint32_t n = hit_node(hit), t = hit_triangle(hit);
        flags[x] = node_vs_triangle(hierarchy, nodeOffset + n, tri_pos, offset + t) ? 0 : 1;
        hits[x] = hit;
    }
    uint32_t acc = 0;
    for (uint32_t x = 0; x < hitTotal; x++) { acc += flags[x]; scan[x] = acc; }
    for (uint32_t x = 0; x < hitTotal; x++) {
        if (x == 0) { if (scan[0] == 0) thits[0] = hits[0]; }
        else if (scan[x] == scan[x - 1]) thits[x - scan[x]] = hits[x];
    }
    *hit_total = hitTotal - scan[hitTotal - 1];
}

// RT:2741-2852
void oracle_HierarchyTraversalWrapper(const float* hierarchy, const float* tri_pos, uint32_t N, uint32_t Tb,
                                      uint32_t offset, uint32_t* flags, uint32_t* scan, uint32_t* hits,
                                      uint32_t* thits, uint32_t* hit_total) {
    (void)Tb;
    std::vector<uint32_t> count, off;
    mirror_levels(N, count, off);
    const uint32_t Bs = B.subdivision;
    for (int level = (int)B.depth - 2; level >= 0; level--) {
        if (*hit_total == 0) return;
        uint32_t hitMax = *hit_total * Bs, hitTotal = *hit_total;
        for (uint32_t x = 0; x < hitTotal; x++) {
            uint32_t hit = thits[x], n = hit_node(hit), t = hit_triangle(hit);
            for (uint32_t i = 0; i < Bs; i++) {
                uint32_t c = n * Bs + i;
                if (c < count[level]) {
                    flags[x * Bs + i] = node_vs_triangle(hierarchy, off[level] + c, tri_pos, offset + t) ? 0 : 1;
                } else {
                    flags[x * Bs + i] = 1;
                }
                hits[x * Bs + i] = create_hit(c, t);
            }
        }
        uint32_t acc = 0;
        for (uint32_t x = 0; x < hitMax; x++) { acc += flags[x]; scan[x] = acc; }
        for (uint32_t x = 0; x < hitMax; x++) {
            if (x == 0) { if (scan[0] == 0) thits[0] = hits[0]; }
            else if (scan[x] == scan[x - 1]) thits[x - scan[x]] = hits[x];
        }
        *hit_total = hitMax - scan[hitMax - 1];
    }
}

void oracle_ShadowRayPreparationWrapper(uint32_t W, uint32_t H, uint32_t L, uint32_t* sflags) {
    std::fill(sflags, sflags + (size_t)W * H * L, 0u);
}

// RT:1498-1614 via RT:2874-2925 (hard shadows)
void oracle_ShadowRayIntersectionWrapper(const float* rays, const uint32_t* skeys, const uint32_t* svalues,
                                         const uint32_t* hits, const float* tri_pos, uint32_t hit_total, uint32_t N,
                                         uint32_t offset, uint32_t W, uint32_t H, uint32_t* sflags) {
    (void)skeys; (void)H;
    for (uint32_t x = 0; x < hit_total; x++) {
        uint32_t n = hit_node(hits[x]), t = hit_triangle(hits[x]);
        f3 v0 = ld3(tri_pos + 12 * (size_t)(offset + t));
        f3 e1 = ld3(tri_pos + 12 * (size_t)(offset + t) + 4) - v0;
        f3 e2 = ld3(tri_pos + 12 * (size_t)(offset + t) + 8) - v0;
        for (uint32_t i = 0; i < B.subdivision; i++) {
            if (n * B.subdivision + i >= N) break;
            uint32_t r = svalues[n * B.subdivision + i];
            f3 o = ld3(rays + 6 * (size_t)r), d = ld3(rays + 6 * (size_t)r + 3);
            f3 P = ld3(B.g_pos + 4 * (size_t)((r % W) + (r / (W * B.light_maximum)) * W));
            float th = ray_triangle(o + d * EPS, d, v0, e1, e2);
            if (th > EPS && th < length(o - P) - EPS * 2.0f) sflags[r] = MISS32;
        }
    }
}

// RT:1616-1709 via RT:2927-2957
void oracle_ShadowRayColoringWrapper(uint32_t W, uint32_t H, uint32_t L, const float* eyep, const uint32_t* sflags,
                                     uint32_t* pixels) {
    f3 eye = ld3(eyep);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            size_t pp = x + (size_t)y * W;
            f3 color{0, 0, 0};
            f3 P = ld3(B.g_pos + 4 * pp), N = ld3(B.g_nrm + 4 * pp);
            if (length(P) != 0.0f) {
                f4 Kd = ld4(B.g_dif + 4 * pp), Ks = ld4(B.g_spec + 4 * pp);
                for (uint32_t l = 0; l < L; l++) {
                    if (sflags[(x + y * W * B.light_maximum) + W * l] != MISS32)
                        color = color + light_term(&B.scene, l, P, N, eye, Kd, Ks, 1.0f, false);
                }
            }
            pixels[pp] = rgb_to_int(color.x * 255.0f, color.y * 255.0f, color.z * 255.0f);
        }
}

void oracle_ReflectionRayPreparationWrapper(uint32_t W, uint32_t H, uint32_t* times) {
    std::fill(times, times + (size_t)W * H * B.light_maximum, MISS32);
}

// RT:1711-1774 via RT:2977-3031 (reference 32-bit packing)
void oracle_ReflectionRayIntersectionWrapper(const float* rays, const uint32_t* skeys, const uint32_t* svalues,
                                             const uint32_t* hits, const float* tri_pos, uint32_t hit_total,
                                             uint32_t N, uint32_t offset, uint32_t W, uint32_t H, uint32_t* times) {
    (void)skeys; (void)W; (void)H;
    for (uint32_t x = 0; x < hit_total; x++) {
        uint32_t n = hit_node(hits[x]), t = hit_triangle(hits[x]);
        f3 v0 = ld3(tri_pos + 12 * (size_t)(offset + t));
        f3 e1 = ld3(tri_pos + 12 * (size_t)(offset + t) + 4) - v0;
        f3 e2 = ld3(tri_pos + 12 * (size_t)(offset + t) + 8) - v0;
        for (uint32_t i = 0; i < B.subdivision; i++) {
            if (n * B.subdivision + i >= N) break;
            uint32_t r = svalues[n * B.subdivision + i];
            f3 o = ld3(rays + 6 * (size_t)r), d = ld3(rays + 6 * (size_t)r + 3);
            float th = ray_triangle(o + d * EPS, d, v0, e1, e2);
            if (th > EPS) {
                uint32_t rec = (uint32_t)pack_hit(th, offset + t, 1);
                if (rec < times[r]) times[r] = rec;
            }
        }
    }
}

// RT:1776-2017 via RT:3033-3101
void oracle_ReflectionRayColoringWrapper(const float* tri_pos, const float* tri_nrm, uint32_t W, uint32_t H,
                                         uint32_t L, const float* eyep, int create_rays, const uint32_t* times,
                                         float* rays, uint32_t* flags, uint32_t* keys, uint32_t* values,
                                         uint32_t* pixels) {
    f3 eye = ld3(eyep);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            uint32_t r = x + y * W * B.light_maximum;
            uint32_t rec = times[r], t = rec & 0x000FFFFFu;
            if (rec == MISS32) { if (create_rays) flags[r] = 1; continue; }
            int32_t m = B.scene.tri_mat[3 * (size_t)t];
            f3 v0 = ld3(tri_pos + 12 * (size_t)t), v1 = ld3(tri_pos + 12 * (size_t)t + 4), v2 = ld3(tri_pos + 12 * (size_t)t + 8);
            f3 o = ld3(rays + 6 * (size_t)r), d = ld3(rays + 6 * (size_t)r + 3);
            float th = ray_triangle(o + d * EPS, d, v0, v1 - v0, v2 - v0);
            f3 P = o + d * (EPS + th);
            float aABC = length(cross(v1 - v0, v2 - v0));
            float aPBC = length(cross(v1 - P, v2 - P));
            float aPCA = length(cross(v0 - P, v2 - P));
            const float* tn = tri_nrm + 12 * (size_t)t;
            f3 N = ((aPBC / aABC) * ld3(tn) + (aPCA / aABC) * ld3(tn + 4)) +
                   ((1.0f - (aPBC / aABC)) - (aPCA / aABC)) * ld3(tn + 8);
            f4 Kd = ld4(B.scene.mat_diffuse + 4 * (size_t)m), Ks = ld4(B.scene.mat_specular + 4 * (size_t)m);
            f3 color{0, 0, 0};
            for (uint32_t l = 0; l < L; l++) color = color + light_term(&B.scene, l, P, N, eye, Kd, Ks, 0.75f, true);
            pixels[x + y * W] += rgb_to_int(color.x * Ks.w, color.y * Ks.w, color.z * Ks.w);
            if (create_rays) {
                f3 nd = reflect(d, N);
                st3(rays + 6 * (size_t)r, P); st3(rays + 6 * (size_t)r + 3, nd);
                keys[r] = reflection_key_ref(P, nd);
                values[r] = r;
                flags[r] = 0;
            }
        }
}

}  // extern "C"
