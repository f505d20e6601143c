// rsh_device.hpp -- FP32 device math and geometry of the RSH pipeline (gfx950).
//
// Every function is a sequence of IEEE-754 single-precision +,-,*,/,sqrt with
// contraction disabled (the library is built with -ffp-contract=off), so the
// host restatement used for testing obtains identical bits.  The reference's
// CUDA transcendental calls (acosf, cosf, sinf, tanf, powf, atan2f in
// projects/CUDA-Engine/source/RayTracer.cu) are replaced by series
// evaluations defined here: acos from the Maclaurin series of asin, sin/cos
// from their Maclaurin series after Cody-Waite reduction by pi/2, pow as
// exp2(y*log2 x) with log from the atanh series and exp from its Maclaurin
// series.  Vector helpers follow helper_math.h semantics
// (dependencies/cuda/include/helper_math.h:1152-1441) with normalize taken as
// v * (1.0f / sqrtf(dot(v, v))).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsh {

struct v3 { float x, y, z; };

__device__ __forceinline__ v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ v3 xyz(float4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len(v3 v) { return sqrtf(dot(v, v)); }
__device__ __forceinline__ v3 normalize(v3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return v * inv; }
__device__ __forceinline__ v3 reflect(v3 i, v3 n) { return i - 2.0f * n * dot(n, i); }
// helper_math.h:1422-1431, pow(x, 2) taken as the correctly rounded x * x;
// (0, 0, 0) on total internal reflection.
__device__ inline v3 refract(v3 i, v3 n, float ior) {
    const float c = dot(i, n);
    const float k = 1.0f - (ior * ior) * (1.0f - c * c);
    if (k < 0.0f) return v3{0.0f, 0.0f, 0.0f};
    return i * ior - n * (ior * c + sqrtf(k));
}
__device__ __forceinline__ float clampf(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }
__device__ __forceinline__ uint32_t clampu(uint32_t f, uint32_t a, uint32_t b) {
    uint32_t m = f < b ? f : b;
    return m > a ? m : a;
}
// Saturating float -> uint32 (NaN/negative -> 0, >= 2^32 -> 0xFFFFFFFF).
__device__ __forceinline__ uint32_t f2u(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

constexpr float EPS = 0.01f;                    // RayTracer.cu:34
constexpr float HALF_PI_F = 3.1415927f / 2.0f;  // Constants.h:2-3
constexpr float PI_F = 3.1415927f;
constexpr uint32_t MISS32 = 0xFFFFFFFFu;
constexpr unsigned long long MISS64 = 0xFFFFFFFFFFFFFFFFull;

// ------------------------------------------------------------------ acos --
__device__ __forceinline__ float asin_series(float s, float z) {
    float p = 0.008390335366129875f;
    p = p * z + 0.009761609137058258f;
    p = p * z + 0.01155180111527443f;
    p = p * z + 0.013964843936264515f;
    p = p * z + 0.017352763563394547f;
    p = p * z + 0.022372158244252205f;
    p = p * z + 0.0303819440305233f;
    p = p * z + 0.0446428582072258f;
    p = p * z + 0.07500000298023224f;
    p = p * z + 0.1666666716337204f;
    return s + s * (z * p);
}

__device__ inline float acos_s(float x) {
    if (x != x) return x;
    float a = fabsf(x);
    if (a > 1.0f) return __builtin_nanf("");
    if (a <= 0.5f) return 1.5707963705062866f - asin_series(x, x * x);
    float z = (1.0f - a) * 0.5f;
    float r = 2.0f * asin_series(sqrtf(z), z);
    return x < 0.0f ? 3.1415927410125732f - r : r;
}

// ------------------------------------------------------- sin / cos / tan --
struct sc_t { float s, c; int q; };

__device__ inline sc_t reduce_sc(float x) {
    float k = rintf(x * 0.6366197466850281f);
    float r = ((x - k * 1.5703125f) - k * 0.0004838705062866211f) - k * -4.371138828673793e-08f;
    float z = r * r;
    float ps = 2.7557318844628753e-06f;
    ps = ps * z + -0.00019841270113829523f;
    ps = ps * z + 0.008333333767950535f;
    ps = ps * z + -0.1666666716337204f;
    float pc = -2.755731998149713e-07f;
    pc = pc * z + 2.4801587642286904e-05f;
    pc = pc * z + -0.0013888889225199819f;
    pc = pc * z + 0.0416666679084301f;
    sc_t o;
    o.s = r + r * (z * ps);
    o.c = (1.0f - 0.5f * z) + (z * z) * pc;
    o.q = ((int)k) & 3;
    return o;
}
__device__ __forceinline__ bool trig_ok(float x) { return x == x && fabsf(x) < 65536.0f; }
__device__ inline float sin_s(float x) {
    if (!trig_ok(x)) return __builtin_nanf("");
    sc_t r = reduce_sc(x);
    return r.q == 0 ? r.s : r.q == 1 ? r.c : r.q == 2 ? -r.s : -r.c;
}
__device__ inline float cos_s(float x) {
    if (!trig_ok(x)) return __builtin_nanf("");
    sc_t r = reduce_sc(x);
    return r.q == 0 ? r.c : r.q == 1 ? -r.s : r.q == 2 ? -r.c : r.s;
}
__device__ inline float tan_s(float x) {
    if (!trig_ok(x)) return __builtin_nanf("");
    sc_t r = reduce_sc(x);
    return (r.q & 1) ? -(r.c / r.s) : (r.s / r.c);
}

// ------------------------------------------------------------------- pow --
__device__ inline float log2_s(float x) {
    int e = 0;
    if (x < 1.1754943508222875e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t b = __float_as_uint(x);
    e += (int)((b >> 23) & 0xFFu) - 127;
    float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.4142135381698608f) { m = m * 0.5f; e += 1; }
    float f = (m - 1.0f) / (m + 1.0f);
    float z = f * f;
    float p = 0.09090909090909091f;
    p = p * z + 0.1111111111111111f;
    p = p * z + 0.14285714285714285f;
    p = p * z + 0.2f;
    p = p * z + 0.3333333432674408f;
    float at = f + f * (z * p);
    return (float)e + (2.0f * at) * 1.4426950216293335f;
}
__device__ inline float exp2_s(float w) {
    if (w != w) return w;
    if (w >= 128.0f) return __builtin_inff();
    if (w < -150.0f) return 0.0f;
    float n = rintf(w);
    float r = w - n;
    float u = r * 0.6931471824645996f;
    float p = 0.00019841270113829523f;
    p = p * u + 0.0013888889225199819f;
    p = p * u + 0.008333333767950535f;
    p = p * u + 0.0416666679084301f;
    p = p * u + 0.1666666716337204f;
    p = p * u + 0.5f;
    p = p * u + 1.0f;
    p = p * u + 1.0f;
    int ni = (int)n;
    if (ni > 127) { p = p * 2.0f; ni -= 1; }
    if (ni < -126) { p = p * 7.888609052210118e-31f; ni += 100; }
    return p * __uint_as_float((uint32_t)(ni + 127) << 23);
}
__device__ inline float pow_s(float x, float y) {
    if (x != x || y != y) return __builtin_nanf("");
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_inff();
    if (x == __builtin_inff()) return y > 0.0f ? __builtin_inff() : 0.0f;
    return exp2_s(y * log2_s(x));
}

// atan2(v, v): the point (v, v) is on the diagonal, so the value is one of
// pi/4 (v > 0), +0 (v = +0), -pi (v = -0), -3pi/4 (v < 0); NaN stays NaN.
__device__ __forceinline__ float atan2_diag(float v) {
    if (v != v) return v;
    if (v > 0.0f) return 0.7853981852531433f;
    if (v < 0.0f) return -2.356194496154785f;
    return (__float_as_uint(v) >> 31) ? -3.1415927410125732f : 0.0f;
}

// ------------------------------------------------------ integer division --
// Division by a divisor c fixed over many numerators: floor(x / c) is the high
// word of x * m, m = ceil(2^32 / c) (c >= 2).  With m c = 2^32 + e, 0 <= e < c,
// and x = q c + r: x m / 2^32 = q + r / c + x e / (c 2^32), whose floor is q
// while r + x e / 2^32 < c, which x (c - 1) < 2^32 guarantees.  Callers check
// that bound for their largest numerator once, on the host (div_hi_ok), and
// pass the verdict to the kernel, so no lane takes a per-value branch.
__host__ __device__ inline bool div_hi_ok(uint64_t max_x, uint32_t c) {
    return c <= 1u || max_x <= 0xFFFFFFFFull / (c - 1u);
}
__device__ __forceinline__ uint32_t div_hi_magic(uint32_t c) {
    return c <= 1u ? 0u : (uint32_t)((0x100000000ull + c - 1ull) / c);
}
__device__ __forceinline__ uint32_t div_hi(uint32_t x, uint32_t c, uint32_t m) {
    return c == 1u ? x : __umulhi(x, m);
}

// ------------------------------------------------------------------ keys --
// RayTracer.cu:147-168 CreateShadowRayIndex
__device__ inline uint32_t shadow_key_ref(v3 d, uint32_t l) {
    uint32_t index = (l + 1u) << 30;
    index += clampu(f2u((atan2_diag(d.x) + HALF_PI_F) * 5.0f), 0u, 31u) << 15;
    index += clampu(f2u(acos_s(d.z) * 5.0f), 0u, 31u);
    return index + 1u;
}
// RayTracer.cu:170-198 CreateReflectionRayIndex
__device__ inline uint32_t reflection_key_ref(v3 o, v3 d) {
    float distance = len(o);
    uint32_t index = clampu(f2u((atan2_diag(o.x) + HALF_PI_F) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u(acos_s(o.z / distance) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u((atan2_diag(d.x) + HALF_PI_F) * 5.0f), 0u, 30u);
    index = (index << 5) | clampu(f2u(acos_s(d.z) * 5.0f), 0u, 30u);
    return index + 1u;
}

// RayTracer.cu:200-227 CreateRefractionRayIndex: the distance from the origin,
// then the reflection key's four fields; 7 + 7 + 7 + 7 + 5 bits in 32, so the
// distance field's top bits leave the word, as in the reference.
__device__ inline uint32_t refraction_key_ref(v3 o, v3 d) {
    float distance = len(o);
    uint32_t index = clampu(f2u(distance), 0u, 128u);
    index = (index << 7) | clampu(f2u((atan2_diag(o.x) + HALF_PI_F) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u(acos_s(o.z / distance) * 20.0f), 0u, 128u);
    index = (index << 7) | clampu(f2u((atan2_diag(d.x) + HALF_PI_F) * 5.0f), 0u, 30u);
    index = (index << 5) | clampu(f2u(acos_s(d.z) * 5.0f), 0u, 30u);
    return index + 1u;
}

struct HashParams {
    uint32_t cell_bits, dir_bits, layout, segment_shift;
    float min_x, min_y, min_z, scale_x, scale_y, scale_z;
};

// Octahedral key: [type:2][cell z|y|x][u|v] (layout 0) or [type][u|v][cell] (layout 1).
__device__ inline uint32_t key_oct(const HashParams& h, v3 o, v3 d, uint32_t type) {
    const uint32_t cb = h.cell_bits, db = h.dir_bits;
    const uint32_t cmax = (1u << cb) - 1u, dmax = (1u << db) - 1u;
    uint32_t cx = min(f2u((o.x - h.min_x) * h.scale_x), cmax);
    uint32_t cy = min(f2u((o.y - h.min_y) * h.scale_y), cmax);
    uint32_t cz = min(f2u((o.z - h.min_z) * h.scale_z), cmax);
    float s = (fabsf(d.x) + fabsf(d.y)) + fabsf(d.z);
    float px = d.x / s, py = d.y / s;
    if (d.z < 0.0f) {
        float ax = fabsf(px), ay = fabsf(py);
        float nx = (1.0f - ay) * (px >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - ax) * (py >= 0.0f ? 1.0f : -1.0f);
        px = nx; py = ny;
    }
    const float fs = (float)(1u << db);
    uint32_t u = min(f2u((px * 0.5f + 0.5f) * fs), dmax);
    uint32_t v = min(f2u((py * 0.5f + 0.5f) * fs), dmax);
    uint32_t cell = (cz << (2 * cb)) | (cy << cb) | cx;
    uint32_t dir = (u << db) | v;
    uint32_t body;
    if (h.layout == 2) {
        // bit-interleaved, most significant level first: z y x u v at each level
        body = 0;
        const uint32_t top = cb > db ? cb : db;
        for (uint32_t i = top; i-- > 0;) {
            if (i < cb) body = (body << 3) | (((cz >> i) & 1u) << 2) | (((cy >> i) & 1u) << 1) | ((cx >> i) & 1u);
            if (i < db) body = (body << 2) | (((u >> i) & 1u) << 1) | ((v >> i) & 1u);
        }
    } else {
        body = h.layout == 0 ? ((cell << (2 * db)) | dir) : ((dir << (3 * cb)) | cell);
    }
    return (type << (3 * cb + 2 * db)) | body;
}

// -------------------------------------------------------------- geometry --
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Correctly rounded 1/x.  For |x| in [2^-125, 2^126) (exponent field 2..252)
// one Newton step on v_rcp_f32, y = y0 + y0 (1 - x y0) in two FMAs, equals the
// IEEE quotient for every input (an exhaustive device check over all such
// floats, rsh_selfcheck_rcp, is part of the test suite); other inputs, zeros,
// denormals, infinities and NaN take the IEEE division.
__device__ __forceinline__ float rcp_rn(float x) {
    const uint32_t ex = (__float_as_uint(x) >> 23) & 0xFFu;
    if (ex - 2u <= 250u) {
        const float y0 = hw_rcp(x);
        return __builtin_fmaf(__builtin_fmaf(-x, y0, 1.0f), y0, y0);
    }
    return 1.0f / x;
}

// RayTracer.cu:257-280 (reciprocal correctly rounded instead of __fdividef)
__device__ __forceinline__ float ray_triangle(v3 o, v3 d, v3 v0, v3 e1, v3 e2) {
    v3 tvec = o - v0;
    v3 pvec = cross(d, e2);
    float det = rcp_rn(dot(e1, pvec));
    float u = dot(tvec, pvec) * det;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    v3 qvec = cross(tvec, e1);
    float v = dot(d, qvec) * det;
    if (v < 0.0f || (u + v) > 1.0f) return -1.0f;
    return dot(e2, qvec) * det;
}

// RayTracer.cu:283-298 SphereNodeIntersection
__device__ __forceinline__ bool sphere_node(float4 sphere, float4 cone, float4 tri, float cosine, float tangent) {
    if (cone.w == HALF_PI_F) return true;
    v3 cd = xyz(cone);
    v3 s2t = xyz(tri) - xyz(sphere);
    float pr = dot(s2t, cd);
    v3 proj = pr * cd;
    return (pr * tangent + (sphere.w + tri.w) / cosine) >= len(s2t - proj);
}

// 1/cos from v_rcp_f32 for sphere_node_fast and the cluster test, or NaN for
// a cosine outside [1e-30, 1e30] (there the estimate is not trusted).
__device__ __forceinline__ float node_rcos(float cosine) {
    const float a = fabsf(cosine);
    return (a >= 1e-30f && a <= 1e30f) ? hw_rcp(cosine) : __int_as_float(0x7FC00000);
}

// sphere_node with the same result, decided from hardware estimates when
// they are clearly on one side.  With rcos = 1/cos and the square root from
// v_rcp_f32 / v_sqrt_f32 (each within 1 ulp), the estimated sides
// lhs' = pr tan + (r + rt) rcos and rhs' = |v| differ from the correctly
// rounded lhs and rhs by at most 2^-21 (|pr tan| + |(r + rt) rcos|) and
// 2^-22 |v| (plus denormal ulps for values below 1e-19).  When
// |lhs' - rhs'| exceeds M = 2^-18 (|pr tan| + |(r + rt) rcos| + |v|) + 2^-60,
// four times those bounds, lhs' >= rhs' decides as lhs >= rhs would;
// otherwise, and whenever a value is NaN or infinite (rcos NaN included),
// the exact expression of sphere_node decides.
__device__ __forceinline__ bool sphere_node_fast(float4 sphere, float4 cone, float4 tri, float cosine, float tangent,
                                                 float rcos) {
    if (cone.w == HALF_PI_F) return true;
    const v3 cd = xyz(cone);
    const v3 s2t = xyz(tri) - xyz(sphere);
    const float pr = dot(s2t, cd);
    const v3 v = s2t - pr * cd;
    const float dd = dot(v, v);
    const float a = pr * tangent;
    const float R = sphere.w + tri.w;
    const float q = R * rcos;
    const float rs = hw_sqrt(dd);
    const float diff = (a + q) - rs;
    const float M = 0x1p-18f * (fabsf(a) + fabsf(q) + rs) + 0x1p-60f;
    if (diff > M) return true;
    if (diff < -M) return false;
    return (a + R / cosine) >= sqrtf(dd);
}

// Ray pre-test before ray_triangle: false only when the ray's line (origin o,
// direction d, any length) passes the sphere (center c, radius r) that holds
// the triangle at a distance beyond r by a margin of 1e-4 (|c - o|_1 + r) +
// 1e-6.  ray_triangle's FP32 evaluation can accept a point off the triangle
// only within ~1e-6 of those magnitudes (its rounding is relative to |o - v0|,
// |d| and the edges), so a rejected ray would be a miss there too.  NaN passes.
// dd = dot(d, d), computed once per ray.
__device__ __forceinline__ bool ray_may_hit_sphere(v3 o, v3 d, float dd, float4 sph) {
    const v3 w = xyz(sph) - o;
    const v3 x = cross(w, d);
    const float rr = sph.w + 1e-4f * (fabsf(w.x) + fabsf(w.y) + fabsf(w.z) + sph.w) + 1e-6f;
    return !(dot(x, x) > rr * rr * dd);
}

// ray_may_hit_sphere on the part of the line a hit can use: also false when
// the sphere's parameter range along the ray, (w.d -+ r|d|) / |d|^2 with
// w = c - o, ends before lo or starts beyond hi, with r widened by
// 1e-3 (|w|_1 + r) + 1e-5 (ten times the line test's margin: ray_triangle's
// parameter carries the same relative rounding as its hit point).  The
// tests compare squares, (w.d - hi |d|^2)^2 > r'^2 |d|^2 with the left side
// positive, so no square root is taken.  lo is the hit threshold EPS; hi the
// shadow ray's length or the parameter of the closest hit found so far (a
// triangle starting beyond it cannot give a smaller record).  NaN or
// infinite bounds never reject.  dd = dot(d, d), computed once per ray.
__device__ __forceinline__ bool ray_may_hit_sphere_range(v3 o, v3 d, float dd, float4 sph, float lo, float hi) {
    const v3 w = xyz(sph) - o;
    const v3 x = cross(w, d);
    const float l1 = fabsf(w.x) + fabsf(w.y) + fabsf(w.z) + sph.w;
    const float rr = sph.w + 1e-4f * l1 + 1e-6f;
    if (dot(x, x) > rr * rr * dd) return false;
    const float rt = sph.w + 1e-3f * l1 + 1e-5f;
    const float A = rt * rt * dd;
    const float wd = dot(w, d);
    const float e = wd - hi * dd;     // > rt |d|: the sphere starts beyond hi
    const float f = lo * dd - wd;     // > rt |d|: the sphere ends before lo
    return !((e > 0.0f && e * e > A) || (f > 0.0f && f * f > A));
}

// Conservative cluster pre-test for sphere_node.  With J = I - cd cd^T the
// triangle test passes iff f(c) = |J (c - s)| - ((c - s).cd) tan <= (r + rt)/cos.
// f is Lipschitz with constant L = sqrt(1 + (tan k)^2 + 2 tan k |1 - k^2|),
// k = |cd| <= sqrt(2).  For a triangle sphere (c, rt) inside a cluster sphere
// (q, rho), |c - q| <= rho - rt, so a passing triangle gives
// f(q) <= r/cos + rho * max(L, 1/cos).  The cluster is rejected only when
// f(q) exceeds that bound by E, a margin of 1e-4 of the magnitudes involved.
// Both evaluations use the hardware square root and reciprocal (v_sqrt_f32,
// v_rcp_f32, within 1 ulp): their error, like the FP32 rounding of either
// evaluation, is ~1e-6 of those magnitudes, far inside E.  Nodes outside the
// argument's assumptions (the HALF_PI shortcut, cos <= 0, tan < 0,
// non-finite values, |cd| far from 1) never reject; a NaN anywhere passes.
// node_bound returns (max(L, 1/cos) slightly inflated, node_rcos(cos)), with
// -1 in the first for a node the argument does not cover (cluster_may_pass
// then never rejects).
__device__ __forceinline__ float2 node_bound(float4 sphere, float4 cone, float cosine, float tangent) {
    const float k = hw_sqrt(dot(xyz(cone), xyz(cone)));
    const float tk = tangent * k;
    const float L = hw_sqrt(1.0f + tk * tk + 2.0f * tk * fabsf(1.0f - k * k));
    const float rc = node_rcos(cosine);
    const float lmax = fmaxf(L, rc) * 1.00001f;
    const bool ok = cone.w != HALF_PI_F && cosine >= 1e-30f && rc == rc && tangent >= 0.0f && k > 0.5f && k < 1.4f &&
                    isfinite(cosine) && isfinite(tangent) && isfinite(lmax) && isfinite(sphere.w) &&
                    isfinite(sphere.x) && isfinite(sphere.y) && isfinite(sphere.z);
    return make_float2(ok ? lmax : -1.0f, rc);
}

// bound = node_bound(sphere, cone, cos, tan)
__device__ __forceinline__ bool cluster_may_pass(float4 sphere, float4 cone, float tangent, float2 bound, float4 cl) {
    const float lmax = bound.x;
    if (!(lmax > 0.0f)) return true;
    const v3 cd = xyz(cone);
    const v3 s2q = xyz(cl) - xyz(sphere);
    const float pr = dot(s2q, cd);
    const v3 v = s2q - pr * cd;
    const float f = hw_sqrt(dot(v, v)) - pr * tangent;
    const float rho = cl.w;
    const float allow = sphere.w * bound.y + rho * lmax;
    const float E = 1e-4f * (hw_sqrt(dot(s2q, s2q)) + rho + (fabsf(pr) + rho) * tangent + (sphere.w + rho) * lmax) +
                    1e-4f;
    return !(f > allow + E);
}

// RayTracer.cu:301-376 CreateTriangleBoundingSphere
__device__ inline float4 triangle_sphere(v3 a, v3 b, v3 c) {
    float dotABAB = dot(b - a, b - a);
    float dotABAC = dot(b - a, c - a);
    float dotACAC = dot(c - a, c - a);
    float d = 2.0f * (dotABAB * dotACAC - dotABAC * dotABAC);
    v3 ref = a, center;
    float radius;
    if (fabsf(d) <= EPS) {
        float dAB = len(a - b), dBC = len(b - c), dCA = len(c - a);
        if (dAB > dBC) {
            if (dAB > dCA) { center = (a + b) * 0.5f; radius = dAB * 0.5f; }
            else { center = (a + c) * 0.5f; radius = dCA * 0.5f; }
        } else {
            if (dBC > dCA) { center = (b + c) * 0.5f; radius = dBC * 0.5f; }
            else { center = (a + c) * 0.5f; radius = dCA * 0.5f; }
        }
        return f4(center, radius);
    }
    float s = (dotABAB * dotACAC - dotACAC * dotABAC) / d;
    float t = (dotACAC * dotABAB - dotABAB * dotABAC) / d;
    if (s <= 0.0f) center = (a + c) * 0.5f;
    else if (t <= 0.0f) center = (a + b) * 0.5f;
    else if (s + t >= 1.0f) { center = (b + c) * 0.5f; ref = b; }
    else center = (a + s * (b - a)) + t * (c - a);
    radius = len(center - ref);
    return f4(center, radius);
}

// RayTracer.cu:379-406 CreateHierarchyCone0 (mode 0) / guarded union (mode 1).
__device__ inline float4 cone0(float4 cone, v3 rdir, float s0, uint32_t mode) {
    (void)s0;
    if (mode == 0) {
        v3 cd = xyz(cone), rd = rdir;
        float spread = acos_s(dot(cd, rd));
        if (cone.w > spread) return cone;
        v3 q = normalize(dot(cd, rd) * cd - rd);
        v3 e = cd * cos_s(cone.w) + q * sin_s(cone.w);
        v3 nd = normalize(rd + e);
        return f4(nd, acos_s(dot(nd, rd)));
    }
    v3 cd = xyz(cone), rd = normalize(rdir);
    float c = clampf(dot(cd, rd), -1.0f, 1.0f);
    float spread = acos_s(c);
    if (cone.w >= spread) return cone;
    if (spread + cone.w >= PI_F) return f4(cd, HALF_PI_F);
    v3 qv = c * cd - rd;
    float qq = dot(qv, qv);
    if (!(qq > 0.0f)) return f4(cd, HALF_PI_F);
    v3 q = qv * (1.0f / sqrtf(qq));
    v3 e = cd * cos_s(cone.w) + q * sin_s(cone.w);
    v3 sum = rd + e;
    float ss = dot(sum, sum);
    if (!(ss > 0.0f)) return f4(cd, HALF_PI_F);
    v3 nd = sum * (1.0f / sqrtf(ss));
    float ns = acos_s(clampf(dot(nd, rd), -1.0f, 1.0f));
    return f4(nd, fminf(ns, HALF_PI_F));
}

// RayTracer.cu:408-425 CreateHierarchyConeN
__device__ inline float4 coneN(float4 c1, float4 c2, uint32_t mode) {
    v3 d1 = xyz(c1), d2 = xyz(c2);
    if (mode == 0) {
        float spread = acos_s(dot(d1, d2)) * 0.5f;
        v3 d = normalize(d1 + d2);
        return f4(d, clampf(spread + fmaxf(c1.w, c2.w), 0.0f, HALF_PI_F));
    }
    float spread = acos_s(clampf(dot(d1, d2), -1.0f, 1.0f)) * 0.5f;
    v3 sum = d1 + d2;
    float ss = dot(sum, sum);
    if (!(ss > 0.0f)) return f4(d1, HALF_PI_F);
    v3 d = sum * (1.0f / sqrtf(ss));
    return f4(d, clampf(spread + fmaxf(c1.w, c2.w), 0.0f, HALF_PI_F));
}

// RayTracer.cu:427-446 CreateHierarchySphere
__device__ __forceinline__ float4 sphere_union(float4 s1, float4 s2) {
    v3 c1 = xyz(s1), c2 = xyz(s2);
    float dist = len(c2 - c1);
    if (dist + s2.w <= s1.w) return s1;
    if (dist + s1.w <= s2.w) return s2;
    return f4((c1 + c2) * 0.5f, dist * 0.5f + fmaxf(s1.w, s2.w));
}

// RayTracer.cu:100-107 RgbToInt
__device__ __forceinline__ uint32_t rgb_to_int(float r, float g, float b) {
    r = clampf(r, 0.0f, 255.0f);
    g = clampf(g, 0.0f, 255.0f);
    b = clampf(b, 0.0f, 255.0f);
    return (uint32_t)((int)r | ((int)g << 8) | ((int)b << 16));
}

// RayTracer.cu:1770 packing (mode 1) and exact 64-bit record (mode 0).
__device__ __forceinline__ unsigned long long pack_hit(float t, uint32_t tri, uint32_t mode) {
    if (mode == 0) return ((unsigned long long)__float_as_uint(t) << 32) | (unsigned long long)tri;
    return (unsigned long long)((f2u(t * 10.0f + 1.0f) << 20) + tri);
}

struct Lights {
    const float4* pos;
    const float4* col;
    const float2* inten;
    uint32_t total;
};

// ----------------------------------------------------------- soft shadows --
// The reference's soft shadows (RayTracer.cu:1550-1603 and 1680-1697, compiled
// out by Constants.h:32): 16 points about the light on a 4 x 4 grid spaced by
// 1/8 of two plane axes, the grid's centre row and column left out.  Cell
// c = i + 4 j lies at offsets (a - 2, b - 2) with a = i + (i >= 2) and
// b = j + (j >= 2).  The first axis is the cross product of a coordinate axis
// with the shadow ray's direction d; the axis is picked by the reference's
// test against min(|x|, max(|y|, |z|)), which chooses x unless |x| is the
// largest component.  The axes are not normalised.
__device__ inline void soft_axes(v3 d, v3& a0, v3& a1) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const float m = fminf(ax, fmaxf(ay, az));
    if (ax == m) a0 = cross(mk(1.0f, 0.0f, 0.0f), d);
    else if (ay == m) a0 = cross(mk(0.0f, 1.0f, 0.0f), d);
    else a0 = cross(mk(0.0f, 0.0f, 1.0f), d);
    a1 = cross(d, a0);
}

__device__ inline v3 soft_point(v3 o, v3 a0, v3 a1, uint32_t c) {
    const uint32_t i = c & 3u, j = c >> 2;
    const float u = (float)((int)(i + (i >= 2u ? 1u : 0u)) - 2), w = (float)((int)(j + (j >= 2u ? 1u : 0u)) - 2);
    return (o + (a0 * u) * 0.125f) + (a1 * w) * 0.125f;
}

// The shadow ray from light position L towards surface point P
// (RayTracer.cu:682-740); soft: from grid cell c about the light instead.
// Ray creation and the gather of sorted shadow rays both call it, so the
// gather recomputes exactly the ray creation made.
__device__ inline void shadow_ray(v3 P, v3 L, bool soft, uint32_t c, v3& o, v3& d) {
    o = L;
    d = normalize(P - o);
    if (soft) {
        v3 a0, a1;
        soft_axes(d, a0, a1);
        o = soft_point(o, a0, a1, c);
        d = normalize(P - o);
    }
}

// Light scale of a pixel whose grid cells in `occluded` (bit c: cell c) are
// shadowed: log2(1 + 2 s), s the sum over open cells of 0.55/16 (the four
// inner cells), 0.5375/16 (the corners) or 0.525/16 (the rest), in cell order.
__device__ inline float soft_factor(uint32_t occluded) {
    float s = 0.0f;
    for (uint32_t c = 0; c < 16; c++) {
        if ((occluded >> c) & 1u) continue;
        if (c == 5 || c == 6 || c == 9 || c == 10) s += 0.55f / 16.0f;
        else if (c == 0 || c == 3 || c == 12 || c == 15) s += 0.5375f / 16.0f;
        else s += 0.525f / 16.0f;
    }
    return log2_s(1.0f + s * 2.0f);
}

// Blinn-Phong term of one light (RayTracer.cu:1655-1702, :1847-1876).
// soft: both components are scaled by `factor` last (soft shadows).
__device__ inline v3 light_term(const Lights& L, uint32_t l, v3 P, v3 N, v3 eye, float4 Kd, float4 Ks,
                                bool reflection, bool soft = false, float factor = 1.0f) {
    v3 color = mk(0.0f, 0.0f, 0.0f);
    v3 Ld = xyz(L.pos[l]) - P;
    float dist = len(Ld);
    Ld = normalize(Ld);
    v3 lc = xyz(L.col[l]);
    float2 I = L.inten[l];
    float att = 1.0f / ((0.75f + dist * 0.0005f) + (dist * dist) * 0.00005f);
    float diffuse = clampf(fmaxf(dot(Ld, N), 0.0f), 0.0f, 1.0f);
    if (diffuse > 0.0f) {
        v3 Hv = normalize(Ld - normalize(P - eye));
        float specular = clampf(pow_s(fmaxf(dot(Hv, N), 0.0f), Ks.w), 0.0f, 1.0f);
        v3 cd, cs;
        if (reflection) {
            cd = ((((xyz(Kd) * lc) * diffuse) * I.x) * att) * 0.75f;
            cs = ((((xyz(Ks) * lc) * specular) * I.y) * att) * 0.75f;
        } else {
            cd = (((xyz(Kd) * lc) * diffuse) * I.x) * att;
            cs = (((xyz(Ks) * lc) * specular) * I.y) * att;
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

}  // namespace rsh
