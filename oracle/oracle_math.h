// oracle_math.h -- FP32 scalar/vector math for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY: this directory is the correctness oracle for the
// ray-space-hierarchy pipeline.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product library never does.
//
// Vector helpers restate helper_math.h (dependencies/cuda/include/helper_math.h):
//   dot            :1248-1251   a.x*b.x + a.y*b.y + a.z*b.z (left to right)
//   length         :1291-1294   sqrtf(dot(v, v))
//   normalize      :1309-1313   v * rsqrtf(dot(v, v)); the host rsqrtf is
//                               1.0f / sqrtf(x) (:57-60), used here on purpose
//   reflect        :1411-1414   i - 2.0f * n * dot(n, i)
//   cross          :1438-1441
//   clamp(float)   :1152-1155   fmaxf(a, fminf(f, b)) with IEEE maxNum/minNum
// The kernels in projects/CUDA-Engine/source/RayTracer.cu call CUDA's acosf,
// cosf, tanf, sinf, powf and atan2f.  Their NVIDIA-specific last-bit results
// cannot be reproduced here, so this file defines its own FP32 functions from
// the Maclaurin series of asin/sin/cos/exp and the atanh series of log,
// evaluated by Horner's rule in a fixed operation order.  Every operation is an
// IEEE-754 single-precision +,-,*,/ or sqrt, so any implementation that
// performs the same sequence with contraction disabled (-ffp-contract=off)
// obtains the same bits.  Accuracy against float64 libm is checked in
// tests/test_oracle_math.py (a few ulp over the domains the pipeline uses).
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>

namespace oracle {

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

static inline f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
static inline f3 xyz(const f4& v) { return f3{v.x, v.y, v.z}; }
static inline f4 mk4(const f3& v, float w) { return f4{v.x, v.y, v.z, w}; }
static inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
static inline f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
static inline f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float length(f3 v) { return sqrtf(dot(v, v)); }
static inline f3 normalize(f3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return v * inv; }
static inline f3 reflect(f3 i, f3 n) { return i - 2.0f * n * dot(n, i); }

// IEEE-754 maxNum/minNum (a NaN operand yields the other operand), the
// semantics of CUDA's device fmaxf/fminf.
static inline float fmax_(float a, float b) { return std::fmax(a, b); }
static inline float fmin_(float a, float b) { return std::fmin(a, b); }
static inline float clampf(float f, float a, float b) { return fmax_(a, fmin_(f, b)); }
static inline uint32_t clampu(uint32_t f, uint32_t a, uint32_t b) {
    uint32_t m = f < b ? f : b;
    return m > a ? m : a;
}

// Float -> uint32 conversion with the GPU's saturating behaviour: NaN and
// negatives give 0, values >= 2^32 give 0xFFFFFFFF, otherwise truncation.
// (The reference relies on this at RayTracer.cu:153-154,188-193,1770; on the
// host the plain C cast would be undefined behaviour.)
static inline uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ---------------------------------------------------------------------------
// acos: acos(x) = pi/2 - asin(x) for |x| <= 1/2, and
//       acos(|x|) = 2 asin(sqrt((1-|x|)/2)) for |x| > 1/2 (pi - that for x < 0).
// asin(s) = s + s*z*P(z), z = s*s, P = Maclaurin coefficients
// c_n = (2n)! / (4^n (n!)^2 (2n+1)), n = 1..10 (truncation < 2e-9 at s = 1/2).
static inline float asin_poly(float s, float z) {
    float p = 0.008390335366129875f;         // 46189/5505024
    p = p * z + 0.009761609137058258f;       // 12155/1245184
    p = p * z + 0.01155180111527443f;        // 6435/557056
    p = p * z + 0.013964843936264515f;       // 143/10240
    p = p * z + 0.017352763563394547f;       // 231/13312
    p = p * z + 0.022372158244252205f;       // 63/2816
    p = p * z + 0.0303819440305233f;         // 35/1152
    p = p * z + 0.0446428582072258f;         // 5/112
    p = p * z + 0.07500000298023224f;        // 3/40
    p = p * z + 0.1666666716337204f;         // 1/6
    return s + s * (z * p);
}

static inline float acos_f(float x) {
    const float PI_F = 3.1415927410125732f;
    const float PIO2_F = 1.5707963705062866f;
    if (x != x) return x;
    float a = std::fabs(x);
    if (a > 1.0f) return std::nanf("");
    if (a <= 0.5f) {
        float z = x * x;
        return PIO2_F - asin_poly(x, z);
    }
    float z = (1.0f - a) * 0.5f;
    float s = sqrtf(z);
    float r = 2.0f * asin_poly(s, z);
    return x < 0.0f ? PI_F - r : r;
}

// ---------------------------------------------------------------------------
// sin/cos/tan: Cody-Waite reduction by pi/2 with a three-part constant
// (1.5703125 has 8 significant bits, so k*PIO2_A is exact for |k| < 2^16),
// then Maclaurin polynomials on |r| <= pi/4 (sin to r^9, cos to r^10).
struct sincos_r { float s, c; int q; };

static inline sincos_r reduce_sincos(float x) {
    const float TWO_OVER_PI = 0.6366197466850281f;
    const float PIO2_A = 1.5703125f;
    const float PIO2_B = 0.0004838705062866211f;
    const float PIO2_C = -4.371138828673793e-08f;
    float k = rintf(x * TWO_OVER_PI);
    float r = ((x - k * PIO2_A) - k * PIO2_B) - k * PIO2_C;
    float z = r * r;
    float ps = 2.7557318844628753e-06f;
    ps = ps * z + -0.00019841270113829523f;
    ps = ps * z + 0.008333333767950535f;
    ps = ps * z + -0.1666666716337204f;
    float s = r + r * (z * ps);
    float pc = -2.755731998149713e-07f;
    pc = pc * z + 2.4801587642286904e-05f;
    pc = pc * z + -0.0013888889225199819f;
    pc = pc * z + 0.0416666679084301f;
    float c = (1.0f - 0.5f * z) + (z * z) * pc;
    sincos_r out;
    out.s = s; out.c = c; out.q = ((int)k) & 3;
    return out;
}

static inline bool trig_arg_ok(float x) { return x == x && std::fabs(x) < 65536.0f; }

static inline float sin_f(float x) {
    if (!trig_arg_ok(x)) return std::nanf("");
    sincos_r r = reduce_sincos(x);
    switch (r.q) { case 0: return r.s; case 1: return r.c; case 2: return -r.s; default: return -r.c; }
}
static inline float cos_f(float x) {
    if (!trig_arg_ok(x)) return std::nanf("");
    sincos_r r = reduce_sincos(x);
    switch (r.q) { case 0: return r.c; case 1: return -r.s; case 2: return -r.c; default: return r.s; }
}
static inline float tan_f(float x) {
    if (!trig_arg_ok(x)) return std::nanf("");
    sincos_r r = reduce_sincos(x);
    return (r.q & 1) ? -(r.c / r.s) : (r.s / r.c);
}

// ---------------------------------------------------------------------------
// pow(x, y) for x >= 0 (the only use is the Blinn-Phong specular term,
// RayTracer.cu:1678,1870,1983): exp2(y * log2(x)).
// log2(x) = e + 2*atanh(f)*log2(e), f = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)).
// exp2(w) = 2^n * exp(r*ln2), n = rint(w), |r| <= 1/2, Maclaurin to u^7.
static inline float log2_f(float x) {
    // x is finite and > 0
    int e = 0;
    if (x < 1.1754943508222875e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t b = fbits(x);
    e += (int)((b >> 23) & 0xFFu) - 127;
    float m = bitsf((b & 0x007FFFFFu) | 0x3F800000u);   // [1, 2)
    if (m > 1.4142135381698608f) { m = m * 0.5f; e += 1; }
    float f = (m - 1.0f) / (m + 1.0f);
    float z = f * f;
    float p = 0.09090909090909091f;          // 1/11
    p = p * z + 0.1111111111111111f;          // 1/9
    p = p * z + 0.14285714285714285f;         // 1/7
    p = p * z + 0.2f;                         // 1/5
    p = p * z + 0.3333333432674408f;          // 1/3
    float at = f + f * (z * p);               // atanh(f)
    return (float)e + (2.0f * at) * 1.4426950216293335f;
}

static inline float exp2_f(float w) {
    if (w != w) return w;
    if (w >= 128.0f) return INFINITY;
    if (w < -150.0f) return 0.0f;
    float n = rintf(w);
    float r = w - n;
    float u = r * 0.6931471824645996f;
    float p = 0.00019841270113829523f;        // 1/7!
    p = p * u + 0.0013888889225199819f;       // 1/6!
    p = p * u + 0.008333333767950535f;        // 1/5!
    p = p * u + 0.0416666679084301f;          // 1/4!
    p = p * u + 0.1666666716337204f;          // 1/3!
    p = p * u + 0.5f;                         // 1/2!
    p = p * u + 1.0f;
    p = p * u + 1.0f;
    int ni = (int)n;
    if (ni > 127) { p = p * 2.0f; ni -= 1; }
    if (ni < -126) { p = p * 7.888609052210118e-31f; ni += 100; }   // 2^-100
    return p * bitsf((uint32_t)(ni + 127) << 23);
}

static inline float pow_f(float x, float y) {
    if (x != x || y != y) return std::nanf("");
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (x < 0.0f) return std::nanf("");
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    if (x == INFINITY) return y > 0.0f ? INFINITY : 0.0f;
    return exp2_f(y * log2_f(x));
}

// atan2(v, v) -- both reference hashes call atan2 with equal arguments
// (RayTracer.cu:153,188,192).  The point (v, v) lies on the diagonal, so the
// result takes one of four values: pi/4 (v > 0), +0 (v = +0), -pi (v = -0),
// -3pi/4 (v < 0), NaN for NaN.  The float constants are the nearest FP32.
static inline float atan2_diag(float v) {
    if (v != v) return v;
    if (v > 0.0f) return 0.7853981852531433f;
    if (v < 0.0f) return -2.356194496154785f;
    return std::signbit(v) ? -3.1415927410125732f : 0.0f;
}

}  // namespace oracle
