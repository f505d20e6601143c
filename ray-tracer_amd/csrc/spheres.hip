// spheres.hip -- object bounding spheres for scene loading (host code).
//
// The reference computes each object's sphere once, on the CPU, as the
// smallest ball enclosing the object's vertices (Gaertner's Miniball,
// projects/CUDA-Engine/source/BoundingSphere.cpp:10-62), then "revises" the
// radius: it starts from squared_radius^2 (quirk Q12: the square of the
// squared radius), raises it to the largest vertex distance and caps it at
// 1e5.  This file computes the same smallest ball with Welzl's move-to-front
// algorithm in double precision:
//   mtf(n, S): B = ball with the support set S on its boundary;
//              for each point p of the first n of the list, in order:
//                if p is outside B: mtf(index of p, S + {p}), move p to the
//                front of the list, B = the ball found;
// the ball of a support set of 1-4 points is the circumsphere of those points
// (the smallest sphere through them, for 2 or 3 points the one centred in
// their affine hull).  Its expected cost is linear in the number of points.
//
// mode 0 (RSH_SPHERE_MINIMAL): centre rounded to float, radius = the largest
//   vertex distance from that float centre (exact and as evaluated in float),
//   one ulp up: every vertex lies in the float sphere.
// mode 1 (RSH_SPHERE_REFERENCE): the reference's radius, min(1e5,
//   max(r^4, largest vertex distance)) with r^2 the ball's squared radius.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/rsh.h"

namespace {

struct P3 {
    double x, y, z;
};
inline P3 sub(const P3& a, const P3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 add(const P3& a, const P3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 mul(const P3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(const P3& a, const P3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross(const P3& a, const P3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Ball {
    P3 c{0, 0, 0};
    double r2 = -1.0;     // empty ball
    bool contains(const P3& p) const {
        const P3 d = sub(p, c);
        const double q = dot(d, d);
        return q <= r2 * (1.0 + 1e-12) + 1e-300;
    }
};

// Smallest ball with every point of s (1-4 affinely independent points, up to
// rounding) on its boundary.  Near-degenerate sets fall back to the smallest
// ball of a sub-pair / sub-triple that holds the rest.
Ball ball_of(const P3* s, int n);

Ball ball2(const P3& a, const P3& b) {
    Ball B;
    B.c = mul(add(a, b), 0.5);
    const P3 d = sub(a, b);
    B.r2 = 0.25 * dot(d, d);
    return B;
}

Ball ball3(const P3& p0, const P3& p1, const P3& p2) {
    const P3 a = sub(p1, p0), b = sub(p2, p0);
    const P3 n = cross(a, b);
    const double nn = dot(n, n);
    const double scale = dot(a, a) * dot(b, b);
    if (!(nn > 1e-24 * scale)) {     // collinear: the farthest pair
        Ball B = ball2(p0, p1), C = ball2(p0, p2), D = ball2(p1, p2);
        if (C.r2 > B.r2) B = C;
        if (D.r2 > B.r2) B = D;
        return B;
    }
    // circumcentre: p0 + (|a|^2 (b x n) + |b|^2 (n x a)) / (2 |n|^2)
    const P3 t = add(mul(cross(b, n), dot(a, a)), mul(cross(n, a), dot(b, b)));
    Ball B;
    const P3 o = mul(t, 0.5 / nn);
    B.c = add(p0, o);
    B.r2 = dot(o, o);
    return B;
}

Ball ball4(const P3& p0, const P3& p1, const P3& p2, const P3& p3) {
    // 2 (p_i - p0) . x = |p_i - p0|^2, centre p0 + x (Cramer's rule)
    const P3 a = sub(p1, p0), b = sub(p2, p0), c = sub(p3, p0);
    const double det = dot(a, cross(b, c));
    const double scale = std::sqrt(dot(a, a) * dot(b, b) * dot(c, c));
    if (!(std::fabs(det) > 1e-12 * scale)) {     // coplanar: the best triple that holds all four
        const P3 pts[4] = {p0, p1, p2, p3};
        Ball best;
        best.r2 = INFINITY;
        for (int skip = 0; skip < 4; skip++) {
            P3 t[3];
            int k = 0;
            for (int i = 0; i < 4; i++)
                if (i != skip) t[k++] = pts[i];
            Ball B = ball3(t[0], t[1], t[2]);
            if (B.contains(pts[skip]) && B.r2 < best.r2) best = B;
        }
        if (std::isfinite(best.r2)) return best;
        return ball3(p0, p1, p2);
    }
    const double aa = dot(a, a), bb = dot(b, b), cc = dot(c, c);
    const P3 x = mul(add(add(mul(cross(b, c), aa), mul(cross(c, a), bb)), mul(cross(a, b), cc)), 0.5 / det);
    Ball B;
    B.c = add(p0, x);
    B.r2 = dot(x, x);
    return B;
}

Ball ball_of(const P3* s, int n) {
    Ball B;
    switch (n) {
        case 0: return B;
        case 1: B.c = s[0]; B.r2 = 0.0; return B;
        case 2: return ball2(s[0], s[1]);
        case 3: return ball3(s[0], s[1], s[2]);
        default: return ball4(s[0], s[1], s[2], s[3]);
    }
}

struct Mtf {
    std::vector<P3> pts;
    P3 sup[4];
    int ns = 0;
    // points [0, end) inside the ball with the support set on its boundary
    Ball run(size_t end) {
        Ball B = ball_of(sup, ns);
        if (ns == 4) return B;
        for (size_t i = 0; i < end; i++) {
            if (!B.contains(pts[i])) {
                sup[ns++] = pts[i];
                B = run(i);
                ns--;
                std::rotate(pts.begin(), pts.begin() + i, pts.begin() + i + 1);     // move to front
            }
        }
        return B;
    }
};

// Sphere of object k (float3[2] at out).
void object_sphere(const float* tri_pos, uint32_t first, uint32_t last, uint32_t mode, float* o) {
    Mtf m;
    m.pts.reserve(3 * (size_t)(last - first + 1));
    for (uint32_t t = first; t <= last; t++)
        for (int v = 0; v < 3; v++) {
            const float* p = tri_pos + 4 * (3 * (size_t)t + v);
            m.pts.push_back({p[0], p[1], p[2]});
        }
    const Ball B = m.run(m.pts.size());
    float c[3] = {(float)B.c.x, (float)B.c.y, (float)B.c.z};
    // largest vertex distance from the float centre: exact (rounded up) and
    // as a float evaluation of |v - c| would give it (the reference's
    // revision step), plus one ulp
    double dmax = 0.0;
    float fmax = 0.0f;
    for (const P3& p : m.pts) {
        const double dx = p.x - c[0], dy = p.y - c[1], dz = p.z - c[2];
        dmax = std::max(dmax, std::sqrt(dx * dx + dy * dy + dz * dz));
        const float fx = (float)p.x - c[0], fy = (float)p.y - c[1], fz = (float)p.z - c[2];
        fmax = std::max(fmax, std::sqrt(fx * fx + fy * fy + fz * fz));
    }
    float r = (float)dmax;
    if ((double)r < dmax) r = std::nextafter(r, INFINITY);
    r = std::nextafter(std::max(r, fmax), INFINITY);
    if (mode == 1) {
        const float sq = (float)B.r2;
        r = std::min(100000.0f, std::max(sq * sq, r));
    }
    o[0] = c[0]; o[1] = c[1]; o[2] = c[2];
    o[3] = r; o[4] = (float)first; o[5] = (float)last;
}

}  // namespace

extern "C" {

// Smallest enclosing ball of the vertices of objects [0, K): object k spans
// triangles ranges[2k] .. ranges[2k + 1] (inclusive) of `tri_pos`, float4 per
// vertex (x, y, z, w), 3 vertices per triangle.  out: float3[2K], the
// reference's object-sphere layout (centre | radius, first, last).
rsh_status rsh_object_spheres(const float* tri_pos, uint32_t T, const uint32_t* ranges, uint32_t K, uint32_t mode,
                              float* out) {
    if (!tri_pos || !ranges || !out || mode > 1) return RSH_INVALID_ARG;
    for (uint32_t k = 0; k < K; k++)
        if (ranges[2 * k] > ranges[2 * k + 1] || ranges[2 * k + 1] >= T) return RSH_INVALID_ARG;
    // objects are independent: a few host threads take every n-th one
    const unsigned nt = std::max(1u, std::min(16u, std::min(std::thread::hardware_concurrency(), K / 8 + 1)));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; w++)
        pool.emplace_back([=] {
            for (uint32_t k = w; k < K; k += nt)
                object_sphere(tri_pos, ranges[2 * k], ranges[2 * k + 1], mode, out + 6 * (size_t)k);
        });
    for (auto& t : pool) t.join();
    return RSH_OK;
}

}  // extern "C"
