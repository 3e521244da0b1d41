// Collocation and Galerkin potential / field integrals of uniformly charged
// triangles (3D) and segments (2D), shared by the host mesh builder (refinement
// decisions) and the gfx950 fill kernels (link coefficients).  Everything is a pure function of plain
// scalars so the same source compiles for the host (g++) and the device (hipcc).
//
// Semantics follow FasterCap 6.0.7:
//   potential_opt_3d      Solver/Potential.cpp:1561-1711  (CPotential::PotentialOpt, 3D)
//   potential_num_3d      Solver/Potential.cpp:2077-2153  (CPotential::PotentialNumerical)
//   enfield_num_3d        Solver/Potential.cpp:2172-2254  (CPotential::EnFieldNumerical)
//   potential_opt_2d      Solver/Potential.cpp:1866-1949  (CPotential::PotentialOpt, 2D)
//   enfield_2d            Solver/Potential.cpp:1967-2060  (CPotential::EnField, 2D)
//   quadrature rules 2/3/5 Solver/Potential.cpp:112-128, 133-149, 181-206
//   constants             Solver/Potential.h:41-57
// Geometry is stored in float (as the reference stores C3DVector_float) and all
// arithmetic is carried out in the template type T (double for the FP64 path).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define FCG_HD __host__ __device__ inline
#define FCG_UNROLL _Pragma("unroll")
#else
#define FCG_HD inline
#define FCG_UNROLL
#endif

#define FCG_FOUR_PI_TIMES_E0 1.1126500559756e-10
#define FCG_PI_TIMES_E0 2.781625139939e-11
#define FCG_TWO_PI_TIMES_E0 5.563250279878e-11
#define FCG_TWO_TIMES_E0 1.7708375634e-11
#define FCG_PI 3.1415926535897932384626433832795
#define FCG_PI_HALF 1.5707963267948966192313216916398
#define FCG_GEO_EPS 1E-6
#define FCG_TOL_EPS 1E-12
#define FCG_LOG_FOUR_PLUS_TWO 3.3862943611198906188344642429164

namespace fcg {

// Quadrature rules on the unit triangle, given as (x, y) barycentric-like
// coordinates with the homogeneous third coordinate equal to one.  Rule ids and
// point orders are the reference's (Strang-Fix #1, #2, #4).
struct QuadRule {
    int n;
    double x[6], y[6], w[6];
};

FCG_HD QuadRule quad_rule(int rule) {
    QuadRule q{};
    if (rule == 2) {
        q.n = 3;
        const double a = 1.0 / 6.0, c = 4.0 / 6.0, w = 1.0 / 3.0;
        q.x[0] = c; q.y[0] = a;
        q.x[1] = a; q.y[1] = c;
        q.x[2] = a; q.y[2] = a;
        q.w[0] = q.w[1] = q.w[2] = w;
    } else if (rule == 3) {
        q.n = 3;
        const double w = 1.0 / 3.0;
        q.x[0] = 0.0; q.y[0] = 0.5;
        q.x[1] = 0.5; q.y[1] = 0.0;
        q.x[2] = 0.5; q.y[2] = 0.5;
        q.w[0] = q.w[1] = q.w[2] = w;
    } else {  // rule 5
        q.n = 6;
        const double a = 0.659027622374092, b = 0.231933368553031, c = 0.109039009072877;
        const double w = 1.0 / 6.0;
        q.x[0] = a; q.y[0] = b;
        q.x[1] = a; q.y[1] = c;
        q.x[2] = b; q.y[2] = a;
        q.x[3] = b; q.y[3] = c;
        q.x[4] = c; q.y[4] = a;
        q.x[5] = c; q.y[5] = b;
        for (int i = 0; i < 6; ++i) q.w[i] = w;
    }
    return q;
}

template <typename T>
struct Vec3 {
    T x, y, z;
};

template <typename T>
FCG_HD Vec3<T> mk3(T x, T y, T z) { return Vec3<T>{x, y, z}; }
template <typename T>
FCG_HD Vec3<T> sub3(Vec3<T> a, Vec3<T> b) { return Vec3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
FCG_HD T dot3(Vec3<T> a, Vec3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
FCG_HD Vec3<T> cross3(Vec3<T> a, Vec3<T> b) {
    return Vec3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T>
FCG_HD T norm3(Vec3<T> a) { return sqrt(dot3(a, a)); }

template <typename T>
FCG_HD Vec3<T> vtx(const float* v, int i) {
    return Vec3<T>{(T)v[3 * i], (T)v[3 * i + 1], (T)v[3 * i + 2]};
}

// Potential at r of a triangle carrying unit total charge (analytic edge sum,
// divided by the area).  v = 9 floats, vertex-major.
template <typename T>
FCG_HD T potential_opt_3d(Vec3<T> r, const float* v) {
    Vec3<T> V[3] = {vtx<T>(v, 0), vtx<T>(v, 1), vtx<T>(v, 2)};
    Vec3<T> side[3] = {sub3(V[1], V[0]), sub3(V[2], V[1]), sub3(V[0], V[2])};
    Vec3<T> n = cross3(side[0], side[1]);
    const T nmod = norm3(n);
    n.x = n.x / nmod;
    n.y = n.y / nmod;
    n.z = n.z / nmod;
    const T nr = dot3(n, r);
    const Vec3<T> rho{r.x - n.x * nr, r.y - n.y * nr, r.z - n.z * nr};
    Vec3<T> Pmp[3], Rmp[3];
    T Rn[3];
    // both edge loops are unrolled so the small arrays live in registers
FCG_UNROLL
    for (int i = 0; i < 3; ++i) {
        const T nv = dot3(n, V[i]);
        const Vec3<T> rhomp{V[i].x - n.x * nv, V[i].y - n.y * nv, V[i].z - n.z * nv};
        Pmp[i] = sub3(rhomp, rho);
        Rmp[i] = sub3(V[i], r);
        Rn[i] = norm3(Rmp[i]);
    }
    const T d = -dot3(n, Rmp[0]);
    const T dnorm = fabs(d);
    T result = 0;
FCG_UNROLL
    for (int i = 0; i < 3; ++i) {
        const int j = (i == 2) ? 0 : i + 1;
        const T sl = norm3(side[i]);
        const Vec3<T> l{side[i].x / sl, side[i].y / sl, side[i].z / sl};
        const Vec3<T> u = cross3(l, n);
        const T P0signed = dot3(Pmp[j], u);
        const T P0 = fabs(P0signed);
        const T lplus = dot3(Pmp[j], l);
        const T lminus = dot3(Pmp[i], l);
        const T R0 = sqrt(P0 * P0 + d * d);
        const T lmax = (lplus > lminus) ? fabs(lplus) : fabs(lminus);
        if (P0 / lmax > (T)FCG_GEO_EPS && fabs((Rn[j] + lplus) / lmax) > (T)FCG_GEO_EPS &&
            fabs((Rn[i] + lminus) / lmax) > (T)FCG_GEO_EPS) {
            const T R0sq = R0 * R0;
            const T tan1 = atan(P0 * lplus / (R0sq + dnorm * Rn[j]));
            const T tan2 = atan(P0 * lminus / (R0sq + dnorm * Rn[i]));
            const T ln = log((Rn[j] + lplus) / (Rn[i] + lminus));
            const T sgn = (T)((P0signed > 0) - (P0signed < 0));
            result += sgn * (P0 * ln - dnorm * (tan1 - tan2));
        }
    }
    return (result * (T)2.0) / nmod;
}

// Shared local frame of the quadrature routines.
template <typename T>
struct TriFrame {
    Vec3<T> v0, x, y;
    T v01, v02, v12, jac, area;
};

template <typename T>
FCG_HD TriFrame<T> tri_frame(const float* v) {
    TriFrame<T> f;
    const Vec3<T> V0 = vtx<T>(v, 0), V1 = vtx<T>(v, 1), V2 = vtx<T>(v, 2);
    const Vec3<T> s1 = sub3(V1, V0), s2 = sub3(V2, V1), s3 = sub3(V0, V2);
    const Vec3<T> normal{s1.y * s2.z - s1.z * s2.y, -s1.x * s2.z + s1.z * s2.x, s1.x * s2.y - s1.y * s2.x};
    const T nm = norm3(normal);
    const Vec3<T> z{normal.x / nm, normal.y / nm, normal.z / nm};
    const T m1 = norm3(s1);
    f.x = Vec3<T>{s1.x / m1, s1.y / m1, s1.z / m1};
    f.y = cross3(f.x, z);
    f.v01 = m1;
    f.v02 = -dot3(s3, f.x);
    f.v12 = -dot3(s3, f.y);
    f.jac = (f.v02 - f.v01) * f.v12 - f.v12 * f.v02;
    f.area = nm / (T)2.0;
    f.v0 = V0;
    return f;
}

template <typename T>
FCG_HD Vec3<T> tri_point(const TriFrame<T>& f, T px, T py) {
    const T p0 = (f.v01 - f.v02) * px + f.v02 * ((T)1.0 - py);
    const T p1 = f.v12 * ((T)1.0 - px - py);
    return Vec3<T>{f.v0.x + f.x.x * p0 + f.y.x * p1, f.v0.y + f.x.y * p0 + f.y.y * p1,
                   f.v0.z + f.x.z * p0 + f.y.z * p1};
}

// Mean of 1/|p - r| over the triangle by the given rule (unit total charge).
template <typename T>
FCG_HD T potential_num_3d(Vec3<T> r, const float* v, int rule) {
    const TriFrame<T> f = tri_frame<T>(v);
    const QuadRule q = quad_rule(rule);
    T result = 0;
    for (int i = 0; i < q.n; ++i) {
        const Vec3<T> p = tri_point(f, (T)q.x[i], (T)q.y[i]);
        const T pot = (T)1 / norm3(sub3(p, r));
        result += (T)q.w[i] * pot;
    }
    result *= f.jac / (T)2.0;
    result /= f.area;
    return result;
}

// Normal electric field component (direction tn) at r of the unit-charge triangle.
template <typename T>
FCG_HD T enfield_num_3d(Vec3<T> r, const float* v, Vec3<T> tn, int rule) {
    const TriFrame<T> f = tri_frame<T>(v);
    const QuadRule q = quad_rule(rule);
    T result = 0;
    for (int i = 0; i < q.n; ++i) {
        const Vec3<T> p = tri_point(f, (T)q.x[i], (T)q.y[i]);
        const Vec3<T> dd = sub3(r, p);
        const T rd = norm3(dd);
        const T pot = dot3(dd, tn) / (rd * rd * rd);
        result += (T)q.w[i] * pot;
    }
    result *= f.jac / (T)2.0;
    result /= f.area;
    return result;
}

// Device forms of the two quadratures above.  Point (px, py) of the local frame
// is the barycentric point py V0 + px V1 + (1 - px - py) V2, and the frame's
// Jacobian equals twice the area, so each rule reduces to a weighted sum over
// points formed from the vertices relative to r: no frame, no normalisations.
// Mathematically identical; results differ from the frame forms by rounding only.
template <typename T>
FCG_HD T inv_sqrt(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrt(x);
#else
    return (T)1 / sqrt(x);
#endif
}

template <typename T, int RULE>
FCG_HD T potential_quad_3d(Vec3<T> r, const float* v) {
    const Vec3<T> a = sub3(vtx<T>(v, 0), r), b = sub3(vtx<T>(v, 1), r), c = sub3(vtx<T>(v, 2), r);
    const QuadRule q = quad_rule(RULE);
    T result = 0;
FCG_UNROLL
    for (int i = 0; i < q.n; ++i) {
        const T w0 = (T)q.y[i], w1 = (T)q.x[i], w2 = (T)1 - w0 - w1;
        const Vec3<T> d{w0 * a.x + w1 * b.x + w2 * c.x, w0 * a.y + w1 * b.y + w2 * c.y,
                        w0 * a.z + w1 * b.z + w2 * c.z};
        result += (T)q.w[i] * inv_sqrt(dot3(d, d));
    }
    return result;
}

template <typename T, int RULE>
FCG_HD T enfield_quad_3d(Vec3<T> r, const float* v, Vec3<T> tn) {
    const Vec3<T> a = sub3(r, vtx<T>(v, 0)), b = sub3(r, vtx<T>(v, 1)), c = sub3(r, vtx<T>(v, 2));
    const QuadRule q = quad_rule(RULE);
    T result = 0;
FCG_UNROLL
    for (int i = 0; i < q.n; ++i) {
        const T w0 = (T)q.y[i], w1 = (T)q.x[i], w2 = (T)1 - w0 - w1;
        const Vec3<T> d{w0 * a.x + w1 * b.x + w2 * c.x, w0 * a.y + w1 * b.y + w2 * c.y,
                        w0 * a.z + w1 * b.z + w2 * c.z};
        const T ir = inv_sqrt(dot3(d, d));
        result += (T)q.w[i] * dot3(d, tn) * (ir * ir * ir);
    }
    return result;
}

// ---------------------------------------------------------------- Galerkin (3D)
// The Galerkin scheme's coefficients (-g): means over the destination triangle
// instead of values at its centroid.
//   auto_3d          Solver/Potential.cpp:994-1017   (CPotential::Auto)
//   mutual_half_3d   Solver/Potential.cpp:2485-2625  (MutualHalfNumerical)
//   mutual_dhalf_3d  Solver/Potential.cpp:2802-2964  (MutualDHalfNumerical)
//   mutual_full_3d   Solver/Potential.cpp:3138-3347, 3790-4003 (Mutual_2thOrd/_4thOrd_FullNum)
//   mutual_dfull_3d  Solver/Potential.cpp:4197-4377, 3349-3577 (MutualDFullNumerical,
//                    MutualD_2thOrd_FullNum)
//   rules 10 and 17  Solver/Potential.cpp:345-389, 717-800
// All include the 1 / (4 pi e0) factor, as the reference's do.  As for the
// collocation quadratures, rule point (px, py) is the barycentric point
// py V0 + px V1 + (1 - px - py) V2 and the frame's Jacobian, twice the area,
// cancels against the area division: each mean is a weighted sum over points.

// Rules 10 (Strang-Fix #8, 9 points, precision 6) and 17 (19 points, precision 9).
struct QuadRule19 {
    int n;
    double x[19], y[19], w[19];
};

FCG_HD QuadRule19 quad_rule19(int rule) {
    QuadRule19 q{};
    if (rule == 10) {
        const double a = 0.124949503233232, b = 0.437525248383384, c = 0.797112651860071,
                     d = 0.165409927389841, e = 0.037477420750088;
        const double u = 0.205950504760887, v = 0.063691414286223;
        const double xs[9] = {a, b, b, c, c, d, d, e, e}, ys[9] = {b, a, b, d, e, c, e, c, d};
        q.n = 9;
        for (int i = 0; i < 9; ++i) {
            q.x[i] = xs[i];
            q.y[i] = ys[i];
            q.w[i] = i < 3 ? u : v;
        }
    } else {  // 17
        const double a = 1.0 / 3.0, b = 0.02063496160252593, c = 0.4896825191987370, d = 0.1258208170141290,
                     e = 0.4370895914929355, f = 0.6235929287619356, g = 0.1882035356190322,
                     r = 0.9105409732110941, s = 0.04472951339445297, t = 0.7411985987844980,
                     u = 0.03683841205473626, v = 0.22196288916076574;
        const double w1 = 0.09713579628279610, w2 = 0.03133470022713983, w3 = 0.07782754100477543,
                     w4 = 0.07964773892720910, w5 = 0.02557767565869810, w6 = 0.04328353937728940;
        const double xs[19] = {a, b, c, c, d, e, e, f, g, g, r, s, s, t, t, u, u, v, v};
        const double ys[19] = {a, c, b, c, e, d, e, g, f, g, s, r, s, u, v, t, v, t, u};
        const double ws[19] = {w1, w2, w2, w2, w3, w3, w3, w4, w4, w4, w5, w5, w5, w6, w6, w6, w6, w6, w6};
        q.n = 19;
        for (int i = 0; i < 19; ++i) {
            q.x[i] = xs[i];
            q.y[i] = ys[i];
            q.w[i] = ws[i];
        }
    }
    return q;
}

// barycentric point (px, py) of triangle v
template <typename T>
FCG_HD Vec3<T> bary_point(const float* v, T px, T py) {
    const T w0 = py, w1 = px, w2 = (T)1 - px - py;
    const Vec3<T> a = vtx<T>(v, 0), b = vtx<T>(v, 1), c = vtx<T>(v, 2);
    return Vec3<T>{w0 * a.x + w1 * b.x + w2 * c.x, w0 * a.y + w1 * b.y + w2 * c.y, w0 * a.z + w1 * b.z + w2 * c.z};
}

// the triangle's own potential averaged over it, closed form in the side lengths
// a, b, c and the semiperimeter p: -(ln(1 - a/p)/a + ln(1 - b/p)/b + ln(1 - c/p)/c) / (3 pi e0)
template <typename T>
FCG_HD T auto_3d(const float* v) {
    const Vec3<T> V0 = vtx<T>(v, 0), V1 = vtx<T>(v, 1), V2 = vtx<T>(v, 2);
    const T a = norm3(sub3(V1, V0)), b = norm3(sub3(V2, V1)), c = norm3(sub3(V0, V2));
    const T ph = (a + b + c) / (T)2;
    return -(((T)1 / a) * log((T)1 - a / ph) + ((T)1 / b) * log((T)1 - b / ph) + ((T)1 / c) * log((T)1 - c / ph)) /
           ((T)3 * (T)FCG_PI_TIMES_E0);
}

// mean over triangle v1 (rule RULE: 10 or 17) of the analytic potential of v2
template <typename T, int RULE>
FCG_HD T mutual_half_3d(const float* v1, const float* v2) {
    const QuadRule19 q = quad_rule19(RULE);
    T result = 0;
    for (int i = 0; i < q.n; ++i)
        result += (T)q.w[i] * potential_opt_3d<T>(bary_point<T>(v1, (T)q.x[i], (T)q.y[i]), v2);
    return result / (T)FCG_FOUR_PI_TIMES_E0;
}

// the same mean of the divided difference of v2's potential along n (step h)
template <typename T, int RULE>
FCG_HD T mutual_dhalf_3d(const float* v1, const float* v2, Vec3<T> n, T h) {
    const QuadRule19 q = quad_rule19(RULE);
    T result = 0;
    for (int i = 0; i < q.n; ++i) {
        const Vec3<T> p = bary_point<T>(v1, (T)q.x[i], (T)q.y[i]);
        const Vec3<T> pp{p.x + n.x * h, p.y + n.y * h, p.z + n.z * h};
        const Vec3<T> pm{p.x - n.x * h, p.y - n.y * h, p.z - n.z * h};
        result += (T)q.w[i] * ((potential_opt_3d<T>(pm, v2) - potential_opt_3d<T>(pp, v2)) / ((T)2 * h));
    }
    return result / (T)FCG_FOUR_PI_TIMES_E0;
}

// Points of the double (product) rules: the 3-point rule 2 and the 6-point
// degree-4 rule (Strang-Fix), with the pair weights the reference tabulates:
// 1/9 for order 2; for order 4 the products of the two point weights
// 0.109951743655322 and 0.223381589678011.
struct FullRule {
    int n;
    double x[6], y[6];
};
FCG_HD FullRule full_rule(int order) {
    FullRule r{};
    if (order == 4) {
        const double a = 0.816847572980459, b = 0.091576213509771, c = 0.108103018168070, d = 0.445948490915965;
        const double xs[6] = {a, b, b, c, d, d}, ys[6] = {b, a, b, d, c, d};
        r.n = 6;
        for (int i = 0; i < 6; ++i) {
            r.x[i] = xs[i];
            r.y[i] = ys[i];
        }
    } else {
        const double a = 0.66666666666666666666666666666667, b = 0.16666666666666666666666666666667;
        const double xs[3] = {a, b, b}, ys[3] = {b, a, b};
        r.n = 3;
        for (int i = 0; i < 3; ++i) {
            r.x[i] = xs[i];
            r.y[i] = ys[i];
        }
    }
    return r;
}
FCG_HD double full_weight(int order, int i, int j) {
    if (order != 4) return 0.11111111111111111111111111111111;
    const bool a = i < 3, b = j < 3;
    return a && b ? 0.012089385932845641681938923684
                  : (a || b ? 0.024561195285594988334146524542 : 0.049899334607075270538656916121);
}

// double mean of 1/r over v1 and v2 by the product rule of order ORDER (2 or 4)
template <typename T, int ORDER>
FCG_HD T mutual_full_3d(const float* v1, const float* v2) {
    const FullRule q = full_rule(ORDER);
    Vec3<T> p2[6];
FCG_UNROLL
    for (int j = 0; j < q.n; ++j) p2[j] = bary_point<T>(v2, (T)q.x[j], (T)q.y[j]);
    T result = 0;
FCG_UNROLL
    for (int i = 0; i < q.n; ++i) {
        const Vec3<T> p1 = bary_point<T>(v1, (T)q.x[i], (T)q.y[i]);
FCG_UNROLL
        for (int j = 0; j < q.n; ++j) {
            const Vec3<T> dd = sub3(p1, p2[j]);
            result += (T)full_weight(ORDER, i, j) * ((T)1 / norm3(dd));
        }
    }
    return result / (T)FCG_FOUR_PI_TIMES_E0;
}

// double mean of (r1 - r2).n / |r1 - r2|^3 over v1 and v2: RULE 10 or 5 with the
// products of the rule's weights, RULE 0 the 3 x 3 table of order 2
template <typename T, int RULE>
FCG_HD T mutual_dfull_3d(const float* v1, const float* v2, Vec3<T> n) {
    double px[19] = {}, py[19] = {}, w[19] = {};
    int m;
    if (RULE == 0) {
        const FullRule f = full_rule(2);
        m = f.n;
        for (int i = 0; i < m; ++i) {
            px[i] = f.x[i];
            py[i] = f.y[i];
        }
    } else if (RULE == 5) {
        const QuadRule q = quad_rule(5);
        m = q.n;
        for (int i = 0; i < m; ++i) {
            px[i] = q.x[i];
            py[i] = q.y[i];
            w[i] = q.w[i];
        }
    } else {
        const QuadRule19 q = quad_rule19(RULE);
        m = q.n;
        for (int i = 0; i < m; ++i) {
            px[i] = q.x[i];
            py[i] = q.y[i];
            w[i] = q.w[i];
        }
    }
    T result = 0;
    for (int i = 0; i < m; ++i) {
        const Vec3<T> p1 = bary_point<T>(v1, (T)px[i], (T)py[i]);
        for (int j = 0; j < m; ++j) {
            const Vec3<T> dd = sub3(p1, bary_point<T>(v2, (T)px[j], (T)py[j]));
            const T rd = norm3(dd);
            const T wij = RULE == 0 ? (T)full_weight(2, i, j) : (T)(w[i] * w[j]);
            result += wij * (dot3(dd, n) / (rd * rd * rd));
        }
    }
    return result / (T)FCG_FOUR_PI_TIMES_E0;
}

// 2D: potential (up to the log constant, unscaled) at r of a unit-charge segment
// v = (x0, y0, z0, x1, y1, z1) float storage, z ignored.
template <typename T>
FCG_HD T potential_opt_2d(T rx, T ry, const float* v) {
    const T sx = (T)v[3] - (T)v[0], sy = (T)v[4] - (T)v[1];
    const T segMod = sqrt(sx * sx + sy * sy);
    const T ux = sx / segMod, uy = sy / segMod;
    const T yx = -uy, yy = ux;
    const T ox = (T)v[0] + sx / (T)2.0, oy = (T)v[1] + sy / (T)2.0;
    const T a = segMod / (T)2.0;
    const T px = rx - ox, py = ry - oy;
    const T rox = px * ux + py * uy;
    const T roy = px * yx + py * yy;
    const T xpa = rox + a, xma = rox - a;
    T result;
    if (fabs(roy) < (T)FCG_TOL_EPS && (fabs(xpa) < (T)FCG_TOL_EPS || fabs(xma) < (T)FCG_TOL_EPS)) {
        result = -a * (log(a * a) + (T)FCG_LOG_FOUR_PLUS_TWO);
    } else {
        const T ysq = roy * roy;
        const T add1 = -xpa * (log(ysq + xpa * xpa) / (T)2.0 - (T)1.0);
        const T add2 = xma * (log(ysq + xma * xma) / (T)2.0 - (T)1.0);
        const T add3 = (fabs(xpa) < (T)FCG_TOL_EPS) ? fabs(roy) * (T)FCG_PI_HALF : roy * atan2(roy, xpa);
        const T add4 = (fabs(xma) < (T)FCG_TOL_EPS) ? -fabs(roy) * (T)FCG_PI_HALF : -roy * atan2(roy, xma);
        result = add1 + add2 + add3 + add4;
    }
    return result / segMod;
}

// 2D: field component along (nx, ny) at r of a unit-charge segment.
template <typename T>
FCG_HD T enfield_2d(T rx, T ry, const float* v, T nx, T ny) {
    const T sx = (T)v[3] - (T)v[0], sy = (T)v[4] - (T)v[1];
    const T segMod = sqrt(sx * sx + sy * sy);
    const T ux = sx / segMod, uy = sy / segMod;
    const T yx = -uy, yy = ux;
    const T ox = (T)v[0] + sx / (T)2.0, oy = (T)v[1] + sy / (T)2.0;
    const T a = segMod / (T)2.0;
    const T px = rx - ox, py = ry - oy;
    const T rox = px * ux + py * uy;
    const T roy = px * yx + py * yy;
    T lnx = nx * ux + ny * uy;
    T lny = nx * yx + ny * yy;
    const T lnm = sqrt(lnx * lnx + lny * lny);
    if (lnm > 0) {
        lnx = lnx / lnm;
        lny = lny / lnm;
    }
    const T xpa = rox + a, xma = rox - a;
    T result;
    if (fabs(roy) < (T)FCG_TOL_EPS && (fabs(xpa) < (T)FCG_TOL_EPS || fabs(xma) < (T)FCG_TOL_EPS)) {
        result = -a * (log(a * a) + (T)FCG_LOG_FOUR_PLUS_TWO);
    } else {
        const T ysq = roy * roy;
        const T ex = (log(ysq + xpa * xpa) - log(ysq + xma * xma)) / (T)2.0;
        const T ey = atan2(roy, xma) - atan2(roy, xpa);
        result = ex * lnx + ey * lny;
    }
    return result / segMod;
}

}  // namespace fcg
