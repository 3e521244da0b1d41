// CPU restatement of FasterCap 6.0.7's BEM solve path: the checker of the gfx950
// path and the reported CPU baseline.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Written independently of the device kernels: plain arrays, scalar loops, the
// reference's sequential summation orders, OpenMP exactly where the reference has
// it (fill over nodes, MVP gather over nodes).
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

namespace {

const double kFourPiE0 = 1.1126500559756e-10;  // Potential.h:43
const double kTwoE0 = 1.7708375634e-11;        // Potential.h:44
const double kPi = 3.1415926535897932384626433832795;
const double kPiE0 = 2.781625139939e-11;  // Potential.h:41
const double kPiHalf = 1.5707963267948966192313216916398;
const double kGeoEps = 1E-6;    // POTENTIAL_GEO_EPS
const double kTolEps = 1E-12;   // POTENTIAL_TOL_EPS
const double kLog4p2 = 3.3862943611198906188344642429164;  // Potential.cpp:39

void vsub(const double* a, const double* b, double* o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
double vdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void vcross(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
double vnorm(const double* a) { return std::sqrt(vdot(a, a)); }

// Quadrature tables: InitNumerical rules 2, 3, 5 (Potential.cpp:112-128, 133-149, 181-206)
// and the Galerkin scheme's rules 10 (Strang-Fix #8, 9 points, :345-389) and 17 (19
// points, precision 9, :717-800).
struct Rule {
    int n;
    double px[19], py[19], w[19];
};
Rule make_rule(int id) {
    Rule r;
    std::memset(&r, 0, sizeof r);
    if (id == 2) {
        r.n = 3;
        r.px[0] = 4.0 / 6.0; r.py[0] = 1.0 / 6.0;
        r.px[1] = 1.0 / 6.0; r.py[1] = 4.0 / 6.0;
        r.px[2] = 1.0 / 6.0; r.py[2] = 1.0 / 6.0;
        for (int i = 0; i < 3; ++i) r.w[i] = 1.0 / 3.0;
    } else if (id == 3) {
        r.n = 3;
        r.px[0] = 0.0; r.py[0] = 0.5;
        r.px[1] = 0.5; r.py[1] = 0.0;
        r.px[2] = 0.5; r.py[2] = 0.5;
        for (int i = 0; i < 3; ++i) r.w[i] = 1.0 / 3.0;
    } else if (id == 5) {
        const double a = 0.659027622374092, b = 0.231933368553031, c = 0.109039009072877;
        const double xs[6] = {a, a, b, b, c, c}, ys[6] = {b, c, a, c, a, b};
        r.n = 6;
        for (int i = 0; i < 6; ++i) {
            r.px[i] = xs[i];
            r.py[i] = ys[i];
            r.w[i] = 1.0 / 6.0;
        }
    } else if (id == 10) {
        const double a = 0.124949503233232, b = 0.437525248383384, c = 0.797112651860071,
                     d = 0.165409927389841, e = 0.037477420750088;
        const double u = 0.205950504760887, v = 0.063691414286223;
        const double xs[9] = {a, b, b, c, c, d, d, e, e}, ys[9] = {b, a, b, d, e, c, e, c, d};
        r.n = 9;
        for (int i = 0; i < 9; ++i) {
            r.px[i] = xs[i];
            r.py[i] = ys[i];
            r.w[i] = i < 3 ? u : v;
        }
    } else {  // 17
        const double a = 1.0 / 3.0, b = 0.02063496160252593, c = 0.4896825191987370, d = 0.1258208170141290,
                     e = 0.4370895914929355, f = 0.6235929287619356, g = 0.1882035356190322,
                     rr = 0.9105409732110941, ss = 0.04472951339445297, t = 0.7411985987844980,
                     u = 0.03683841205473626, v = 0.22196288916076574;
        const double w1 = 0.09713579628279610, w2 = 0.03133470022713983, w3 = 0.07782754100477543,
                     w4 = 0.07964773892720910, w5 = 0.02557767565869810, w6 = 0.04328353937728940;
        const double xs[19] = {a, b, c, c, d, e, e, f, g, g, rr, ss, ss, t, t, u, u, v, v};
        const double ys[19] = {a, c, b, c, e, d, e, g, f, g, ss, rr, ss, u, v, t, v, t, u};
        const double ws[19] = {w1, w2, w2, w2, w3, w3, w3, w4, w4, w4, w5, w5, w5, w6, w6, w6, w6, w6, w6};
        r.n = 19;
        for (int i = 0; i < 19; ++i) {
            r.px[i] = xs[i];
            r.py[i] = ys[i];
            r.w[i] = ws[i];
        }
    }
    return r;
}
const Rule kRule2 = make_rule(2), kRule3 = make_rule(3), kRule5 = make_rule(5), kRule10 = make_rule(10),
           kRule17 = make_rule(17);
const Rule& rule_of(int id) {
    switch (id) {
        case 2: return kRule2;
        case 3: return kRule3;
        case 10: return kRule10;
        case 17: return kRule17;
        default: return kRule5;
    }
}

// Quadrature of a kernel over the triangle in the reference's local frame.
template <class F>
double tri_quadrature(const float* v, int rule, F kernel) {
    double V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) V[i][k] = v[3 * i + k];
    double s1[3], s2[3], s3[3], nrm[3], z[3], x[3], y[3];
    vsub(V[1], V[0], s1);
    vsub(V[2], V[1], s2);
    vsub(V[0], V[2], s3);
    nrm[0] = s1[1] * s2[2] - s1[2] * s2[1];
    nrm[1] = -s1[0] * s2[2] + s1[2] * s2[0];
    nrm[2] = s1[0] * s2[1] - s1[1] * s2[0];
    const double nm = vnorm(nrm);
    for (int k = 0; k < 3; ++k) z[k] = nrm[k] / nm;
    const double m1 = vnorm(s1);
    for (int k = 0; k < 3; ++k) x[k] = s1[k] / m1;
    vcross(x, z, y);
    const double d01 = m1, d02 = -vdot(s3, x), d12 = -vdot(s3, y);
    const double jac = (d02 - d01) * d12 - d12 * d02;
    const Rule& R = rule_of(rule);
    double acc = 0;
    for (int i = 0; i < R.n; ++i) {
        const double a0 = (d01 - d02) * R.px[i] + d02 * (1.0 - R.py[i]);
        const double a1 = d12 * (1.0 - R.px[i] - R.py[i]);
        double p[3];
        for (int k = 0; k < 3; ++k) p[k] = V[0][k] + x[k] * a0 + y[k] * a1;
        acc += R.w[i] * kernel(p);
    }
    acc *= jac / 2.0;
    return acc / (nm / 2.0);
}

struct Job {
    const fcg_job_desc* d;
    const fcg_node_geom& node(int i) const { return d->nodes[i]; }
};

void centroid(const fcg_node_geom& g, double* c) {
    c[0] = g.c[0];
    c[1] = g.c[1];
    c[2] = g.c[2];
}

}  // namespace

extern "C" {

double orc_potential_opt_3d(const double* r, const float* v) {
    double V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) V[i][k] = v[3 * i + k];
    double side[3][3];
    for (int i = 0; i < 3; ++i) vsub(V[(i + 1) % 3], V[i], side[i]);
    double n[3];
    vcross(side[0], side[1], n);
    const double nmod = vnorm(n);
    for (int k = 0; k < 3; ++k) n[k] /= nmod;
    const double nr = vdot(n, r);
    double rho[3];
    for (int k = 0; k < 3; ++k) rho[k] = r[k] - n[k] * nr;
    double P[3][3], Rv[3][3], Rlen[3];
    for (int i = 0; i < 3; ++i) {
        const double nv = vdot(n, V[i]);
        for (int k = 0; k < 3; ++k) {
            P[i][k] = (V[i][k] - n[k] * nv) - rho[k];
            Rv[i][k] = V[i][k] - r[k];
        }
        Rlen[i] = vnorm(Rv[i]);
    }
    const double d = -vdot(n, Rv[0]);
    const double ad = std::fabs(d);
    double total = 0.0;
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        const double len = vnorm(side[i]);
        double l[3], u[3];
        for (int k = 0; k < 3; ++k) l[k] = side[i][k] / len;
        vcross(l, n, u);
        const double p0s = vdot(P[j], u);
        const double p0 = std::fabs(p0s);
        const double lp = vdot(P[j], l);
        const double lm = vdot(P[i], l);
        const double r0 = std::sqrt(p0 * p0 + d * d);
        const double lmax = lp > lm ? std::fabs(lp) : std::fabs(lm);
        if (!(p0 / lmax > kGeoEps)) continue;
        if (!(std::fabs((Rlen[j] + lp) / lmax) > kGeoEps)) continue;
        if (!(std::fabs((Rlen[i] + lm) / lmax) > kGeoEps)) continue;
        const double t1 = std::atan(p0 * lp / (r0 * r0 + ad * Rlen[j]));
        const double t2 = std::atan(p0 * lm / (r0 * r0 + ad * Rlen[i]));
        const double lg = std::log((Rlen[j] + lp) / (Rlen[i] + lm));
        const double s = (p0s > 0) ? 1.0 : ((p0s < 0) ? -1.0 : 0.0);
        total += s * (p0 * lg - ad * (t1 - t2));
    }
    return total * 2.0 / nmod;
}

double orc_potential_num_3d(const double* r, const float* v, int rule) {
    return tri_quadrature(v, rule, [&](const double* p) {
        double dd[3];
        vsub(p, r, dd);
        return 1.0 / vnorm(dd);
    });
}

double orc_enfield_num_3d(const double* r, const float* v, const double* n, int rule) {
    return tri_quadrature(v, rule, [&](const double* p) {
        double dd[3];
        vsub(r, p, dd);
        const double rd = vnorm(dd);
        return vdot(dd, n) / (rd * rd * rd);
    });
}

// ---------------------------------------------------------------- Galerkin (3D)
// CPotential::Auto (Potential.cpp:994-1017): the self coefficient of a uniformly
// charged triangle (its potential averaged over itself), closed form in the side
// lengths and the semiperimeter.
double orc_auto_3d(const float* v) {
    double V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) V[i][k] = v[3 * i + k];
    double s1[3], s2[3], s3[3];
    vsub(V[1], V[0], s1);
    vsub(V[2], V[1], s2);
    vsub(V[0], V[2], s3);
    const double a = vnorm(s1), b = vnorm(s2), c = vnorm(s3);
    const double ph = (a + b + c) / 2;
    return -((1 / a) * std::log(1 - a / ph) + (1 / b) * std::log(1 - b / ph) + (1 / c) * std::log(1 - c / ph)) /
           (3 * kPiE0);
}

// CPotential::MutualHalfNumerical (Potential.cpp:2485-2625): the mean over triangle
// 1 (rule `rule`, the local frame) of triangle 2's analytic potential.
double orc_mutual_half_3d(const float* v1, const float* v2, int rule) {
    return tri_quadrature(v1, rule, [&](const double* p) { return orc_potential_opt_3d(p, v2); }) / kFourPiE0;
}

// CPotential::MutualDHalfNumerical (Potential.cpp:2802-2964): the same mean of the
// divided difference of triangle 2's potential along n (step h).
double orc_mutual_dhalf_3d(const float* v1, const float* v2, const double* n, double h, int rule) {
    return tri_quadrature(v1, rule,
                          [&](const double* p) {
                              double pp[3], pm[3];
                              for (int k = 0; k < 3; ++k) {
                                  pp[k] = p[k] + n[k] * h;
                                  pm[k] = p[k] - n[k] * h;
                              }
                              return (orc_potential_opt_3d(pm, v2) - orc_potential_opt_3d(pp, v2)) / (2.0 * h);
                          }) /
           kFourPiE0;
}

extern "C++" {
namespace {
// Rule points on a triangle as the double-quadrature routines map them
// (Potential.cpp:3179-3260, the tNtransform form): the local frame, the 2x3
// homogeneous transform and its Jacobian; nmod = |side1 x side2|.
void full_points(const float* v, int n, const double* px, const double* py, double (*out)[3], double* jac,
                 double* nmod) {
    double V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) V[i][k] = v[3 * i + k];
    double s1[3], s2[3], s3[3], nrm[3], z[3], x[3], y[3];
    vsub(V[1], V[0], s1);
    vsub(V[2], V[1], s2);
    vsub(V[0], V[2], s3);
    vcross(s1, s2, nrm);
    const double nm = vnorm(nrm);
    for (int k = 0; k < 3; ++k) z[k] = nrm[k] / nm;
    const double m1 = vnorm(s1);
    for (int k = 0; k < 3; ++k) x[k] = s1[k] / m1;
    vcross(x, z, y);
    const double l1x = m1, l2x = -vdot(s3, x), l2y = -vdot(s3, y);
    const double t0x = l1x - l2x, t0y = -l2x, t0z = l2x;
    const double t1x = -l2y, t1y = -l2y, t1z = l2y;
    *jac = t0x * t1y - t1x * t0y;
    *nmod = nm;
    for (int i = 0; i < n; ++i) {
        const double qx = t0x * px[i] + t0y * py[i] + t0z * 1.0;
        const double qy = t1x * px[i] + t1y * py[i] + t1z * 1.0;
        for (int k = 0; k < 3; ++k) out[i][k] = V[0][k] + x[k] * qx + y[k] * qy;
    }
}

// the double rules of Mutual_2thOrd_FullNum / MutualD_2thOrd_FullNum (3 points per
// triangle, :3147-3152) and Mutual_4thOrd_FullNum (6 points, :3799-3812): points
// and the weights of each point pair, i-major
const double kFull2P[3][2] = {{0.66666666666666666666666666666667, 0.16666666666666666666666666666667},
                              {0.16666666666666666666666666666667, 0.66666666666666666666666666666667},
                              {0.16666666666666666666666666666667, 0.16666666666666666666666666666667}};
const double kFull2W = 0.11111111111111111111111111111111;
const double kFull4P[6][2] = {{0.816847572980459, 0.091576213509771}, {0.091576213509771, 0.816847572980459},
                              {0.091576213509771, 0.091576213509771}, {0.108103018168070, 0.445948490915965},
                              {0.445948490915965, 0.108103018168070}, {0.445948490915965, 0.445948490915965}};
const double kFull4Waa = 0.012089385932845641681938923684, kFull4Wab = 0.024561195285594988334146524542,
             kFull4Wbb = 0.049899334607075270538656916121;
double full4_weight(int i, int j) {
    const bool a = i < 3, b = j < 3;
    return a && b ? kFull4Waa : (a || b ? kFull4Wab : kFull4Wbb);
}

// sum over the point pairs of w(i, j) K(p_i - q_j), the Jacobians and the normal
// moduli applied as the reference applies them
template <class W, class K>
double full_pairs(const float* v1, const float* v2, int n, const double* px, const double* py, W weight, K kernel) {
    double p1[19][3], p2[19][3], j1, j2, n1, n2;
    full_points(v1, n, px, py, p1, &j1, &n1);
    full_points(v2, n, px, py, p2, &j2, &n2);
    double result = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double dd[3];
            vsub(p1[i], p2[j], dd);
            result = result + weight(i, j) * kernel(dd);
        }
    result = result * j1 * j2;
    result = result / (n1 * n2);
    return result / kFourPiE0;
}
}  // namespace
}  // extern "C++"

// Mutual_2thOrd_FullNum (Potential.cpp:3138-3347) and Mutual_4thOrd_FullNum
// (:3790-4003): the double mean of 1/r by a product rule (order 2: 3 x 3 points,
// order 4: 6 x 6 points).
double orc_mutual_full_3d(const float* v1, const float* v2, int order) {
    double px[6] = {}, py[6] = {};
    const int n = order == 4 ? 6 : 3;
    for (int i = 0; i < n; ++i) {
        px[i] = order == 4 ? kFull4P[i][0] : kFull2P[i][0];
        py[i] = order == 4 ? kFull4P[i][1] : kFull2P[i][1];
    }
    return full_pairs(
        v1, v2, n, px, py, [&](int i, int j) { return order == 4 ? full4_weight(i, j) : kFull2W; },
        [](const double* dd) { return 1 / vnorm(dd); });
}

// MutualDFullNumerical (Potential.cpp:4197-4377, rule >= 1) and
// MutualD_2thOrd_FullNum (:3349-3577, rule 0): the double mean of (r1 - r2).n /
// |r1 - r2|^3, with the product of the rule's weights (or the 3 x 3 table).
double orc_mutual_dfull_3d(const float* v1, const float* v2, const double* n, int rule) {
    double px[19] = {}, py[19] = {}, w[19] = {};
    int m;
    if (rule == 0) {
        m = 3;
        for (int i = 0; i < 3; ++i) {
            px[i] = kFull2P[i][0];
            py[i] = kFull2P[i][1];
        }
    } else {
        const Rule& R = rule_of(rule);
        m = R.n;
        for (int i = 0; i < m; ++i) {
            px[i] = R.px[i];
            py[i] = R.py[i];
            w[i] = R.w[i];
        }
    }
    return full_pairs(
        v1, v2, m, px, py, [&](int i, int j) { return rule == 0 ? kFull2W : w[i] * w[j]; },
        [&](const double* dd) {
            const double rd = vnorm(dd);
            return vdot(dd, n) / (rd * rd * rd);
        });
}

// Local segment frame shared by the 2D routines (Potential.cpp:1866-1899).
static void seg_frame(const double* r, const float* v, double* a, double* xo, double* yo, double* len, double* ex,
                      double* ey) {
    const double sx = (double)v[3] - (double)v[0], sy = (double)v[4] - (double)v[1];
    *len = std::sqrt(sx * sx + sy * sy);
    ex[0] = sx / *len;
    ex[1] = sy / *len;
    ey[0] = -ex[1];
    ey[1] = ex[0];
    const double ox = (double)v[0] + sx / 2.0, oy = (double)v[1] + sy / 2.0;
    *a = *len / 2.0;
    const double px = r[0] - ox, py = r[1] - oy;
    *xo = px * ex[0] + py * ex[1];
    *yo = px * ey[0] + py * ey[1];
}

double orc_potential_opt_2d(const double* r, const float* v) {
    double a, x, y, len, ex[2], ey[2];
    seg_frame(r, v, &a, &x, &y, &len, ex, ey);
    const double xp = x + a, xm = x - a;
    double res;
    if (std::fabs(y) < kTolEps && (std::fabs(xp) < kTolEps || std::fabs(xm) < kTolEps)) {
        res = -a * (std::log(a * a) + kLog4p2);
    } else {
        const double y2 = y * y;
        res = -xp * (std::log(y2 + xp * xp) / 2.0 - 1.0) + xm * (std::log(y2 + xm * xm) / 2.0 - 1.0);
        res += (std::fabs(xp) < kTolEps) ? std::fabs(y) * kPiHalf : y * std::atan2(y, xp);
        res += (std::fabs(xm) < kTolEps) ? -std::fabs(y) * kPiHalf : -y * std::atan2(y, xm);
    }
    return res / len;
}

double orc_enfield_2d(const double* r, const float* v, const double* n) {
    double a, x, y, len, ex[2], ey[2];
    seg_frame(r, v, &a, &x, &y, &len, ex, ey);
    double nx = n[0] * ex[0] + n[1] * ex[1];
    double ny = n[0] * ey[0] + n[1] * ey[1];
    const double nn = std::sqrt(nx * nx + ny * ny);
    if (nn > 0) {
        nx /= nn;
        ny /= nn;
    }
    const double xp = x + a, xm = x - a;
    double res;
    if (std::fabs(y) < kTolEps && (std::fabs(xp) < kTolEps || std::fabs(xm) < kTolEps)) {
        res = -a * (std::log(a * a) + kLog4p2);
    } else {
        const double y2 = y * y;
        const double fx = (std::log(y2 + xp * xp) - std::log(y2 + xm * xm)) / 2.0;
        const double fy = std::atan2(y, xm) - std::atan2(y, xp);
        res = fx * nx + fy * ny;
    }
    return res / len;
}

double orc_pot_estimate(const fcg_job_desc* d, int i, int j) {
    const fcg_node_geom& p1 = d->nodes[i];
    const fcg_node_geom& p2 = d->nodes[j];
    double c1[3], c2[3], dist[3];
    centroid(p1, c1);
    centroid(p2, c2);
    vsub(c1, c2, dist);
    const bool diel1 = (p1.flags & FCG_NODE_DIEL) != 0;
    if (d->dim == 2) {
        dist[2] = 0;
        const double rd = std::sqrt(dist[0] * dist[0] + dist[1] * dist[1]);
        const double rmax = p1.maxside > p2.maxside ? p1.maxside : p2.maxside;
        if (!diel1) {
            if (rd / rmax < 4.0) return orc_potential_opt_2d(c1, p2.v);
            return -std::log(rd);
        }
        if (rd / rmax < 4.0) return orc_enfield_2d(c1, p2.v, p1.dn);
        return (dist[0] * p1.dn[0] + dist[1] * p1.dn[1]) / (rd * rd);
    }
    const double rd = vnorm(dist);
    const double rmax = p1.maxside > p2.maxside ? p1.maxside : p2.maxside;
    const double ratio = rd / rmax;
    if (d->scheme == FCG_SCHEME_GALERKIN) {  // the Galerkin branches (Autorefine.cpp:3570-3674)
        if (!diel1) {
            if (ratio < 1.0) return orc_mutual_half_3d(p1.v, p2.v, 17);
            if (ratio < 2.0) return orc_mutual_half_3d(p1.v, p2.v, 10);
            if (ratio < 4.0) return orc_mutual_full_3d(p1.v, p2.v, 4);
            if (ratio < 10.0) return orc_mutual_full_3d(p1.v, p2.v, 2);
            return 1 / (kFourPiE0 * rd);
        }
        if (ratio < 1.0) return orc_mutual_dhalf_3d(p1.v, p2.v, p1.dn, rd / 20, 17);
        if (ratio < 2.0) return orc_mutual_dfull_3d(p1.v, p2.v, p1.dn, 10);
        if (ratio < 4.0) return orc_mutual_dfull_3d(p1.v, p2.v, p1.dn, 5);
        if (ratio < 10.0) return orc_mutual_dfull_3d(p1.v, p2.v, p1.dn, 0);
        return vdot(dist, p1.dn) / (kFourPiE0 * rd * rd * rd);
    }
    if (!diel1) {
        if (ratio < 1.0) return orc_potential_opt_3d(c1, p2.v) / kFourPiE0;
        if (ratio < 2.0) return orc_potential_num_3d(c1, p2.v, 5) / kFourPiE0;
        if (ratio < 4.0) return orc_potential_num_3d(c1, p2.v, 3) / kFourPiE0;
        if (ratio < 10.0) return orc_potential_num_3d(c1, p2.v, 2) / kFourPiE0;
        return 1 / (kFourPiE0 * rd);
    }
    const double* n1 = p1.dn;
    if (ratio < 1.0) {
        const double h = rd / 20;
        double pp[3], pm[3];
        for (int k = 0; k < 3; ++k) {
            pp[k] = c1[k] + n1[k] * h;
            pm[k] = c1[k] - n1[k] * h;
        }
        const double vp = orc_potential_opt_3d(pp, p2.v);
        const double vm = orc_potential_opt_3d(pm, p2.v);
        return (vm - vp) / (2 * h * kFourPiE0);
    }
    if (ratio < 2.0) return orc_enfield_num_3d(c1, p2.v, n1, 5) / kFourPiE0;
    if (ratio < 4.0) return orc_enfield_num_3d(c1, p2.v, n1, 3) / kFourPiE0;
    if (ratio < 10.0) return orc_enfield_num_3d(c1, p2.v, n1, 2) / kFourPiE0;
    return vdot(dist, n1) / (kFourPiE0 * rd * rd * rd);
}

double orc_self(const fcg_job_desc* d, int leaf) { return orc_node_self(d, d->leaf_node[leaf]); }

double orc_node_self(const fcg_job_desc* d, int node) {
    const fcg_node_geom& g = d->nodes[node];
    if (g.flags & FCG_NODE_DIEL) {
        const double eb = d->group_outperm[g.cond], ea = d->group_inperm[g.cond];
        if (d->dim == 3) return (ea + eb) / (kTwoE0 * (eb - ea) * g.area);
        return kPi * (ea + eb) / ((eb - ea) * g.area);
    }
    double c[3];
    centroid(g, c);
    if (d->dim == 3 && d->scheme == FCG_SCHEME_GALERKIN) return orc_auto_3d(g.v);  // Autorefine.cpp:2399-2401
    if (d->dim == 3) return orc_potential_opt_3d(c, g.v) / kFourPiE0;
    return orc_potential_opt_2d(c, g.v);
}

int orc_fill(const fcg_job_desc* d, double* coef, double* self_coef, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const long M = d->n_nodes;
#pragma omp parallel for
    for (long i = 0; i < M; ++i) {
        for (int64_t k = d->row_ptr[i]; k < d->row_ptr[i + 1]; ++k) coef[k] = orc_pot_estimate(d, (int)i, d->col[k]);
    }
    if (self_coef)
        for (int l = 0; l < d->n_leaves; ++l) self_coef[l] = orc_self(d, l);
    return 0;
}

// the hierarchical product without the 2D fix-up
static void mvp_tree(const fcg_job_desc* d, const double* coef, const double* self_coef, const double* q, double* v,
                     int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int M = d->n_nodes, N = d->n_leaves;
    std::vector<double> chg(M), pot(M);
    // up-pass (MultiplyHierarchical.cpp:72-127): post-order numbering means children precede parents
    for (int i = 0; i < M; ++i) {
        if (d->left[i] < 0) {
            const int k = d->leaf_lo[i];
            chg[i] = q[k];
            pot[i] = self_coef[k] * chg[i];
        } else {
            chg[i] = chg[d->left[i]] + chg[d->right[i]];
            pot[i] = 0.0;
        }
    }
    // gather (MultiplyHierarchical.cpp:141-226), sequential per node, OpenMP over nodes
#pragma omp parallel for
    for (long i = 0; i < M; ++i) {
        double acc = pot[i];
        for (int64_t k = d->row_ptr[i]; k < d->row_ptr[i + 1]; ++k) acc += chg[d->col[k]] * coef[k];
        pot[i] = acc;
    }
    // down-pass (MultiplyHierarchical.cpp:231-283): parents after children in post-order,
    // so walk the numbering backwards (pre-order visits parents first)
    for (int i = M - 1; i >= 0; --i) {
        if (d->left[i] >= 0) {
            pot[d->left[i]] += pot[i];
            pot[d->right[i]] += pot[i];
        }
    }
    for (int k = 0; k < N; ++k) v[k] = pot[d->leaf_node[k]];
}

// 2D zero-total-charge row of one N-vector (MultiplyHierarchical.cpp:305-330)
static void fix_2d(const fcg_job_desc* d, const double* q, double* v) {
    if (d->dim != 2) return;
    const int N = d->n_leaves;
    int first_cond = N;
    for (int g = 0; g < d->n_groups; ++g)
        if (!d->group_is_diel[g]) {
            first_cond = d->group_leaf_start[g];
            break;
        }
    const double last = v[N - 1];
    for (int i = first_cond; i < N - 1; ++i) v[i] -= last;
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += q[i];
    v[N - 1] = 1.0 * s;
}

int orc_mvp(const fcg_job_desc* d, const double* coef, const double* self_coef, const double* q, double* v,
            int nthreads) {
    mvp_tree(d, coef, self_coef, q, v, nthreads);
    fix_2d(d, q, v);
    return 0;
}

// Block preconditioner data (see orc_block_precond).
struct BlockPre {
    int nblocks = 0;
    std::vector<int> starts;
    std::vector<long long> poff;
    std::vector<double> P;
};
// Two-level ("super") preconditioner data (see orc_super_precond).
struct SuperPre {
    int m = 0;
    std::vector<int> nodes, starts;
    std::vector<double> Sinv;
};

// gmresPrecondSFastAll (SolveCapacitance.cpp:2288-2574) on any real system of
// size n: right preconditioning (x0 = P b when a preconditioner is given), MGS
// Arnoldi, Givens rotations, stop when |g_{i+1}| / |b| < tol.  Returns the step count.
using Op = std::function<void(const std::vector<double>&, std::vector<double>&)>;
static int gmres_op(long n, const Op& A, const Op* P, const std::vector<double>& b, std::vector<double>& x,
                    double tol) {
    auto nrm = [&](const std::vector<double>& a) {
        double s = 0;
        for (long i = 0; i < n; ++i) s += a[i] * a[i];
        return std::sqrt(s);
    };
    std::vector<double> x0(n, 0.0), r(n), v(n), w(n), pq(n);
    double normr, normb;
    if (P) {
        (*P)(b, x0);
        A(x0, v);
        for (long k = 0; k < n; ++k) r[k] = b[k] - v[k];
        normr = nrm(r);
        normb = nrm(b);
    } else {
        r = b;
        normr = nrm(r);
        normb = normr;
    }
    const int maxit = (int)(n < 1000 ? n : 1000);
    std::vector<std::vector<double>> Q;
    std::vector<std::vector<double>> H;
    std::vector<double> c(maxit + 1), s(maxit + 1), g(maxit + 2, 0.0);
    Q.emplace_back(n);
    for (long k = 0; k < n; ++k) Q[0][k] = r[k] / normr;
    g[0] = normr;
    if (normr / normb < tol) {
        x = x0;
        return 0;
    }
    int i;
    double rel = normr / normb;
    for (i = 0; i < maxit; ++i) {
        if (P) {
            (*P)(Q[i], pq);
            A(pq, v);
        } else {
            A(Q[i], v);
        }
        w = v;
        H.emplace_back(i + 2, 0.0);
        std::vector<double>& h = H[i];
        for (int j = 0; j <= i; ++j) {
            double sdot = 0;
            for (long k = 0; k < n; ++k) sdot += Q[j][k] * w[k];
            h[j] = sdot;
            for (long k = 0; k < n; ++k) v[k] = v[k] - h[j] * Q[j][k];
        }
        h[i + 1] = nrm(v);
        Q.emplace_back(n);
        for (long k = 0; k < n; ++k) Q[i + 1][k] = v[k] / h[i + 1];
        for (int j = 0; j < i; ++j) {
            const double t1 = h[j], t2 = h[j + 1];
            h[j] = c[j] * t1 - s[j] * t2;
            h[j + 1] = c[j] * t2 + s[j] * t1;
        }
        const double t1 = h[i], t2 = h[i + 1];
        const double len = std::sqrt(t1 * t1 + t2 * t2);
        c[i] = t1 / len;
        s[i] = -t2 / len;
        h[i] = c[i] * t1 - s[i] * t2;
        h[i + 1] = 0;
        const double gi = g[i];
        g[i] = c[i] * gi;
        g[i + 1] = s[i] * gi;
        rel = std::fabs(g[i + 1]) / normb;
        if (rel < tol) break;
    }
    int iters = i;
    if (i >= maxit) iters = i - 1;
    std::vector<double> z(g.begin(), g.begin() + iters + 1);
    for (int ii = iters; ii >= 0; --ii) {
        z[ii] = z[ii] / H[ii][ii];
        for (int j = ii - 1; j >= 0; --j) z[j] = z[j] - H[ii][j] * z[ii];
    }
    std::vector<double> y(n, 0.0);
    for (long k = 0; k < n; ++k) {
        double t = 0;
        for (int j = 0; j <= iters; ++j) t += Q[j][k] * z[j];
        y[k] = t;
    }
    if (P) {
        (*P)(y, x);
        for (long k = 0; k < n; ++k) x[k] = x0[k] + x[k];
    } else {
        x = y;
    }
    return iters + 1;
}

// the real system: no preconditioner, Jacobi, or the block / two-level preconditioner
static int gmres(const fcg_job_desc* d, const double* coef, const double* self_coef, const std::vector<double>& b,
                 std::vector<double>& x, double tol, int precond, int nthreads, const BlockPre* bp = nullptr,
                 const SuperPre* sp = nullptr) {
    const long n = d->n_leaves;
    const Op A = [&](const std::vector<double>& in, std::vector<double>& out) {
        orc_mvp(d, coef, self_coef, in.data(), out.data(), nthreads);
    };
    const Op P = [&](const std::vector<double>& in, std::vector<double>& out) {
        if (sp) {
            orc_super_apply(d, sp->m, sp->nodes.data(), sp->starts.data(), sp->Sinv.data(), in.data(), out.data());
            return;
        }
        if (bp) {
            orc_block_apply(bp->nblocks, bp->starts.data(), bp->poff.data(), bp->P.data(), in.data(), out.data());
            return;
        }
        for (long i = 0; i < n; ++i) out[i] = in[i] / self_coef[i];
    };
    return gmres_op(n, A, precond ? &P : nullptr, b, x, tol);
}

int orc_solve(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int precond,
              double* cap, int* rows, int* cols, int* iters, int nthreads) {
    return orc_solve_rows(d, coef, self_coef, tol, precond, -1, cap, rows, cols, iters, nthreads);
}

static int solve_rows(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int precond,
                      int max_rows, double* cap, int* rows, int* cols, int* iters, int nthreads,
                      const BlockPre* bp, const SuperPre* sp = nullptr);

int orc_solve_super(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int dim,
                    double* cap, int* rows, int* cols, int* iters, int nthreads) {
    SuperPre sp;
    sp.nodes.assign((size_t)std::max(dim, 1), 0);
    sp.starts.assign((size_t)std::max(dim, 1) + 1, 0);
    sp.Sinv.assign((size_t)dim * (size_t)dim, 0.0);
    if (orc_super_precond(d, dim, sp.nodes.data(), sp.starts.data(), &sp.m, sp.Sinv.data()) != 0)
        return solve_rows(d, coef, self_coef, tol, 0, -1, cap, rows, cols, iters, nthreads, nullptr);  // unavailable
    return solve_rows(d, coef, self_coef, tol, 3, -1, cap, rows, cols, iters, nthreads, nullptr, &sp);
}

int orc_solve_rows(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int precond,
                   int max_rows, double* cap, int* rows, int* cols, int* iters, int nthreads) {
    return solve_rows(d, coef, self_coef, tol, precond, max_rows, cap, rows, cols, iters, nthreads, nullptr);
}

int orc_solve_block(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int bsize,
                    double* cap, int* rows, int* cols, int* iters, int nthreads) {
    BlockPre bp;
    const long n = d->n_leaves;
    bp.starts.assign((size_t)n + 1, 0);
    bp.poff.assign((size_t)n + 1, 0);
    bp.P.assign((size_t)n * (size_t)(bsize < 2 ? 2 : bsize), 0.0);
    const int rc = orc_block_precond(d, bsize, bp.starts.data(), &bp.nblocks, bp.poff.data(), bp.P.data());
    if (rc) return rc;
    return solve_rows(d, coef, self_coef, tol, 2, -1, cap, rows, cols, iters, nthreads, &bp);
}

static int solve_rows(const fcg_job_desc* d, const double* coef, const double* self_coef, double tol, int precond,
                      int max_rows, double* cap, int* rows, int* cols, int* iters, int nthreads,
                      const BlockPre* bp, const SuperPre* sp) {
    const long n = d->n_leaves;
    int ncond = 0;
    for (int g = 0; g < d->n_groups; ++g)
        if (!d->group_is_diel[g]) ++ncond;
    const int nrow = (d->dim == 2 && ncond > 1) ? ncond - 1 : ncond;
    *rows = nrow;
    *cols = ncond;
    std::vector<double> b(n, 0.0), x(n, 0.0);
    int row = 0;
    for (int g = 0; g < d->n_groups; ++g) {
        if (d->group_is_diel[g]) continue;
        const int lo = d->group_leaf_start[g], hi = d->group_leaf_start[g + 1];
        if (d->dim == 2 && hi == n) continue;  // last conductor row is the balance row (SolveCapacitance.cpp:1233-1240)
        if (max_rows >= 0 && row >= max_rows) break;
        for (int k = lo; k < hi; ++k) b[k] = 1.0;
        const int it = gmres(d, coef, self_coef, b, x, tol, precond, nthreads, bp, sp);
        if (iters) iters[row] = it;
        int col = 0;
        for (int g2 = 0; g2 < d->n_groups; ++g2) {
            if (d->group_is_diel[g2]) continue;
            double sum = 0.0;
            for (int k = d->group_leaf_start[g2]; k < d->group_leaf_start[g2 + 1]; ++k)
                sum += d->leaf_outperm[k] * x[k];
            cap[(size_t)row * ncond + col] = sum;
            ++col;
        }
        for (int k = lo; k < hi; ++k) b[k] = 0.0;
        ++row;
    }
    return 0;
}

int orc_compressed_dense(const fcg_job_desc* d, const double* coef, const double* self_coef, double* P) {
    const long N = d->n_leaves;
    std::memset(P, 0, sizeof(double) * N * N);
    for (long i = 0; i < N; ++i) P[i * N + i] = self_coef[i];
    for (int a = 0; a < d->n_nodes; ++a)
        for (int64_t k = d->row_ptr[a]; k < d->row_ptr[a + 1]; ++k) {
            const int b = d->col[k];
            for (int i = d->leaf_lo[a]; i < d->leaf_hi[a]; ++i)
                for (int j = d->leaf_lo[b]; j < d->leaf_hi[b]; ++j) P[(long)i * N + j] += coef[k];
        }
    return 0;
}

int orc_leaf_dense(const fcg_job_desc* d, double* P, int nthreads) {
    const long N = d->n_leaves;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < N; ++i)
        for (long j = 0; j < N; ++j)
            P[i * N + j] = (i == j) ? orc_self(d, (int)i) : orc_pot_estimate(d, d->leaf_node[i], d->leaf_node[j]);
    return 0;
}

long long orc_link_coverage(const fcg_job_desc* d) {
    const long N = d->n_leaves;
    std::vector<unsigned char> cnt((size_t)N * N, 0);
    for (int a = 0; a < d->n_nodes; ++a)
        for (int64_t k = d->row_ptr[a]; k < d->row_ptr[a + 1]; ++k) {
            const int b = d->col[k];
            for (int i = d->leaf_lo[a]; i < d->leaf_hi[a]; ++i)
                for (int j = d->leaf_lo[b]; j < d->leaf_hi[b]; ++j) {
                    unsigned char& c = cnt[(size_t)i * N + j];
                    if (c < 255) ++c;
                }
        }
    long long bad = 0;
    for (long i = 0; i < N; ++i)
        for (long j = 0; j < N; ++j) bad += (cnt[(size_t)i * N + j] != (i == j ? 0 : 1));
    return bad;
}

}  // extern "C"

// ---------------------------------------------------------------- block preconditioner
// RecurseComputePrecond (Solver/SolveCapacitance.cpp:1094-1104, 1876-1993): each
// conductor's tree is walked depth first, left before right, conductors in leaf
// order; the first subtree met with at most bsize leaves forms one block of
// consecutive leaves.  A leaf met before any block is open (its parent has more
// than bsize leaves) forms a block of one: the reference opens a block there but
// never computes it.
extern "C" int orc_block_partition(const fcg_job_desc* d, int bsize, int* starts, int* nblocks) {
    if (bsize < 2) bsize = 2;
    if (bsize > 128) bsize = 128;
    std::vector<int> roots;
    for (int i = 0; i < d->n_nodes; ++i)
        if (d->parent[i] < 0) roots.push_back(i);
    std::sort(roots.begin(), roots.end(), [&](int a, int b) { return d->leaf_lo[a] < d->leaf_lo[b]; });
    int nb = 0;
    std::vector<int> stack;
    for (int r : roots) {
        stack.push_back(r);
        while (!stack.empty()) {
            const int a = stack.back();
            stack.pop_back();
            const int cnt = d->leaf_hi[a] - d->leaf_lo[a];
            if (cnt <= bsize || d->left[a] < 0) {
                starts[nb++] = d->leaf_lo[a];
                continue;
            }
            stack.push_back(d->right[a]);  // left is visited first
            stack.push_back(d->left[a]);
        }
    }
    starts[nb] = d->n_leaves;
    *nblocks = nb;
    return 0;
}

// InvertMatrix (Solver/SolveCapacitance.cpp:2116-2258): right-looking LU with
// partial pivoting (first largest |entry| of the column), then the inverse column
// by column from forward and backward substitution of the unit vectors.
static void invert_block(std::vector<double>& A, int n, double* inv) {
    std::vector<int> piv((size_t)n);
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = std::fabs(A[(size_t)j * n + j]);
        for (int i = j + 1; i < n; ++i)
            if (std::fabs(A[(size_t)i * n + j]) > best) {
                best = std::fabs(A[(size_t)i * n + j]);
                p = i;
            }
        piv[(size_t)j] = p;
        if (p != j)
            for (int c = 0; c < n; ++c) std::swap(A[(size_t)j * n + c], A[(size_t)p * n + c]);
        if (j < n - 1) {
            const double rcp = 1.0 / A[(size_t)j * n + j];
            for (int i = j + 1; i < n; ++i) A[(size_t)i * n + j] *= rcp;
            for (int i = j + 1; i < n; ++i)
                for (int c = j + 1; c < n; ++c) A[(size_t)i * n + c] -= A[(size_t)i * n + j] * A[(size_t)j * n + c];
        }
    }
    std::vector<double> y((size_t)n);
    for (int k = 0; k < n; ++k) {
        for (int i = 0; i < n; ++i) y[(size_t)i] = (i == k) ? 1.0 : 0.0;
        for (int i = 0; i < n; ++i) {  // L (unit diagonal), pivots applied in order
            const int ip = piv[(size_t)i];
            double sum = y[(size_t)ip];
            y[(size_t)ip] = y[(size_t)i];
            for (int j = 0; j < i; ++j) sum -= A[(size_t)i * n + j] * y[(size_t)j];
            y[(size_t)i] = sum;
        }
        for (int i = n - 1; i >= 0; --i) {  // U
            double sum = y[(size_t)i];
            for (int j = i + 1; j < n; ++j) sum -= A[(size_t)i * n + j] * y[(size_t)j];
            y[(size_t)i] = sum / A[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + k] = y[(size_t)i];
    }
}

// ComputeBlockPrecond (Solver/SolveCapacitance.cpp:2068-2113): the block's dense
// collocation matrix (self terms on the diagonal, PotEstimateOpt between leaves
// elsewhere; the reference's two-sided call returns the two one-sided values),
// inverted.
extern "C" int orc_block_precond(const fcg_job_desc* d, int bsize, int* starts, int* nblocks, long long* poff,
                                 double* P) {
    orc_block_partition(d, bsize, starts, nblocks);
    const int nb = *nblocks;
    poff[0] = 0;
    for (int b = 0; b < nb; ++b) {
        const long long n = starts[b + 1] - starts[b];
        poff[b + 1] = poff[b] + n * n;
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int b = 0; b < nb; ++b) {
        const int b0 = starts[b], n = starts[b + 1] - starts[b];
        std::vector<double> A((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                A[(size_t)i * n + j] = (i == j) ? orc_self(d, b0 + i)
                                                : orc_pot_estimate(d, d->leaf_node[b0 + i], d->leaf_node[b0 + j]);
        invert_block(A, n, P + poff[b]);
    }
    return 0;
}

// ComputePrecondVectFast block branch (Solver/SolveCapacitance.cpp:3527-3544).
extern "C" void orc_block_apply(int nblocks, const int* starts, const long long* poff, const double* P,
                                const double* q, double* out) {
    for (int b = 0; b < nblocks; ++b) {
        const int b0 = starts[b], n = starts[b + 1] - starts[b];
        const double* Pb = P + poff[b];
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += Pb[(size_t)i * n + j] * q[b0 + j];
            out[b0 + i] = s;
        }
    }
}

// ---------------------------------------------------------------- two-level preconditioner
// Super elements (MakePreconditioner, Solver/SolveCapacitance.cpp:1067-1104, and the
// super branch of RecurseComputePrecond, :1876-1993): with R tree roots (conductors
// and dielectric interfaces) and dimension dim, every root's tree is walked to
// level L = log2(dim / R) (integer division; the root is level 0): a non-leaf at
// level L is one element standing for all its leaves, a leaf at level <= L is
// one.  Elements follow leaf order, so each covers consecutive leaves.  With
// R >= dim the preconditioner is unavailable (returns 1; the reference then drops
// it and solves unpreconditioned).
extern "C" int orc_super_partition(const fcg_job_desc* d, int dim, int* nodes, int* starts, int* m) {
    std::vector<int> roots;
    for (int i = 0; i < d->n_nodes; ++i)
        if (d->parent[i] < 0) roots.push_back(i);
    std::sort(roots.begin(), roots.end(), [&](int a, int b) { return d->leaf_lo[a] < d->leaf_lo[b]; });
    *m = 0;
    const unsigned R = (unsigned)roots.size();
    if (dim <= 0 || R >= (unsigned)dim) return 1;
    const int L = (int)(unsigned)(std::log((double)((unsigned)dim / R)) / std::log(2.0));
    int cnt = 0;
    std::vector<std::pair<int, int>> stack;
    for (int r : roots) {
        stack.push_back({r, 0});
        while (!stack.empty()) {
            const auto [a, lev] = stack.back();
            stack.pop_back();
            if (d->left[a] < 0 || lev == L) {
                nodes[cnt] = a;
                starts[cnt] = d->leaf_lo[a];
                ++cnt;
                continue;
            }
            stack.push_back({d->right[a], lev + 1});
            stack.push_back({d->left[a], lev + 1});
        }
    }
    starts[cnt] = d->n_leaves;
    *m = cnt;
    return 0;
}

// ComputeSuperPrecond (Solver/SolveCapacitance.cpp:2007-2066): the elements' dense
// collocation matrix (self potential of each element node on the diagonal,
// PotEstimateOpt between element nodes elsewhere), inverted (InvertMatrix).
// Sinv: m x m row-major (capacity dim * dim).
extern "C" int orc_super_precond(const fcg_job_desc* d, int dim, int* nodes, int* starts, int* m, double* Sinv) {
    if (orc_super_partition(d, dim, nodes, starts, m) != 0) return 1;
    const int n = *m;
    std::vector<double> A((size_t)n * n);
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            A[(size_t)i * n + j] = (i == j) ? orc_node_self(d, nodes[i]) : orc_pot_estimate(d, nodes[i], nodes[j]);
    invert_block(A, n, Sinv);
    return 0;
}

// ComputePrecondVectFast super branch (Solver/SolveCapacitance.cpp:3545-3592):
// collect each element's area-weighted input, multiply by the off-diagonal part
// of the inverse, distribute back with the leaf's area weight, and add the
// diagonal entry times the leaf's own weighted input.
extern "C" void orc_super_apply(const fcg_job_desc* d, int m, const int* nodes, const int* starts,
                                const double* Sinv, const double* q, double* out) {
    std::vector<double> pot((size_t)m, 0.0), chg((size_t)m, 0.0);
    for (int s = 0; s < m; ++s) {
        const double dimS = d->nodes[nodes[s]].area;
        for (int i = starts[s]; i < starts[s + 1]; ++i) pot[(size_t)s] += q[i] * d->nodes[d->leaf_node[i]].area / dimS;
    }
    for (int s = 0; s < m; ++s)
        for (int t = 0; t < m; ++t)
            if (s != t) chg[(size_t)s] += Sinv[(size_t)s * m + t] * pot[(size_t)t];
    for (int s = 0; s < m; ++s) {
        const double dimS = d->nodes[nodes[s]].area;
        for (int i = starts[s]; i < starts[s + 1]; ++i) {
            const double w = d->nodes[d->leaf_node[i]].area / dimS;
            out[i] = chg[(size_t)s] * w + Sinv[(size_t)s * m + s] * q[i] * w;
        }
    }
}


// ------------------------------------------------------------------ complex permittivities

// SelfPotential of a leaf of a complex job (Solver/Autorefine.cpp:2358-2393, 2D
// :2435-2471): a dielectric panel's self term is the complex number
// (ea + eb) / (2 e0 (eb - ea) A) in 3D, pi (ea + eb) / ((eb - ea) L) in 2D
// (complex division by the conjugate over the squared modulus); a conductor
// panel's is real.
extern "C" void orc_self_cplx(const fcg_job_desc* d, int leaf, double* re, double* im) {
    const int node = d->leaf_node[leaf];
    const fcg_node_geom& g = d->nodes[node];
    *im = 0.0;
    if (!(g.flags & FCG_NODE_DIEL) || !d->group_outperm_im) {
        *re = orc_node_self(d, node);
        return;
    }
    const double sr = d->group_inperm[g.cond] + d->group_outperm[g.cond];
    const double si = d->group_inperm_im[g.cond] + d->group_outperm_im[g.cond];
    const double dr = d->group_outperm[g.cond] - d->group_inperm[g.cond];
    const double di = d->group_outperm_im[g.cond] - d->group_inperm_im[g.cond];
    const double scale = d->dim == 3 ? 1.0 / (kTwoE0 * g.area) : kPi / g.area;
    const double mod2 = dr * dr + di * di;
    *re = scale * (sr * dr + si * di) / mod2;
    *im = scale * (si * dr - sr * di) / mod2;
}

// MultiplyMatByVec_fast, complex branch (Solver/MultiplyHierarchical.cpp:332-455):
// q and v are [Re(N) | Im(N)]; v_re = R q_re - D q_im, v_im = R q_im + D q_re with
// D = diag(self_im), then each half's 2D fix-up.
extern "C" int orc_mvp_cplx(const fcg_job_desc* d, const double* coef, const double* self_re, const double* self_im,
                            const double* q, double* v, int nthreads) {
    const int N = d->n_leaves;
    const double *qr = q, *qi = q + N;
    double *vr = v, *vi = v + N;
    mvp_tree(d, coef, self_re, qr, vr, nthreads);
    for (int i = 0; i < N; ++i) vr[i] -= self_im[i] * qi[i];
    fix_2d(d, qr, vr);
    mvp_tree(d, coef, self_re, qi, vi, nthreads);
    for (int i = 0; i < N; ++i) vi[i] += self_im[i] * qr[i];
    fix_2d(d, qi, vi);
    return 0;
}

// CSolveCap::Solve for a complex job (Solver/SolveCapacitance.cpp:1138-1366):
// GMRES on the real 2N system with b = [1 on the conductor | 0], Jacobi divides
// both halves by the real self term (ComputePrecondVectFast, :3498-3526), and
// C = sum of the complex outer permittivity times the complex charge.  cap
// receives the real matrix (rows x n_cond) followed by the imaginary one.
extern "C" int orc_solve_cplx(const fcg_job_desc* d, const double* coef, const double* self_re,
                              const double* self_im, double tol, int precond, double* cap, int* rows, int* cols,
                              int* iters, int nthreads) {
    const long N = d->n_leaves, n = 2 * N;
    int ncond = 0;
    for (int g = 0; g < d->n_groups; ++g)
        if (!d->group_is_diel[g]) ++ncond;
    const int nrow = (d->dim == 2 && ncond > 1) ? ncond - 1 : ncond;
    *rows = nrow;
    *cols = ncond;
    const Op A = [&](const std::vector<double>& in, std::vector<double>& out) {
        orc_mvp_cplx(d, coef, self_re, self_im, in.data(), out.data(), nthreads);
    };
    const Op P = [&](const std::vector<double>& in, std::vector<double>& out) {
        for (long i = 0; i < n; ++i) out[i] = in[i] / self_re[i % N];
    };
    std::vector<double> b(n, 0.0), x(n, 0.0);
    int row = 0;
    for (int g = 0; g < d->n_groups; ++g) {
        if (d->group_is_diel[g]) continue;
        const int lo = d->group_leaf_start[g], hi = d->group_leaf_start[g + 1];
        if (d->dim == 2 && hi == N) continue;
        for (int k = lo; k < hi; ++k) b[k] = 1.0;
        const int it = gmres_op(n, A, precond ? &P : nullptr, b, x, tol);
        if (iters) iters[row] = it;
        int col = 0;
        for (int g2 = 0; g2 < d->n_groups; ++g2) {
            if (d->group_is_diel[g2]) continue;
            double sr = 0.0, si = 0.0;
            for (int k = d->group_leaf_start[g2]; k < d->group_leaf_start[g2 + 1]; ++k) {
                const double wr = d->leaf_outperm[k], wi = d->leaf_outperm_im[k];
                sr += wr * x[k] - wi * x[k + N];
                si += wi * x[k] + wr * x[k + N];
            }
            cap[(size_t)row * ncond + col] = sr;
            cap[(size_t)nrow * ncond + (size_t)row * ncond + col] = si;
            ++col;
        }
        for (int k = lo; k < hi; ++k) b[k] = 0.0;
        ++row;
    }
    return 0;
}
