// Host mesh builder: FastCap parser, super hierarchy, adaptive discretization and
// hierarchical link discovery, flattened into the post-order / destination-CSR
// arrays the device library consumes.
//
// Follows FasterCap 6.0.7 (file:line citations per function).  Geometry is stored
// as float and every arithmetic expression is evaluated in double, rounding only
// when a value is stored into a float field (the reference's C3DVector_float
// storage; the exact operator semantics of its missing Geometry package are not
// available, so this promotion rule is the one chosen here).
#include "fastercap_host.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../common/potential.h"

namespace {

constexpr double kPanelEps = 1E-12;                          // AUTOPANEL_EPS, AutoElement.h:58
constexpr double kTwoDivFourthRoot27 = 0.87738267530166164054614593453133;  // AutoPanel.cpp:36
constexpr double kTwoDivFourthRoot3 = 1.5196713713031850946623755013091;    // AutoPanel.cpp:37
constexpr double kM05PlusSqr2Div4 = -0.14644660940672623779957781894758;
constexpr double k05PlusSqr2Div4 = 0.85355339059327376220042218105242;
constexpr double kSqr2 = 1.4142135623730950488016887242097;
constexpr double kQuadSideTol = 0.2;   // AUTOREFINE_QUAD_SIDE_TOL, Autorefine.cpp:72
constexpr double kQuadAngleTol = 0.1;  // AUTOREFINE_QUAD_ANGLE_TOL, Autorefine.cpp:74
constexpr double kRelDielTol = 0.001;  // AUTOREFINE_REL_DIEL_TOL, Autorefine.cpp:78
constexpr double kMax2DDim = 0.5;      // AUTOREFINE_MAX_2D_DIM, Autorefine.cpp:104
constexpr int kMaxParseLevel = 128;    // AUTOREFINE_MAX_PARSE_LEVEL

constexpr uint8_t kLeaf = FCG_NODE_LEAF;
constexpr uint8_t kDiel = FCG_NODE_DIEL;
constexpr uint8_t kNormalDir = FCG_NODE_OUTPERM_NORMAL_DIR;
constexpr uint8_t kElementLevel = FCG_NODE_OUTPERM_ELEMENT_LEVEL;
constexpr uint8_t kSuper = FCG_NODE_SUPER;

thread_local std::string g_err;

struct V3 {
    double x = 0, y = 0, z = 0;
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double mod(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) {
    const double m = mod(a);
    return m > 0 ? a / m : a;
}

struct F3 {
    float x = 0, y = 0, z = 0;
    V3 d() const { return {x, y, z}; }
    void set(V3 a) {
        x = (float)a.x;
        y = (float)a.y;
        z = (float)a.z;
    }
};

// One tree node (CAutoElement + CAutoPanel / CAutoSegment fields, AutoElement.h:398-406,
// AutoPanel.h:196-199, AutoSegment.h).
struct Elem {
    F3 v[3];
    F3 c;
    F3 n;
    double dim = 0;      // m_dDimension: area (3D) or length (2D)
    double maxside = 0;  // m_dMaxSideLen (3D); 2D uses dim
    uint8_t type = kLeaf;
    uint8_t maxSideIdx = 0;
    uint8_t dielIndex = 0;
    Elem* left = nullptr;
    Elem* right = nullptr;
    uint64_t linkStart = 0, linkEnd = 0;
    int64_t numChildren = 0;
    int32_t leafIndex = -1;
    int32_t nodeId = -1;
    int32_t group = -1;
    bool isLeaf() const { return (type & kLeaf) != 0; }
};

struct Group {
    std::string name;
    bool isDiel = false;
    double outperm[2] = {1.0, 0.0};
    double inperm[2] = {0.0, 0.0};
    V3 dielRef;
    std::vector<Elem*> panels;
    std::vector<std::pair<double, double>> surfOutperm;  // m_dSurfOutperm (AutoConductor.h:484)
    Elem* top = nullptr;
    int64_t inputPanels = 0;
    int64_t leafNum = 0;
};

struct BBox {
    V3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    void add(V3 p) {
        for (int i = 0; i < 3; ++i) {
            if (p[i] < lo[i]) lo[i] = p[i];
            if (p[i] > hi[i]) hi[i] = p[i];
        }
    }
    // C3DBBox::MaxSide (Geometry package, not in the reference tree): index of the
    // largest extent; ties resolved towards the lower axis index (x, then y, then z).
    int maxSide() const {
        int m = 0;
        double e[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
        if (e[1] > e[m]) m = 1;
        if (e[2] > e[m]) m = 2;
        return m;
    }
};

struct Builder {
    bool hasCplx = false;  // m_ucHasCmplxPerm: some permittivity read so far is complex
    fch_options opt{};
    int dim = 3;
    std::deque<Elem> arena;
    std::deque<std::unique_ptr<Group>> groups;  // deque order: dielectrics (front), conductors
    double maxSide = 0, maxArea = 0, maxLength = 0, totalArea = 0;
    double meshEps = 0.1, linkEps = 0.1, curv = 3.0;
    double maxMeshEps = 0;
    int level = 0, maxLevel = 0;
    bool computeLinks = false;
    int64_t linksNum = 0, panelNum = 0, nodeNum = 0;
    int64_t inputPanelNum = 0, rawCond = 0, rawDiel = 0;
    int64_t condNum = 0, dielNum = 0, groupDielNum = 0;
    int parseLevel = -1;
    int groupNum[kMaxParseLevel + 2];
    BBox global2D;
    std::string baseDir;
    std::vector<int32_t> colOut;  // record-pass link targets, indexed by link slot
    // flattened output
    std::vector<Elem*> nodes;     // post-order
    int64_t baseLinks = 0, basePanels = 0, countPanels = 0;

    Elem* newElem() {
        arena.emplace_back();
        return &arena.back();
    }

    // ---------------------------------------------------------------- 3D geometry
    // CAutoPanel::MaxSide (AutoPanel.cpp:91-131)
    static int maxSide3(Elem& e, double l0, double l1, double l2) {
        int minside;
        if (l0 >= l1 && l0 >= l2) {
            e.maxSideIdx = 0;
            e.maxside = l0;
            minside = (l1 >= l2) ? 2 : 1;
        } else if (l1 >= l0 && l1 >= l2) {
            e.maxSideIdx = 1;
            e.maxside = l1;
            minside = (l0 >= l2) ? 2 : 0;
        } else {
            e.maxSideIdx = 2;
            e.maxside = l2;
            minside = (l0 >= l1) ? 1 : 0;
        }
        return minside;
    }

    // CAutoPanel::CalculateNormal, Newell's method accumulated into float storage
    // (AutoPanel.cpp:135-185)
    static double newellNormal(const Elem& e, F3& normal) {
        normal = F3{};
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1 > 2) ? 0 : i + 1;
            const V3 a = e.v[i].d(), b = e.v[j].d();
            normal.x = (float)((double)normal.x + (a.y - b.y) * (a.z + b.z));
            normal.y = (float)((double)normal.y + (a.z - b.z) * (a.x + b.x));
            normal.z = (float)((double)normal.z + (a.x - b.x) * (a.y + b.y));
        }
        const double nm = mod(normal.d());
        if (nm > kPanelEps) normal.set(normal.d() / nm);
        return nm;
    }

    // CAutoPanel::CalcPanelGeomPar (AutoPanel.cpp:51-88)
    static double calcPanelGeom(Elem& e) {
        V3 side[3] = {e.v[1].d() - e.v[0].d(), e.v[2].d() - e.v[1].d(), e.v[0].d() - e.v[2].d()};
        double sl[3] = {mod(side[0]), mod(side[1]), mod(side[2])};
        const double nm = newellNormal(e, e.n);
        e.dim = nm / 2.0;
        e.c.set((e.v[0].d() + e.v[1].d() + e.v[2].d()) / 3.0);
        const int b = maxSide3(e, sl[0], sl[1], sl[2]);
        int a, c;
        if (b == 0) {
            a = 2;
            c = 1;
        } else if (b == 1) {
            a = 0;
            c = 2;
        } else {
            a = 1;
            c = 0;
        }
        return (sl[a] * sl[a] + sl[c] * sl[c] - sl[b] * sl[b]) / (2.0 * sl[a] * sl[c]);
    }

    // GetGeoNormal (AutoPanel.h:95-107, AutoSegment.h:92-105)
    V3 geoNormal(const Elem& e) const {
        if (e.type & kSuper) {
            F3 u;
            if (dim == 3) {
                newellNormal(e, u);
            } else {
                u.x = (float)((double)e.v[0].y - (double)e.v[1].y);
                u.y = (float)((double)e.v[1].x - (double)e.v[0].x);
                u.z = 0;
                const double nm = mod(u.d());
                if (nm > kPanelEps) u.set(u.d() / nm);
            }
            return u.d();
        }
        return e.n.d();
    }

    // GetDielNormal (AutoPanel.h:109-134, AutoSegment.h:107-133)
    static V3 dielNormal(const Elem& e) {
        V3 u;
        if (e.type & kSuper) {
            const double m = mod(e.n.d());
            u = (m < kPanelEps) ? V3{0, 0, 0} : e.n.d() / m;
        } else {
            u = e.n.d();
        }
        if ((e.type & kNormalDir) != kNormalDir) u = -u;
        return u;
    }

    // CAutoPanel::MakeSuperPanel (AutoPanel.cpp:209-380)
    void makeSuperPanel(Elem& s, Elem* L, Elem* R) {
        if ((R->type & kDiel) && (L->type & kDiel))
            s.type |= (kSuper | kDiel | kNormalDir);
        else
            s.type |= kSuper;
        s.type &= ~kLeaf;
        s.left = L;
        s.right = R;
        s.dim = L->dim + R->dim;
        s.c.set((L->c.d() * L->dim + R->c.d() * R->dim) / s.dim);
        const V3 lN = (L->type & kNormalDir) ? L->n.d() : -L->n.d();
        const V3 rN = (R->type & kNormalDir) ? R->n.d() : -R->n.d();
        s.n.set((lN * L->dim + rN * R->dim) / s.dim);
        const double mn = mod(s.n.d());
        V3 normal;
        if (mn < kPanelEps) {
            if (std::fabs(lN.x) > kPanelEps) {
                normal = {(-lN.y - lN.z) / lN.x, 1.0, 1.0};
            } else if (std::fabs(lN.y) > kPanelEps) {
                normal = {1.0, (-lN.x - lN.z) / lN.y, 1.0};
            } else if (std::fabs(lN.z) > kPanelEps) {
                normal = {1.0, 1.0, (-lN.x - lN.y) / lN.z};
            } else {
                normal = geoNormal(*L) * L->dim + geoNormal(*R) * R->dim;
                normal = normalized(normal);
                if (mod(normal) < kPanelEps) normal = {0.0, 0.0, 1.0};
            }
        } else {
            normal = s.n.d() / mn;
        }
        V3 cmToCs = (L->dim < R->dim) ? L->c.d() - s.c.d() : R->c.d() - s.c.d();
        if (mod(cross(normal, cmToCs)) < kPanelEps) {
            cmToCs.x = normal.x * 0.5 + normal.y * kM05PlusSqr2Div4 + normal.z * k05PlusSqr2Div4;
            cmToCs.y = normal.x * 0.5 + normal.y * k05PlusSqr2Div4 + normal.z * kM05PlusSqr2Div4;
            cmToCs.z = normal.x * (-kSqr2) + normal.y * 0.5 + normal.z * 0.5;
        }
        const V3 xv = normalized(cross(cmToCs, normal));
        const V3 yv = normalized(cross(normal, xv));
        const double Rc = kTwoDivFourthRoot27 * std::sqrt(s.dim);
        const V3 C = s.c.d();
        s.v[0].set(C + yv * Rc);
        const double a = kTwoDivFourthRoot3 * std::sqrt(s.dim);
        s.v[1].set(C - yv * (Rc / 2.0) - xv * (a / 2.0));
        s.v[2].set(C - yv * (Rc / 2.0) + xv * (a / 2.0));
        s.maxSideIdx = 0;
        s.maxside = a;
    }

    // CAutoPanel::Subdivide (AutoPanel.cpp:382-523)
    void subdivide3(Elem& p) {
        if (!p.isLeaf()) return;
        Elem* L = newElem();
        Elem* R = newElem();
        p.left = L;
        p.right = R;
        const V3 v0 = p.v[0].d(), v1 = p.v[1].d(), v2 = p.v[2].d();
        double ls[3], rs[3];
        const double half = p.maxside / 2.0;
        if (p.maxSideIdx == 0) {
            const V3 mid = (v1 + v0) / 2.0;
            const double nsl = mod(v2 - mid);
            L->v[0] = p.v[0];
            L->v[1].set(mid);
            L->v[2] = p.v[2];
            ls[0] = half;
            ls[1] = nsl;
            ls[2] = mod(v0 - v2);
            R->v[0].set(mid);
            R->v[1] = p.v[1];
            R->v[2] = p.v[2];
            rs[0] = half;
            rs[1] = mod(v2 - v1);
            rs[2] = nsl;
        } else if (p.maxSideIdx == 1) {
            const V3 mid = (v2 + v1) / 2.0;
            const double nsl = mod(v0 - mid);
            L->v[0] = p.v[1];
            L->v[1].set(mid);
            L->v[2] = p.v[0];
            ls[0] = half;
            ls[1] = nsl;
            ls[2] = mod(v1 - v0);
            R->v[0].set(mid);
            R->v[1] = p.v[2];
            R->v[2] = p.v[0];
            rs[0] = half;
            rs[1] = mod(v0 - v2);
            rs[2] = nsl;
        } else {
            const V3 mid = (v0 + v2) / 2.0;
            const double nsl = mod(v1 - mid);
            L->v[0] = p.v[2];
            L->v[1].set(mid);
            L->v[2] = p.v[1];
            ls[0] = half;
            ls[1] = nsl;
            ls[2] = mod(v2 - v1);
            R->v[0].set(mid);
            R->v[1] = p.v[0];
            R->v[2] = p.v[1];
            rs[0] = half;
            rs[1] = mod(v1 - v0);
            rs[2] = nsl;
        }
        L->type = p.type;
        R->type = p.type;
        L->dielIndex = R->dielIndex = p.dielIndex;
        L->group = R->group = p.group;
        p.type &= ~kLeaf;
        L->dim = R->dim = p.dim / 2.0;
        L->n = R->n = p.n;
        L->c.set((L->v[0].d() + L->v[1].d() + L->v[2].d()) / 3.0);
        R->c.set((R->v[0].d() + R->v[1].d() + R->v[2].d()) / 3.0);
        maxSide3(*L, ls[0], ls[1], ls[2]);
        maxSide3(*R, rs[0], rs[1], rs[2]);
    }

    // ---------------------------------------------------------------- 2D geometry
    // CAutoSegment::CalcSegmentGeomPar / CalculateNormal (AutoSegment.cpp:35-67)
    static void calcSegmentGeom(Elem& e) {
        e.dim = mod(e.v[1].d() - e.v[0].d());
        e.maxside = e.dim;
        e.c.set((e.v[0].d() + e.v[1].d()) / 2.0);
        e.n.x = (float)((double)e.v[0].y - (double)e.v[1].y);
        e.n.y = (float)((double)e.v[1].x - (double)e.v[0].x);
        e.n.z = 0;
        const double nm = mod(e.n.d());
        if (nm > kPanelEps) e.n.set(e.n.d() / nm);
    }

    // CAutoSegment::MakeSuperSegment (AutoSegment.cpp:69-150)
    void makeSuperSegment(Elem& s, Elem* L, Elem* R) {
        if ((R->type & kDiel) && (L->type & kDiel))
            s.type |= (kSuper | kDiel | kNormalDir);
        else
            s.type |= kSuper;
        s.type &= ~kLeaf;
        s.left = L;
        s.right = R;
        s.dim = L->dim + R->dim;
        s.maxside = s.dim;
        s.c.set((L->c.d() * L->dim + R->c.d() * R->dim) / s.dim);
        const V3 lN = (L->type & kNormalDir) ? L->n.d() : -L->n.d();
        const V3 rN = (R->type & kNormalDir) ? R->n.d() : -R->n.d();
        s.n.set((lN * L->dim + rN * R->dim) / s.dim);
        const double mn = mod(s.n.d());
        const V3 normal = (mn < kPanelEps) ? lN : s.n.d() / mn;
        const V3 seg{normal.y, -normal.x, 0.0};
        s.v[0].set(s.c.d() - seg * (s.dim / 2.0));
        s.v[1].set(s.c.d() + seg * (s.dim / 2.0));
    }

    // CAutoSegment::Subdivide (AutoSegment.cpp:152-230)
    void subdivide2(Elem& p) {
        if (!p.isLeaf()) return;
        Elem* L = newElem();
        Elem* R = newElem();
        p.left = L;
        p.right = R;
        const V3 mid = (p.v[1].d() + p.v[0].d()) / 2.0;
        L->v[0] = p.v[0];
        L->v[1].set(mid);
        R->v[0].set(mid);
        R->v[1] = p.v[1];
        L->type = R->type = p.type;
        L->dielIndex = R->dielIndex = p.dielIndex;
        L->group = R->group = p.group;
        p.type &= ~kLeaf;
        L->dim = R->dim = p.dim / 2.0;
        L->maxside = L->dim;
        R->maxside = R->dim;
        L->n = R->n = p.n;
        L->c.set((L->v[0].d() + L->v[1].d()) / 2.0);
        R->c.set((R->v[0].d() + R->v[1].d()) / 2.0);
    }

    void subdivide(Elem& p) {
        if (dim == 3)
            subdivide3(p);
        else
            subdivide2(p);
    }

    // ---------------------------------------------------------------- criteria
    static void vfloat9(const Elem& e, float* v9) {
        for (int i = 0; i < 3; ++i) {
            v9[3 * i] = e.v[i].x;
            v9[3 * i + 1] = e.v[i].y;
            v9[3 * i + 2] = e.v[i].z;
        }
    }

    // CAutoRefine::RefineCriteria 3D (Autorefine.cpp:2928-3003) and 2D (:3007-3092)
    bool refineCriteria(double* c1, double* c2, const Elem& p1, const Elem& p2, V3 dist, double rdist,
                        double rmax, double eps, double ccoeff) const {
        bool force = false;
        const V3 n1 = dielNormal(p1), n2 = dielNormal(p2);
        const bool d1 = (p1.type & kDiel) != 0, d2 = (p2.type & kDiel) != 0;
        double pot1, pot2;
        float v1[9], v2[9];
        vfloat9(p1, v1);
        vfloat9(p2, v2);
        using fcg::Vec3;
        if (dim == 3) {
            if (rdist / rmax < 2.0) {
                const Vec3<double> r1{p1.c.d().x, p1.c.d().y, p1.c.d().z};
                const Vec3<double> r2{p2.c.d().x, p2.c.d().y, p2.c.d().z};
                pot1 = d1 ? fcg::enfield_num_3d<double>(r1, v2, Vec3<double>{n1.x, n1.y, n1.z}, 2)
                          : fcg::potential_num_3d<double>(r1, v2, 2);
                pot2 = d2 ? fcg::enfield_num_3d<double>(r2, v1, Vec3<double>{n2.x, n2.y, n2.z}, 2)
                          : fcg::potential_num_3d<double>(r2, v1, 2);
            } else {
                pot1 = d1 ? dot(dist, n1) / (rdist * rdist * rdist) : 1.0 / rdist;
                pot2 = d2 ? dot(-dist, n2) / (rdist * rdist * rdist) : 1.0 / rdist;
            }
            if (d1)
                *c1 = std::fabs(pot1 * p2.dim * p2.maxside / (maxSide * maxArea * ccoeff));
            else
                *c1 = std::fabs(pot1 * p2.dim / (maxArea * ccoeff));
            if (*c1 > eps) force = true;
            if (d2)
                *c2 = std::fabs(pot2 * p1.dim * p1.maxside / (maxSide * maxArea * ccoeff));
            else
                *c2 = std::fabs(pot2 * p1.dim / (maxArea * ccoeff));
            if (*c2 > eps) force = true;
        } else {
            if (rdist / rmax < 2.0) {
                pot1 = d1 ? fcg::enfield_2d<double>(p1.c.d().x, p1.c.d().y, v2, n1.x, n1.y)
                          : fcg::potential_opt_2d<double>(p1.c.d().x, p1.c.d().y, v2);
                pot2 = d2 ? fcg::enfield_2d<double>(p2.c.d().x, p2.c.d().y, v1, n2.x, n2.y)
                          : fcg::potential_opt_2d<double>(p2.c.d().x, p2.c.d().y, v1);
            } else {
                pot1 = d1 ? dot(dist, n1) / (rdist * rdist) : -std::log(rdist);
                pot2 = d2 ? dot(-dist, n2) / (rdist * rdist) : -std::log(rdist);
            }
            if (d1)
                *c1 = std::fabs(pot1 * p2.dim * p2.dim / (maxLength * maxLength * ccoeff));
            else
                *c1 = std::fabs((pot1 + 0.0) * p2.dim / (maxLength * ccoeff));
            if (*c1 > eps) force = true;
            if (d2)
                *c2 = std::fabs(pot2 * p1.dim * p1.dim / (maxLength * maxLength * ccoeff));
            else
                *c2 = std::fabs((pot2 + 0.0) * p1.dim / (maxLength * ccoeff));
            if (*c2 > eps) force = true;
        }
        return force;
    }

    // ---------------------------------------------------------------- discretization
    // CAutoRefine::DiscretizeMutual 3D/2D (Autorefine.cpp:2732-2827, 2830-2921)
    void discretizeMutual(Elem* p1, Elem* p2, bool selfCond) {
        ++level;
        if (level > maxLevel) maxLevel = level;
        const V3 dist = p1->c.d() - p2->c.d();
        const double rdist = mod(dist);
        const double r1 = (dim == 3) ? p1->maxside : p1->dim;
        const double r2 = (dim == 3) ? p2->maxside : p2->dim;
        const double rmax = (r1 > r2) ? r1 : r2;
        double cc = 1.0;
        if (selfCond) cc = (curv - 1.0) * (dot(geoNormal(*p1), geoNormal(*p2)) + 1.0) + 1.0;
        double c1 = 0, c2 = 0;
        const bool force = refineCriteria(&c1, &c2, *p1, *p2, dist, rdist, rmax, meshEps, cc);
        if (force) {
            if (p1->dim > p2->dim) {
                subdivide(*p1);
                discretizeMutual(p1->left, p2, selfCond);
                discretizeMutual(p1->right, p2, selfCond);
            } else {
                subdivide(*p2);
                discretizeMutual(p1, p2->left, selfCond);
                discretizeMutual(p1, p2->right, selfCond);
            }
        } else {
            if (!(p1->type & kSuper) && c1 > maxMeshEps) maxMeshEps = c1;
            if (!(p2->type & kSuper) && c2 > maxMeshEps) maxMeshEps = c2;
        }
        --level;
    }

    // CAutoRefine::DiscretizeSelf (Autorefine.cpp:2638-2682)
    void discretizeSelf(Elem* p) {
        ++level;
        if (level > maxLevel) maxLevel = level;
        p->linkStart = p->linkEnd = 0;
        if (p->isLeaf() && level == 1) subdivide(*p);
        if (!p->isLeaf()) {
            discretizeMutual(p->left, p->right, true);
            discretizeSelf(p->left);
            discretizeSelf(p->right);
        } else {
            ++panelNum;
        }
        ++nodeNum;
        --level;
    }

    // ---------------------------------------------------------------- links
    // CAutoRefine::RefineMutual 3D/2D (Autorefine.cpp:3096-3235, 3239-3447)
    void refineMutual(Elem* p1, Elem* p2) {
        ++level;
        if (level > maxLevel) maxLevel = level;
        const V3 dist = p1->c.d() - p2->c.d();
        const double rdist = mod(dist);
        const double r1 = (dim == 3) ? p1->maxside : p1->dim;
        const double r2 = (dim == 3) ? p2->maxside : p2->dim;
        const double rmax = (r1 > r2) ? r1 : r2;
        bool force = false;
        if (!p1->isLeaf() || !p2->isLeaf()) {
            double c1, c2;
            force = refineCriteria(&c1, &c2, *p1, *p2, dist, rdist, rmax, linkEps, 1.0);
        }
        if (force) {
            int which;
            if (!p1->isLeaf() && p2->isLeaf())
                which = 1;
            else if (p1->isLeaf() && !p2->isLeaf())
                which = 2;
            else if (dim == 3)
                which = (p1->dim > p2->dim) ? 1 : 2;
            else
                which = (r1 > r2) ? 1 : 2;
            if (which == 1) {
                refineMutual(p1->left, p2);
                refineMutual(p1->right, p2);
            } else {
                refineMutual(p1, p2->left);
                refineMutual(p1, p2->right);
            }
        } else {
            if (!computeLinks) {
                linksNum += 2;
            } else {
                colOut[p1->linkEnd] = p2->nodeId;
                colOut[p2->linkEnd] = p1->nodeId;
            }
            p1->linkEnd++;
            p2->linkEnd++;
        }
        --level;
    }

    // CAutoRefine::RefineSelf (Autorefine.cpp:2115-2229): structure part only; the self
    // coefficient is computed by the device fill.
    void refineSelf(Elem* p) {
        ++level;
        if (level > maxLevel) maxLevel = level;
        if (!p->isLeaf()) {
            refineMutual(p->left, p->right);
            refineSelf(p->left);
            refineSelf(p->right);
        }
        --level;
    }

    // CAutoRefine::ComputeLinks (Autorefine.cpp:1132-1171)
    void computeLinksPass() {
        for (size_t i = 0; i < groups.size(); ++i) {
            for (size_t j = i + 1; j < groups.size(); ++j) refineMutual(groups[i]->top, groups[j]->top);
            refineSelf(groups[i]->top);
        }
    }

    // CAutoRefine::RecurseIndex (Autorefine.cpp:1173-1220)
    void recurseIndex(Elem* p, int32_t group) {
        p->group = group;
        if (p->isLeaf()) {
            p->leafIndex = (int32_t)(basePanels + countPanels);
            ++countPanels;
            p->numChildren = 1;
        } else {
            recurseIndex(p->left, group);
            recurseIndex(p->right, group);
            p->numChildren = p->left->numChildren + p->right->numChildren;
        }
        p->nodeId = (int32_t)nodes.size();
        nodes.push_back(p);
        const uint64_t n = p->linkEnd - p->linkStart;
        p->linkEnd = (uint64_t)baseLinks;
        p->linkStart = (uint64_t)baseLinks;
        baseLinks += (int64_t)n;
    }

    // ---------------------------------------------------------------- super hierarchy
    // CAutoRefine::RecurBuild3DSuperHier / RecurBuild2DSuperHier (Autorefine.cpp:6547-6714)
    Elem* recurBuild(std::vector<Elem*>& a1, std::vector<Elem*>& a2, size_t first, size_t num, const BBox& bb) {
        if (num == 1) return a1[first];
        const int mc = bb.maxSide();
        const double mid = (bb.hi[mc] + bb.lo[mc]) / 2.0;
        size_t lo = first, hi = first + num - 1;
        for (size_t i = first; i < first + num; ++i) {
            if (a1[i]->c.d()[mc] <= mid)
                a2[lo++] = a1[i];
            else
                a2[hi--] = a1[i];
        }
        BBox lb, rb;
        for (size_t i = first; i < lo; ++i) {
            a1[i] = a2[i];
            lb.add(a1[i]->c.d());
        }
        for (size_t i = lo; i < first + num; ++i) {
            a1[i] = a2[i];
            rb.add(a1[i]->c.d());
        }
        if (lo == first || lo == first + num) {
            // every centroid on one side of the mid plane: coincident centroids
            throw std::runtime_error("coincident panel centroids while building the super hierarchy");
        }
        Elem* L = recurBuild(a1, a2, first, lo - first, lb);
        Elem* R = recurBuild(a1, a2, lo, first + num - lo, rb);
        Elem* s = newElem();
        if (dim == 3)
            makeSuperPanel(*s, L, R);
        else
            makeSuperSegment(*s, L, R);
        return s;
    }

    // CAutoRefine::BuildSuperHierarchy (Autorefine.cpp:6410-6544)
    void buildSuperHierarchy() {
        maxArea = maxSide = maxLength = totalArea = 0;
        double scale = 1.0;
        if (dim == 2) scale = kMax2DDim / mod(global2D.hi - global2D.lo);
        for (auto& gp : groups) {
            Group& g = *gp;
            if (g.panels.empty()) throw std::runtime_error("group '" + g.name + "' has no panels");
            if (dim == 2) {
                g.dielRef = g.dielRef * scale;
                for (Elem* e : g.panels) {
                    e->dim *= scale;
                    e->maxside = e->dim;
                    e->v[0].set(e->v[0].d() * scale);
                    e->v[1].set(e->v[1].d() * scale);
                    e->c.set(e->c.d() * scale);
                }
            }
            BBox bb;
            for (Elem* e : g.panels) bb.add(e->c.d());
            std::vector<Elem*> a1 = g.panels, a2(g.panels.size());
            g.top = recurBuild(a1, a2, 0, a1.size(), bb);
            g.panels.clear();
            if (dim == 3) {
                if (g.top->maxside > maxSide) maxSide = g.top->maxside;
                if (g.top->dim > maxArea) maxArea = g.top->dim;
                totalArea += g.top->dim;
            } else {
                if (g.top->dim > maxLength) maxLength = g.top->dim;
            }
        }
    }

    // ---------------------------------------------------------------- parser
    // CAutoRefine::GetConductor (Autorefine.cpp:6279-6365) and FindConductorFromName
    Group* getGroup(const std::string& name, bool isDiel, double outRe, double outIm, double inRe, double inIm,
                    V3 dielRef, uint8_t* dielIndex) {
        Group* g = nullptr;
        for (auto& gp : groups)
            if (gp->name == name) {
                g = gp.get();
                break;
            }
        if (!g) {
            auto ng = std::make_unique<Group>();
            ng->name = name;
            ng->isDiel = isDiel;
            ng->outperm[0] = outRe;
            ng->outperm[1] = outIm;
            if (isDiel) {
                ng->inperm[0] = inRe;
                ng->inperm[1] = inIm;
                ng->dielRef = dielRef;
            }
            g = ng.get();
            if (isDiel) {
                ++dielNum;
                groups.push_front(std::move(ng));
            } else {
                ++condNum;
                groups.push_back(std::move(ng));
            }
        }
        *dielIndex = 0;
        if (!isDiel) {
            size_t k = 0;
            for (; k < g->surfOutperm.size(); ++k)
                if (g->surfOutperm[k].first == outRe && g->surfOutperm[k].second == outIm) break;
            if (k == g->surfOutperm.size()) {
                if (g->surfOutperm.size() < 255) {
                    g->surfOutperm.emplace_back(outRe, outIm);
                } else {
                    k = g->surfOutperm.size() - 1;
                    std::fprintf(stderr, "Error: maximum number (255) of mediums with different dielectric constants\n");
                }
            }
            *dielIndex = (uint8_t)k;
        }
        return g;
    }

    // CAutoRefine::CreatePanel (Autorefine.cpp:5990-6083)
    void createPanel(const F3 vtx[3], Group* g, uint8_t dielIndex, bool useLocal, V3 localRef) {
        Elem* e = newElem();
        for (int i = 0; i < 3; ++i) e->v[i] = vtx[i];
        calcPanelGeom(*e);
        if (g->isDiel) {
            float tmpref;
            if (useLocal) {
                tmpref = (float)dot(localRef - e->v[0].d(), e->n.d());
                e->type |= kElementLevel;
            } else {
                tmpref = (float)dot(g->dielRef - e->v[0].d(), e->n.d());
            }
            e->type |= (tmpref > 0) ? (kDiel | kNormalDir) : kDiel;
        } else {
            e->dielIndex = dielIndex;
        }
        ++inputPanelNum;
        g->panels.push_back(e);
        g->inputPanels++;
    }

    // CAutoRefine::CreateSegment (Autorefine.cpp:6193-6275)
    void createSegment(const F3 vtx[2], Group* g, uint8_t dielIndex, bool useLocal, V3 localRef) {
        Elem* e = newElem();
        e->v[0] = vtx[0];
        e->v[1] = vtx[1];
        calcSegmentGeom(*e);
        if (g->isDiel) {
            const V3 ref = useLocal ? V3{localRef.x, localRef.y, 0} : V3{g->dielRef.x, g->dielRef.y, 0};
            const float tmpref = (float)dot(ref - e->v[0].d(), e->n.d());
            e->type |= (tmpref > 0) ? (kDiel | kNormalDir) : kDiel;
        } else {
            e->dielIndex = dielIndex;
        }
        ++inputPanelNum;
        g->panels.push_back(e);
        g->inputPanels++;
    }

    struct FileText {
        std::vector<std::string> lines;
        std::map<std::string, size_t> sections;  // sub-file name -> first line index after 'F'
    };
    std::map<std::string, std::shared_ptr<FileText>> fileCache;

    std::shared_ptr<FileText> loadFile(const std::string& name) {
        auto it = fileCache.find(name);
        if (it != fileCache.end()) return it->second;
        std::string path = name;
        if (!name.empty() && name[0] != '/' && !baseDir.empty()) path = baseDir + "/" + name;
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open file '" + path + "'");
        auto ft = std::make_shared<FileText>();
        char buf[4096];
        while (std::fgets(buf, sizeof buf, f)) ft->lines.emplace_back(buf);
        std::fclose(f);
        // CAutoRefine::CreateFileMap (Autorefine.cpp:4676-4744)
        for (size_t i = 0; i < ft->lines.size(); ++i) {
            const std::string& l = ft->lines[i];
            if (!l.empty() && l[0] == 'F') {
                char t[512], nm[512];
                if (std::sscanf(l.c_str(), "%511s %511s", t, nm) == 2) ft->sections[nm] = i + 1;
            }
        }
        fileCache[name] = ft;
        return ft;
    }

    static bool readPerm(const char*& p, float* re, float* im, bool* cplx) {
        int skip = 0;
        if (std::sscanf(p, "%f%n", re, &skip) < 1) return false;
        p += skip;
        skip = 0;
        if (std::sscanf(p, " - j %f%n", im, &skip) == 1) {
            *im = -*im;
            *cplx = true;
            p += skip;
        } else if (skip = 0, std::sscanf(p, " + j %f%n", im, &skip) == 1) {
            *cplx = true;
            p += skip;
        } else {
            *im = 0.0f;
        }
        return true;
    }

    // CAutoRefine::Parse3DInputFile / Parse2DInputFile (Autorefine.cpp:4746-5568, 5571-5988)
    void parseFile(const std::string& name, const std::shared_ptr<FileText>& parent, bool isDiel, V3 offset,
                   double outRe, double outIm, const std::string& groupname, double inRe, double inIm, V3 dielRef) {
        ++parseLevel;
        if (parseLevel >= kMaxParseLevel) throw std::runtime_error("input file recursion too deep");
        std::shared_ptr<FileText> ft;
        size_t line = 0;
        if (parent && parent->sections.count(name)) {
            ft = parent;
            line = parent->sections.at(name);
        } else {
            ft = loadFile(name);
        }
        for (; line < ft->lines.size(); ++line) {
            const std::string& ls = ft->lines[line];
            const char c0 = ls.empty() ? '\0' : ls[0];
            if (c0 == 'E' || c0 == 'e' || c0 == 'F' || c0 == 'f') break;
            const char* p = ls.c_str() + 1;
            int skip = 0;
            char nm[1024] = {0}, tmp[1024] = {0};
            bool& cplx = hasCplx;
            if (c0 == 'C') {
                if (std::sscanf(p, "%1023s%n", nm, &skip) < 1) continue;
                p += skip;
                float oRe = 1.0f, oIm = 0.0f;
                readPerm(p, &oRe, &oIm, &cplx);
                float x = 0, y = 0, z = 0;
                skip = 0;
                if (dim == 3) {
                    std::sscanf(p, "%f%f%f%n", &x, &y, &z, &skip);
                } else {
                    std::sscanf(p, "%f%f%n", &x, &y, &skip);
                }
                p += skip;
                V3 loff = V3{x, y, (dim == 3) ? (double)z : 0.0} + offset;
                std::string lg = (parseLevel == 0) ? std::string("g") : groupname;
                lg += std::to_string(groupNum[parseLevel]) + "_";
                tmp[0] = 0;
                skip = 0;
                std::sscanf(p, "%1023s%n", tmp, &skip);
                if (tmp[0] != '+') groupNum[parseLevel]++;
                groupNum[parseLevel + 1] = 1;
                parseFile(nm, ft, false, loff, oRe, oIm, lg, 0.0, 0.0, V3{});
            } else if (c0 == 'D') {
                if (std::sscanf(p, "%1023s%n", nm, &skip) < 1) continue;
                p += skip;
                float oRe = 1, oIm = 0, iRe = 1, iIm = 0;
                readPerm(p, &oRe, &oIm, &cplx);
                readPerm(p, &iRe, &iIm, &cplx);
                float x = 0, y = 0, z = 0;
                V3 loff, lref;
                skip = 0;
                if (dim == 3) {
                    std::sscanf(p, "%f%f%f%n", &x, &y, &z, &skip);
                    p += skip;
                    loff = V3{x, y, z} + offset;
                    skip = 0;
                    std::sscanf(p, "%f%f%f%n", &x, &y, &z, &skip);
                    p += skip;
                    lref = V3{x, y, z} + offset;
                } else {
                    std::sscanf(p, "%f%f%n", &x, &y, &skip);
                    p += skip;
                    loff = V3{x, y, 0} + offset;
                    skip = 0;
                    std::sscanf(p, "%f%f%n", &x, &y, &skip);
                    p += skip;
                    lref = V3{(double)x + offset.x, (double)y + offset.y, 0.0};
                }
                tmp[0] = 0;
                skip = 0;
                std::sscanf(p, "%1023s%n", tmp, &skip);
                if (tmp[0] == '-') {
                    std::swap(oRe, iRe);
                    std::swap(oIm, iIm);
                }
                // a dummy interface (Autorefine.cpp:4986-5004): non-positive real part, or
                // no contrast (with complex values read so far: in both parts)
                bool dummy = (iRe <= 0.0f || oRe <= 0.0f);
                if (cplx)
                    dummy = dummy || (std::fabs(iRe - oRe) / iRe < kRelDielTol && std::fabs(iIm - oIm) / iIm < kRelDielTol);
                else
                    dummy = dummy || std::fabs(iRe - oRe) / iRe < kRelDielTol;
                if (dummy) {
                    std::fprintf(stderr, "Warning: dummy dielectric-dielectric interface found in the input file, skipping\n");
                } else {
                    std::string lg = "diel" + std::to_string(groupDielNum);
                    groupDielNum++;
                    parseFile(nm, ft, true, loff, oRe, oIm, lg, iRe, iIm, lref);
                }
            } else if (c0 == 'T' && dim == 3) {
                if (std::sscanf(p, "%1023s%n", tmp, &skip) < 1) continue;
                p += skip;
                F3 vtx[3];
                for (int i = 0; i < 3; ++i) {
                    float x = 0, y = 0, z = 0;
                    skip = 0;
                    std::sscanf(p, "%f%f%f%n", &x, &y, &z, &skip);
                    p += skip;
                    vtx[i].set(V3{x, y, z} + offset);
                }
                float x, y, z;
                skip = 0;
                bool useLocal = (std::sscanf(p, "%f%f%f%n", &x, &y, &z, &skip) == 3);
                V3 lref = useLocal ? V3{x, y, z} + offset : V3{};
                std::string full = groupname + (isDiel ? std::string() : std::string(tmp));
                uint8_t di;
                Group* g = getGroup(full, isDiel, outRe, outIm, inRe, inIm, dielRef, &di);
                createPanel(vtx, g, di, useLocal, lref);
                if (isDiel)
                    ++rawDiel;
                else
                    ++rawCond;
            } else if (c0 == 'Q' && dim == 3) {
                if (std::sscanf(p, "%1023s%n", tmp, &skip) < 1) continue;
                p += skip;
                V3 q[4];
                for (int i = 0; i < 4; ++i) {
                    float x = 0, y = 0, z = 0;
                    skip = 0;
                    std::sscanf(p, "%f%f%f%n", &x, &y, &z, &skip);
                    p += skip;
                    q[i] = V3{x, y, z} + offset;
                }
                float x, y, z;
                skip = 0;
                bool useLocal = (std::sscanf(p, "%f%f%f%n", &x, &y, &z, &skip) == 3);
                V3 lref = useLocal ? V3{x, y, z} + offset : V3{};
                std::string full = groupname + (isDiel ? std::string() : std::string(tmp));
                uint8_t di;
                Group* g = getGroup(full, isDiel, outRe, outIm, inRe, inIm, dielRef, &di);
                addQuad(q, g, di, useLocal, lref);
                if (isDiel)
                    ++rawDiel;
                else
                    ++rawCond;
            } else if (c0 == 'S' && dim == 2) {
                if (std::sscanf(p, "%1023s%n", tmp, &skip) < 1) continue;
                p += skip;
                F3 vtx[2];
                for (int i = 0; i < 2; ++i) {
                    float x = 0, y = 0;
                    skip = 0;
                    std::sscanf(p, "%f%f%n", &x, &y, &skip);
                    p += skip;
                    vtx[i].set(V3{(double)x + offset.x, (double)y + offset.y, 0.0});
                    global2D.add(vtx[i].d());
                }
                float x, y;
                skip = 0;
                bool useLocal = (std::sscanf(p, "%f%f%n", &x, &y, &skip) == 2);
                V3 lref = useLocal ? V3{(double)x + offset.x, (double)y + offset.y, 0.0} : V3{};
                std::string full = groupname + (isDiel ? std::string() : std::string(tmp));
                uint8_t di;
                Group* g = getGroup(full, isDiel, outRe, outIm, inRe, inIm, dielRef, &di);
                createSegment(vtx, g, di, useLocal, lref);
                if (isDiel)
                    ++rawDiel;
                else
                    ++rawCond;
            } else if (c0 == 'N' && !isDiel) {
                char cn[1024] = {0}, nn[1024] = {0};
                if (std::sscanf(p, "%1023s %1023s", cn, nn) != 2) continue;
                renameGroup(groupname + cn, groupname + nn);
            }
        }
        --parseLevel;
    }

    // 'N' statement: rename, or merge into an existing group (MergeConductors,
    // Autorefine.cpp:6380-6408)
    void renameGroup(const std::string& from, const std::string& to) {
        if (from == to) return;
        Group* src = nullptr;
        Group* dst = nullptr;
        for (auto& gp : groups) {
            if (gp->name == from) src = gp.get();
            if (gp->name == to) dst = gp.get();
        }
        if (!src) return;
        if (!dst) {
            src->name = to;
            return;
        }
        dst->panels.insert(dst->panels.end(), src->panels.begin(), src->panels.end());
        dst->inputPanels += src->inputPanels;
        for (auto it = groups.begin(); it != groups.end(); ++it)
            if (it->get() == src) {
                groups.erase(it);
                break;
            }
        --condNum;
    }

    // 'Q' panel split (Autorefine.cpp:5115-5354).  Non-planar quads split along the
    // shorter diagonal; near-rectangles use the strip path; any other planar quad is
    // split along the shorter diagonal too, because the reference's Delaunay
    // triangulator (C2DTriangulate, Geometry package) is not available.
    void addQuad(V3 q[4], Group* g, uint8_t di, bool useLocal, V3 lref) {
        // plane through the Newell normal and the vertex centroid
        V3 n{0, 0, 0}, cen{0, 0, 0};
        for (int i = 0; i < 4; ++i) {
            const int j = (i + 1) % 4;
            n.x += (q[i].y - q[j].y) * (q[i].z + q[j].z);
            n.y += (q[i].z - q[j].z) * (q[i].x + q[j].x);
            n.z += (q[i].x - q[j].x) * (q[i].y + q[j].y);
            cen = cen + q[i];
        }
        const double nm = mod(n);
        if (nm < kPanelEps) {
            std::fprintf(stderr, "Warning: degenerate quadrilateral panel 'Q' found (co-linear vertexes), skipping\n");
            return;
        }
        n = n / nm;
        cen = cen / 4.0;
        bool planar = true;
        for (int i = 0; i < 4; ++i)
            if (std::fabs(dot(q[i] - cen, n)) > kPanelEps) planar = false;
        F3 t[3];
        auto tri = [&](V3 a, V3 b, V3 c) {
            t[0].set(a);
            t[1].set(b);
            t[2].set(c);
            createPanel(t, g, di, useLocal, lref);
        };
        auto diagSplit = [&]() {
            const double d1 = mod(q[2] - q[0]), d2 = mod(q[3] - q[1]);
            if (d1 <= d2) {
                tri(q[0], q[1], q[2]);
                tri(q[2], q[3], q[0]);
            } else {
                tri(q[0], q[1], q[3]);
                tri(q[1], q[2], q[3]);
            }
        };
        if (!planar) {
            diagSplit();
            return;
        }
        V3 sides[4] = {q[1] - q[0], q[2] - q[1], q[3] - q[2], q[0] - q[3]};
        double ds[4];
        for (int i = 0; i < 4; ++i) ds[i] = mod(sides[i]);
        const double cosang = dot(sides[0], sides[1]) / (ds[0] * ds[1]);
        if (std::fabs(ds[2] - ds[0]) < ds[0] * kQuadSideTol && std::fabs(ds[3] - ds[1]) < ds[1] * kQuadSideTol &&
            std::fabs(cosang) < kQuadAngleTol) {
            if (ds[0] < ds[1]) {
                V3 t0 = q[0];
                q[0] = q[1];
                q[1] = q[2];
                q[2] = q[3];
                q[3] = t0;
                V3 s0 = sides[0];
                sides[0] = sides[1];
                sides[1] = sides[2];
                sides[2] = sides[3];
                sides[3] = s0;
                double m0 = ds[0];
                ds[0] = ds[1];
                ds[1] = ds[2];
                ds[2] = ds[3];
                ds[3] = m0;
            }
            long numsegs = (long)(ds[0] / (ds[1] * 5.0) + 0.5);
            V3 nodes4[4];
            nodes4[0] = q[0];
            nodes4[3] = q[3];
            if (numsegs == 0) numsegs = 1;
            const V3 span01 = sides[0] / (double)numsegs;
            const V3 span32 = -sides[2] / (double)numsegs;
            for (long i = 0; i < numsegs - 1; ++i) {
                nodes4[1] = nodes4[0] + span01;
                nodes4[2] = nodes4[3] + span32;
                tri(nodes4[0], nodes4[1], nodes4[2]);
                tri(nodes4[0], nodes4[2], nodes4[3]);
                nodes4[0] = nodes4[1];
                nodes4[3] = nodes4[2];
            }
            nodes4[1] = q[1];
            nodes4[2] = q[2];
            tri(nodes4[0], nodes4[1], nodes4[2]);
            tri(nodes4[0], nodes4[2], nodes4[3]);
        } else {
            diagSplit();
        }
    }
};

}  // namespace

struct fch_mesh {
    // a build stopped after the super hierarchy (refine_on_device): the builder,
    // its coarse tree in flat form and the elements behind the flat indices
    std::unique_ptr<Builder> pending;
    std::vector<fcg_refine_node> coarse;
    std::vector<Elem*> coarse_elems;
    std::vector<int32_t> coarse_tops;
    fcg_refine_params refine_params{};
    double t_read_s = 0, t_superh_s = 0;

    int dim = 3;
    std::vector<fcg_node_geom> nodes;
    std::vector<int32_t> left, right, parent, leaf_lo, leaf_hi, leaf_node;
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;
    std::vector<int32_t> group_is_diel, group_leaf_start;
    std::vector<double> group_outperm, group_inperm, leaf_outperm;
    bool cplx = false;  // some permittivity is complex: the *_im arrays below are exported
    std::vector<double> group_outperm_im, group_inperm_im, leaf_outperm_im;
    std::vector<std::string> group_names;
    // for the FastCap-format writers: each group's outer media (conductors),
    // dielectric reference point, and each leaf's medium index
    std::vector<std::vector<std::pair<double, double>>> group_surf;
    std::vector<V3> group_diel_ref;
    std::vector<uint8_t> leaf_diel_index;
    fch_stats stats{};
    fcg_link_params link_params{};
};

extern "C" {

void fch_default_options(fch_options* o) {
    o->mesh_eps = 0.1;
    o->eps_ratio = 1.0;
    o->curv_coeff = 3.0;
    o->verbose = 0;
    o->links_on_device = 0;
    o->refine_on_device = 0;
}

const char* fch_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {
using clk = std::chrono::steady_clock;
double seconds(clk::time_point a, clk::time_point c) { return std::chrono::duration<double>(c - a).count(); }

// runs f, mapping the builder's exceptions to FC_* codes
template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        g_err = "out of memory";
        return FCG_OUT_OF_MEMORY;
    } catch (const std::runtime_error& e) {
        g_err = e.what();
        if (g_err.rfind("cannot open", 0) == 0) return FCG_CANNOT_OPEN_FILE;
        return FCG_FILE_ERROR;
    } catch (...) {
        g_err = "unhandled exception";
        return FCG_EXCEPTION_ERROR;
    }
}

// RecurseIndex, the link record pass (unless the links come from the device)
// and the flattening into the mesh arrays
std::unique_ptr<fch_mesh> finish_build(Builder& b, bool links_on_device, double t_read, double t_superh,
                                       double t_disc) {
        auto t3 = clk::now();
        // AutoRefineLinks structure part: RecurseIndex, then the record pass.
        b.nodes.reserve((size_t)b.nodeNum);
        b.baseLinks = 0;
        b.basePanels = 0;
        for (size_t gi = 0; gi < b.groups.size(); ++gi) {
            b.countPanels = 0;
            b.recurseIndex(b.groups[gi]->top, (int32_t)gi);
            b.basePanels += b.countPanels;
            b.groups[gi]->leafNum = b.countPanels;
        }
        if ((int64_t)b.nodes.size() != b.nodeNum || b.basePanels != b.panelNum || b.baseLinks != b.linksNum)
            throw std::runtime_error("internal error: node/panel/link counts disagree between passes");
        if (b.linksNum > (int64_t)INT32_MAX * 8 || b.nodeNum > (int64_t)INT32_MAX)
            throw std::runtime_error("mesh too large");
        b.colOut.assign((size_t)b.linksNum, -1);
        b.computeLinks = true;
        if (!links_on_device) b.computeLinksPass();
        auto t4 = clk::now();

        auto m = std::make_unique<fch_mesh>();
        m->dim = b.dim;
        m->link_params.dim = b.dim;
        m->link_params.link_eps = b.linkEps;
        m->link_params.max_side = b.maxSide;
        m->link_params.max_area = b.maxArea;
        m->link_params.max_length = b.maxLength;
        const int64_t M = b.nodeNum, N = b.panelNum, L = b.linksNum;
        m->nodes.resize((size_t)M);
        m->left.assign((size_t)M, -1);
        m->right.assign((size_t)M, -1);
        m->parent.assign((size_t)M, -1);
        m->leaf_lo.resize((size_t)M);
        m->leaf_hi.resize((size_t)M);
        m->leaf_node.resize((size_t)N);
        m->row_ptr.resize((size_t)M + 1);
        for (int64_t i = 0; i < M; ++i) {
            Elem* e = b.nodes[(size_t)i];
            fcg_node_geom& g = m->nodes[(size_t)i];
            std::memset(&g, 0, sizeof g);
            for (int k = 0; k < 3; ++k) {
                g.v[3 * k] = e->v[k].x;
                g.v[3 * k + 1] = e->v[k].y;
                g.v[3 * k + 2] = e->v[k].z;
            }
            g.c[0] = e->c.x;
            g.c[1] = e->c.y;
            g.c[2] = e->c.z;
            g.flags = e->type;
            g.cond = e->group;
            g.maxside = (b.dim == 3) ? e->maxside : e->dim;
            g.area = e->dim;
            const V3 dn = Builder::dielNormal(*e);
            g.dn[0] = dn.x;
            g.dn[1] = dn.y;
            g.dn[2] = dn.z;
            if (!e->isLeaf()) {
                m->left[(size_t)i] = e->left->nodeId;
                m->right[(size_t)i] = e->right->nodeId;
                m->parent[(size_t)e->left->nodeId] = (int32_t)i;
                m->parent[(size_t)e->right->nodeId] = (int32_t)i;
            }
            m->row_ptr[(size_t)i] = (int64_t)e->linkStart;
        }
        m->row_ptr[(size_t)M] = L;
        for (int64_t i = 0; i < M; ++i) {
            Elem* e = b.nodes[(size_t)i];
            if (e->linkEnd != (uint64_t)m->row_ptr[(size_t)i + 1])
                throw std::runtime_error("internal error: link ranges are not contiguous");
        }
        // leaf ranges: post-order children precede parents
        for (int64_t i = 0; i < M; ++i) {
            Elem* e = b.nodes[(size_t)i];
            if (e->isLeaf()) {
                m->leaf_lo[(size_t)i] = e->leafIndex;
                m->leaf_hi[(size_t)i] = e->leafIndex + 1;
                m->leaf_node[(size_t)e->leafIndex] = (int32_t)i;
            } else {
                m->leaf_lo[(size_t)i] = m->leaf_lo[(size_t)e->left->nodeId];
                m->leaf_hi[(size_t)i] = m->leaf_hi[(size_t)e->right->nodeId];
            }
        }
        m->col.resize((size_t)L);
        for (int64_t k = 0; k < L; ++k) {
            if (b.colOut[(size_t)k] < 0) throw std::runtime_error("internal error: unrecorded link slot");
            m->col[(size_t)k] = b.colOut[(size_t)k];
        }
        const size_t G = b.groups.size();
        m->group_is_diel.resize(G);
        m->group_leaf_start.resize(G + 1);
        m->group_outperm.resize(G);
        m->group_inperm.resize(G);
        m->group_names.resize(G);
        m->leaf_outperm.assign((size_t)N, 0.0);
        m->cplx = b.hasCplx;
        m->group_outperm_im.assign(G, 0.0);
        m->group_inperm_im.assign(G, 0.0);
        m->leaf_outperm_im.assign((size_t)N, 0.0);
        int64_t ls = 0;
        for (size_t gi = 0; gi < G; ++gi) {
            Group& g = *b.groups[gi];
            m->group_is_diel[gi] = g.isDiel ? 1 : 0;
            m->group_leaf_start[gi] = (int32_t)ls;
            m->group_outperm[gi] = g.outperm[0];
            m->group_inperm[gi] = g.inperm[0];
            m->group_outperm_im[gi] = g.outperm[1];
            m->group_inperm_im[gi] = g.inperm[1];
            m->group_names[gi] = g.name;
            m->group_surf.push_back(g.surfOutperm);
            m->group_diel_ref.push_back(g.dielRef);
            ls += g.leafNum;
        }
        m->leaf_diel_index.assign((size_t)N, 0);
        m->group_leaf_start[G] = (int32_t)ls;
        for (int64_t i = 0; i < M; ++i) {
            Elem* e = b.nodes[(size_t)i];
            if (!e->isLeaf()) continue;
            Group& g = *b.groups[(size_t)e->group];
            m->leaf_diel_index[(size_t)e->leafIndex] = e->dielIndex;
            if (!g.isDiel) {
                m->leaf_outperm[(size_t)e->leafIndex] = g.surfOutperm[e->dielIndex].first;
                m->leaf_outperm_im[(size_t)e->leafIndex] = g.surfOutperm[e->dielIndex].second;
            }
        }
        fch_stats& s = m->stats;
        s.input_raw_cond_panels = b.rawCond;
        s.input_raw_diel_panels = b.rawDiel;
        s.input_panels = b.inputPanelNum;
        s.panels = N;
        s.nodes = M;
        s.links = L;
        s.n_conductors = (int32_t)b.condNum;
        s.n_dielectrics = (int32_t)b.dielNum;
        s.dim = b.dim;
        s.max_level = b.maxLevel;
        s.max_mesh_eps = b.maxMeshEps;
        s.t_read_s = t_read;
        s.t_superh_s = t_superh;
        s.t_discretize_s = t_disc;
        s.t_links_s = seconds(t3, t4);
        return m;
}
}  // namespace

extern "C" {

int fch_build(const char* path, const fch_options* opts, fch_mesh** out) {
    g_err.clear();
    if (!path || !out) {
        g_err = "null argument";
        return FCG_INVALID_ARGUMENT;
    }
    *out = nullptr;
    return guarded([&]() -> int {
        auto bp = std::make_unique<Builder>();
        Builder& b = *bp;
        fch_options o;
        if (opts)
            o = *opts;
        else
            fch_default_options(&o);
        b.opt = o;
        b.meshEps = o.mesh_eps;
        // CSolveCap::Run: link eps = ratio * mesh eps, 0 -> 0.01; curvature >= 1 (SolveCapacitance.cpp:108-119)
        b.linkEps = o.eps_ratio * o.mesh_eps;
        if (b.linkEps == 0.0) b.linkEps = 0.01;
        b.curv = o.curv_coeff < 1.0 ? 1.0 : o.curv_coeff;
        for (int i = 0; i < kMaxParseLevel + 2; ++i) b.groupNum[i] = 1;

        std::string p(path);
        const size_t slash = p.find_last_of('/');
        std::string fname = p;
        if (slash != std::string::npos) {
            b.baseDir = p.substr(0, slash);
            fname = p.substr(slash + 1);
        }
        auto t0 = clk::now();
        // ReadFastCapFile: 2D when the first line mentions "2D" (Autorefine.cpp:4650)
        {
            auto ft = b.loadFile(fname);
            const std::string first = ft->lines.empty() ? std::string() : ft->lines[0];
            b.dim = (first.find("2D") != std::string::npos || first.find("2d") != std::string::npos) ? 2 : 3;
        }
        b.parseFile(fname, nullptr, false, V3{}, 1.0, 0.0, "", 0.0, 0.0, V3{});
        if (b.groups.empty()) throw std::runtime_error("no panels found in the input file");
        auto t1 = clk::now();
        b.buildSuperHierarchy();
        auto t2 = clk::now();

        if (o.refine_on_device) {
            // stop here: the coarse tree goes to the device (fch_get_coarse_tree,
            // fcg_refine, fch_set_refined_tree); pre-order, children after parents
            auto m = std::make_unique<fch_mesh>();
            m->dim = b.dim;
            m->t_read_s = seconds(t0, t1);
            m->t_superh_s = seconds(t1, t2);
            for (size_t gi = 0; gi < b.groups.size(); ++gi) {
                std::vector<Elem*> st{b.groups[gi]->top};
                m->coarse_tops.push_back((int32_t)m->coarse.size());
                while (!st.empty()) {
                    Elem* e = st.back();
                    st.pop_back();
                    e->nodeId = (int32_t)m->coarse.size();
                    fcg_refine_node r{};
                    for (int k = 0; k < 3; ++k) {
                        r.v[3 * k] = e->v[k].x;
                        r.v[3 * k + 1] = e->v[k].y;
                        r.v[3 * k + 2] = e->v[k].z;
                    }
                    r.c[0] = e->c.x;
                    r.c[1] = e->c.y;
                    r.c[2] = e->c.z;
                    r.n[0] = e->n.x;
                    r.n[1] = e->n.y;
                    r.n[2] = e->n.z;
                    r.dim = e->dim;
                    r.maxside = e->maxside;
                    r.left = r.right = -1;
                    r.group = (int32_t)gi;
                    r.type = e->type;
                    r.max_side_idx = e->maxSideIdx;
                    r.diel_index = e->dielIndex;
                    m->coarse.push_back(r);
                    m->coarse_elems.push_back(e);
                    if (!e->isLeaf()) {
                        st.push_back(e->right);
                        st.push_back(e->left);
                    }
                }
            }
            // children indices, now that every element has one
            for (size_t i = 0; i < m->coarse.size(); ++i) {
                Elem* e = m->coarse_elems[i];
                if (e->isLeaf()) continue;
                m->coarse[i].left = e->left->nodeId;
                m->coarse[i].right = e->right->nodeId;
            }
            fcg_refine_params& rp = m->refine_params;
            rp.dim = b.dim;
            rp.n_groups = (int32_t)b.groups.size();
            rp.mesh_eps = b.meshEps;
            rp.curv = b.curv;
            rp.max_side = b.maxSide;
            rp.max_area = b.maxArea;
            rp.max_length = b.maxLength;
            m->pending = std::move(bp);
            *out = m.release();
            return FCG_OK;
        }

        // AutoRefinePanels (Autorefine.cpp:293-479): mutual then self discretization, then
        // the counting pass of ComputeLinks.
        b.level = 0;
        b.maxLevel = 0;
        b.maxMeshEps = 0.0;
        for (size_t i = 0; i < b.groups.size(); ++i)
            for (size_t j = i + 1; j < b.groups.size(); ++j)
                b.discretizeMutual(b.groups[i]->top, b.groups[j]->top, false);
        for (auto& g : b.groups) {
            b.level = 0;
            b.discretizeSelf(g->top);
        }
        b.computeLinks = false;
        b.linksNum = 0;
        if (!o.links_on_device) b.computeLinksPass();
        auto t3 = clk::now();
        auto m = finish_build(b, o.links_on_device != 0, seconds(t0, t1), seconds(t1, t2), seconds(t2, t3));
        *out = m.release();
        return FCG_OK;
    });
}

int fch_get_coarse_tree(const fch_mesh* m, const fcg_refine_node** nodes, int* n_nodes, const int32_t** tops,
                        fcg_refine_params* params) {
    if (!m || !nodes || !n_nodes || !tops || !params) return FCG_INVALID_ARGUMENT;
    if (!m->pending) {
        g_err = "mesh is not waiting for a refinement";
        return FCG_INVALID_ARGUMENT;
    }
    *nodes = m->coarse.data();
    *n_nodes = (int)m->coarse.size();
    *tops = m->coarse_tops.data();
    *params = m->refine_params;
    return FCG_OK;
}

int fch_set_refined_tree(fch_mesh* m, const fcg_refine_node* nodes, int n, double max_mesh_eps, int32_t max_level,
                         double t_discretize_s) {
    if (!m || !nodes || !m->pending) return FCG_INVALID_ARGUMENT;
    const int K0 = (int)m->coarse.size();
    if (n < K0) return FCG_INVALID_ARGUMENT;
    return guarded([&]() -> int {
        Builder& b = *m->pending;
        std::vector<Elem*> E((size_t)n);
        for (int i = 0; i < n; ++i) E[(size_t)i] = i < K0 ? m->coarse_elems[(size_t)i] : b.newElem();
        for (int i = 0; i < n; ++i) {
            const fcg_refine_node& r = nodes[i];
            Elem& e = *E[(size_t)i];
            if (i >= K0) {
                for (int k = 0; k < 3; ++k) {
                    e.v[k].x = r.v[3 * k];
                    e.v[k].y = r.v[3 * k + 1];
                    e.v[k].z = r.v[3 * k + 2];
                }
                e.c = F3{r.c[0], r.c[1], r.c[2]};
                e.n = F3{r.n[0], r.n[1], r.n[2]};
                e.dim = r.dim;
                e.maxside = r.maxside;
                e.maxSideIdx = r.max_side_idx;
                e.dielIndex = r.diel_index;
                e.group = r.group;
            }
            e.type = r.type;
            if ((r.left < 0) != (r.right < 0) || (r.left >= 0 && (r.left <= i || r.right <= i || r.left >= n ||
                                                                   r.right >= n)) ||
                ((r.left < 0) != e.isLeaf()))
                throw std::runtime_error("refined tree: bad children");
            e.left = r.left >= 0 ? E[(size_t)r.left] : nullptr;
            e.right = r.right >= 0 ? E[(size_t)r.right] : nullptr;
        }
        // DiscretizeSelf's counts: every node of every group tree, and the leaves
        b.nodeNum = b.panelNum = 0;
        for (auto& g : b.groups) {
            std::vector<Elem*> st{g->top};
            while (!st.empty()) {
                Elem* e = st.back();
                st.pop_back();
                ++b.nodeNum;
                if (e->isLeaf()) {
                    ++b.panelNum;
                } else {
                    st.push_back(e->right);
                    st.push_back(e->left);
                }
            }
        }
        b.maxMeshEps = max_mesh_eps;
        b.maxLevel = max_level;
        auto fm = finish_build(b, true, m->t_read_s, m->t_superh_s, t_discretize_s);
        *m = std::move(*fm);
        return FCG_OK;
    });
}

int fch_get_desc(const fch_mesh* m, fcg_job_desc* d) {
    if (!m || !d) return FCG_INVALID_ARGUMENT;
    if (m->pending) {
        g_err = "mesh is waiting for its refined tree (fch_set_refined_tree)";
        return FCG_INVALID_ARGUMENT;
    }
    std::memset(d, 0, sizeof *d);
    d->dim = m->dim;
    d->n_leaves = (int32_t)m->leaf_node.size();
    d->n_nodes = (int32_t)m->nodes.size();
    d->n_groups = (int32_t)m->group_is_diel.size();
    d->n_links = (int64_t)m->col.size();
    d->nodes = m->nodes.data();
    d->left = m->left.data();
    d->right = m->right.data();
    d->parent = m->parent.data();
    d->leaf_lo = m->leaf_lo.data();
    d->leaf_hi = m->leaf_hi.data();
    d->leaf_node = m->leaf_node.data();
    d->row_ptr = m->row_ptr.data();
    d->col = m->col.data();
    d->group_is_diel = m->group_is_diel.data();
    d->group_leaf_start = m->group_leaf_start.data();
    d->group_outperm = m->group_outperm.data();
    d->group_inperm = m->group_inperm.data();
    d->leaf_outperm = m->leaf_outperm.data();
    d->group_outperm_im = m->cplx ? m->group_outperm_im.data() : nullptr;
    d->group_inperm_im = m->cplx ? m->group_inperm_im.data() : nullptr;
    d->leaf_outperm_im = m->cplx ? m->leaf_outperm_im.data() : nullptr;
    return FCG_OK;
}

int fch_get_link_params(const fch_mesh* m, fcg_link_params* p) {
    if (!m || !p) return FCG_INVALID_ARGUMENT;
    *p = m->link_params;
    return FCG_OK;
}

int fch_set_links(fch_mesh* m, const int64_t* row_ptr, const int32_t* col, int64_t L, int32_t max_level,
                  double t_links_s) {
    if (!m || L < 0) return FCG_INVALID_ARGUMENT;
    if (!row_ptr && !col) {
        m->stats.links = L;
        m->stats.max_level = std::max(m->stats.max_level, max_level);
        m->stats.t_links_s = t_links_s;
        return FCG_OK;
    }
    if (!row_ptr || (L > 0 && !col)) return FCG_INVALID_ARGUMENT;
    const size_t M = m->nodes.size();
    if (row_ptr[0] != 0 || row_ptr[M] != L) return FCG_INVALID_ARGUMENT;
    for (size_t i = 0; i < M; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return FCG_INVALID_ARGUMENT;
    for (int64_t k = 0; k < L; ++k)
        if (col[k] < 0 || (size_t)col[k] >= M) return FCG_INVALID_ARGUMENT;
    m->row_ptr.assign(row_ptr, row_ptr + M + 1);
    m->col.assign(col, col + L);
    m->stats.links = L;
    m->stats.max_level = std::max(m->stats.max_level, max_level);
    m->stats.t_links_s = t_links_s;
    return FCG_OK;
}

int fch_get_stats(const fch_mesh* m, fch_stats* s) {
    if (!m || !s) return FCG_INVALID_ARGUMENT;
    *s = m->stats;
    return FCG_OK;
}

const char* fch_group_name(const fch_mesh* m, int g) {
    if (!m || g < 0 || (size_t)g >= m->group_names.size()) return nullptr;
    return m->group_names[(size_t)g].c_str();
}

// CAutoRefine::OutputFastCapFile / OutputPanelFile / OutputPanelTree
// (Solver/Autorefine.cpp:1230-1470): a list file <base>_<suffix>.lst naming one
// panel file per conductor surface (outer medium) and per dielectric, each
// listing its group's leaves in tree order; with charges, every panel line ends
// with the leaf's charge density.
int fch_write_fastcap(const fch_mesh* m, const char* filein, const char* suffix, const double* charges) {
    if (!m || !filein || !suffix || m->pending) return FCG_INVALID_ARGUMENT;
    constexpr double kZ2D = 1.0;                     // AUTOREFINE_2D_ZCOORD
    constexpr double kTwoPiE0 = 5.563250279878e-11;  // TWO_PI_TIMES_E0, Solver/Potential.h:42
    std::string base = filein;
    const size_t dot = base.rfind('.');
    if (dot != std::string::npos) base.resize(dot);
    const std::string lst = base + "_" + suffix + ".lst";
    FILE* fl = std::fopen(lst.c_str(), "w");
    if (!fl) {
        std::fprintf(stderr, "  Warning: cannot open \"%s\" for writing, skipping\n", lst.c_str());
        return FCG_FILE_ERROR;
    }
    std::printf("  generating list file \"%s\"\n", lst.c_str());
    std::fprintf(fl, "* Refined FastCap file derived from %s\n*\n* Total number of panels %lu\n*\n", filein,
                 (unsigned long)m->leaf_node.size());
    auto panel_file = [&](const std::string& name, size_t g, int diel_index) {
        FILE* f = std::fopen(name.c_str(), "w");
        if (!f) {
            std::fprintf(stderr, "  Warning: cannot open \"%s\" for writing, skipping\n", name.c_str());
            return;
        }
        std::fprintf(f, "0 Refined FastCap file derived from %s, referenced in list file %s\n*\n", filein,
                     lst.c_str());
        const char* cname = m->group_names[g].c_str();
        for (int32_t k = m->group_leaf_start[g]; k < m->group_leaf_start[g + 1]; ++k) {
            if (diel_index >= 0 && m->leaf_diel_index[(size_t)k] != diel_index) continue;
            const fcg_node_geom& e = m->nodes[(size_t)m->leaf_node[(size_t)k]];
            if (m->dim == 3) {
                std::fprintf(f, "T %s  %g %g %g  %g %g %g  %g %g %g", cname, e.v[0], e.v[1], e.v[2], e.v[3], e.v[4],
                             e.v[5], e.v[6], e.v[7], e.v[8]);
            } else {
                std::fprintf(f, "Q %s  %g %g %g %g %g %g %g %g %g %g %g %g", cname, e.v[0], e.v[1], 0.0, e.v[3],
                             e.v[4], 0.0, e.v[3], e.v[4], kZ2D, e.v[0], e.v[1], kZ2D);
            }
            if (e.flags & FCG_NODE_OUTPERM_ELEMENT_LEVEL) {
                // local dielectric reference point: centroid plus dielectric normal
                const double rx = (double)e.c[0] + e.dn[0], ry = (double)e.c[1] + e.dn[1];
                const double rz = m->dim == 3 ? (double)e.c[2] + e.dn[2] : kZ2D / 2.0;
                std::fprintf(f, "  %g %g %g", rx, ry, rz);
            }
            if (charges) {
                const double q = m->dim == 3 ? charges[k] : charges[k] * kTwoPiE0;
                std::fprintf(f, "  %e\n", q / e.area);
            } else {
                std::fprintf(f, "\n");
            }
        }
        std::fclose(f);
    };
    for (size_t g = 0; g < m->group_names.size(); ++g) {
        std::string cf = base + "_" + m->group_names[g] + "_" + suffix;
        // the reference strips a path up to the last backslash only
        const size_t bs = cf.rfind('\\');
        const std::string nopath = bs != std::string::npos ? cf.substr(bs + 1) : cf;
        if (!m->group_is_diel[g]) {
            const auto& surf = m->group_surf[g];
            for (size_t i = 0; i < surf.size(); ++i) {
                const std::string ext = "_" + std::to_string(i) + ".txt";
                std::fprintf(fl, "C %s  %g  0.0 0.0 0.0%s\n", (nopath + ext).c_str(), surf[i].first,
                             i + 1 == surf.size() ? "" : " +");
                panel_file(cf + ext, g, (int)i);
            }
        } else {
            const V3& r = m->group_diel_ref[g];
            std::fprintf(fl, "D %s  %g %g  0.0 0.0 0.0  %g %g %g\n", (nopath + ".txt").c_str(), m->group_outperm[g],
                         m->group_inperm[g], r.x, r.y, m->dim == 3 ? r.z : kZ2D / 2.0);
            panel_file(cf + ".txt", g, -1);
        }
    }
    std::fclose(fl);
    return FCG_OK;
}

void fch_free(fch_mesh* m) { delete m; }

}  // extern "C"
