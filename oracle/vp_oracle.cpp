// FP64 CPU oracle for the vortex-particle kernels of CPanel.
//
// TEST INFRASTRUCTURE ONLY: restates the reference's particle kernels, the
// particle octree, its pseudo-particle "expansions" and the Barnes-Hut
// traversal so the HIP product path can be checked.  The product never links
// or calls this code.  Citations are file:line in jakeddavis/CPanel
// (CPanel/ and octree/include/).
//
// Orientation of the pair kernels.  The Barnes-Hut traversal calls the
// kernels on the SOURCE (the tree member or the node's pseudo-particle) with
// the TARGET as argument (node.h:236-346): member->velInfl(target),
// member->vortexStretching(target), member->viscousDiffusion(target).  The
// oracle's pair functions take (target, source) and evaluate exactly that
// call.  (The reference's non-accelerated loop in cpCaseVP.cpp:490-495 calls
// particles[i]->vortexStretching(particles[j]) with i the target, i.e. the
// other orientation; the accelerated path is the one restated here.)
#include <cmath>
#include <cstdint>
#include <vector>

#include "cpanel_oracle.h"

namespace {

const double PI = 3.14159265358979323846;
const double CORE_OVERLAP = 1.5;  // particle::coreOverlap (particle.h:33)
const double NU = 1.983e-5;       // particle::viscousDiffusion (particle.cpp:126)

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

struct Part {
    V3 pos, str;
    double radius;
};

// particle::velInflAlgSmooth (particle.cpp:31-38): source p at point poi
V3 vel_alg(const Part& p, V3 poi) {
    double sigma = CORE_OVERLAP * p.radius;
    V3 dist = poi - p.pos;
    double r = norm(dist);
    return (-1 / (4 * PI) * (std::pow(r, 2) + 2.5 * std::pow(sigma, 2)) /
            std::pow(std::pow(r, 2) + std::pow(sigma, 2), 2.5)) *
           cross(dist, p.str);
}

// The configured unsteady case's particle kernel: velInflAlgSmooth's
// algebraic smoothing (particle.cpp:31-38) with one core radius `core` for
// every particle in place of 1.5 x the particle's radius.
V3 vel_alg_core(V3 spos, V3 sstr, V3 poi, double core) {
    V3 dist = poi - spos;
    double r2 = dot(dist, dist), s2 = core * core;
    return (-1 / (4 * PI) * (r2 + 2.5 * s2) / std::pow(r2 + s2, 2.5)) * cross(dist, sstr);
}

// (alpha_t . grad) of vel_alg_core at the target: with x = poi - spos,
// u = f(r) x x alpha_s, f = -(r^2 + 2.5 s^2) / (4 pi (r^2 + s^2)^2.5),
// d u / d x_k = f'(r) x_k / r (x x alpha_s) + f e_k x alpha_s, and
// f'(r) / r = (3 r^2 + 10.5 s^2) / (4 pi (r^2 + s^2)^3.5)
V3 stretch_alg_core(V3 spos, V3 sstr, V3 tpos, V3 tstr, double core) {
    V3 x = tpos - spos;
    double r2 = dot(x, x), s2 = core * core;
    double f = -1 / (4 * PI) * (r2 + 2.5 * s2) / std::pow(r2 + s2, 2.5);
    double h = 1 / (4 * PI) * (3 * r2 + 10.5 * s2) / std::pow(r2 + s2, 3.5);
    return (h * dot(tstr, x)) * cross(x, sstr) + f * cross(tstr, sstr);
}

// particle::velInfl(particle*) (particle.cpp:42-58): source s on target t
V3 vel_gauss(const Part& s, const Part& t) {
    double sigma = std::sqrt(std::pow(CORE_OVERLAP * s.radius, 2) + std::pow(CORE_OVERLAP * t.radius, 2)) /
                   std::sqrt(2);
    V3 dist = t.pos - s.pos;
    double rho = norm(dist) / sigma;
    double K = (1 / (4 * PI * rho) * std::erf(rho / std::pow(2, 0.5)) -
                1 / (std::pow(2 * PI, 1.5)) * std::exp(-0.5 * rho * rho)) /
               (rho * rho);
    return (-1 / std::pow(sigma, 3) * K) * cross(dist, s.str);
}

// particle::velInfl(const Vector3d&) (particle.cpp:60-77): Gaussian at a point
V3 vel_gauss_point(const Part& s, V3 poi) {
    if (norm(s.pos - poi) == 0) return {0, 0, 0};
    double sigma = CORE_OVERLAP * s.radius;
    V3 dist = poi - s.pos;
    double rho = norm(dist) / sigma;
    double K = (1 / (4 * PI * rho) * std::erf(rho / std::pow(2, 0.5)) -
                1 / (std::pow(2 * PI, 1.5)) * std::exp(-0.5 * rho * rho)) /
               (rho * rho);
    return (-1 / std::pow(sigma, 3) * K) * cross(dist, s.str);
}

// particle::vortexStretching (particle.cpp:80-118), called as
// source->vortexStretching(target): Xi = source, Xj = target; the cross
// matrix is the target's strength and it multiplies the source's strength.
V3 stretch(const Part& s, const Part& t) {
    V3 Xi = s.pos, Xj = t.pos;
    if (norm(Xi - Xj) > 5 * s.radius) return {0, 0, 0};
    double sigij = std::sqrt(std::pow(CORE_OVERLAP * s.radius, 2) + std::pow(CORE_OVERLAP * t.radius, 2)) /
                   std::sqrt(2);
    double rho = norm(Xi - Xj) / sigij;
    double xi = 1 / (std::pow(2 * PI, 1.5)) * std::exp(-rho * rho / 2);
    double G = 1 / (4 * PI * rho) * std::erf(rho / std::sqrt(2));
    double K = (G - xi) / (rho * rho);
    double F = (3 * K - xi) / (rho * rho);
    double M[3][3];
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
            double d = -1 / (sigij * sigij) * F * (comp(Xi, k) - comp(Xj, k)) * (comp(Xi, l) - comp(Xj, l));
            M[k][l] = k == l ? K - 1 / (sigij * sigij) * F * (comp(Xi, k) - comp(Xj, k)) * (comp(Xi, l) - comp(Xj, l))
                             : d;
        }
    V3 a = t.str;
    double At[3][3] = {{0, -a.z, a.y}, {a.z, 0, -a.x}, {-a.y, a.x, 0}};
    double s3 = 1 / std::pow(sigij, 3);
    double G3[3][3];
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
            double acc = At[k][0] * M[0][l];
            acc += At[k][1] * M[1][l];
            acc += At[k][2] * M[2][l];
            G3[k][l] = s3 * acc;
        }
    V3 out;
    double o[3];
    for (int k = 0; k < 3; k++) o[k] = G3[k][0] * s.str.x + G3[k][1] * s.str.y + G3[k][2] * s.str.z;
    out = {o[0], o[1], o[2]};
    return out;
}

// particle::viscousDiffusion (particle.cpp:122-140), source->viscousDiffusion(target)
V3 diffusion(const Part& s, const Part& t) {
    V3 Xi = s.pos, Xj = t.pos;
    double sigij = std::sqrt(std::pow(CORE_OVERLAP * s.radius, 2) + std::pow(CORE_OVERLAP * t.radius, 2)) /
                   std::sqrt(2);
    double rho = norm(Xi - Xj) / sigij;
    double Vi = 4 * PI / 3 * std::pow(s.radius, 3);
    double Vj = 4 * PI / 3 * std::pow(t.radius, 3);
    double xi = 1 / (std::pow(2 * PI, 1.5) * std::pow(rho, 3)) * std::exp(-rho * rho / 2);
    return (2 * NU / (sigij * sigij) * xi) * (Vi * t.str - Vj * s.str);
}

// ------------------------------------------------------------------ octree
// node<particle> (octree/include/node.h:30-214) and octree<particle>
// (octree.h:39-195) restated with the same insertion semantics.
struct Node {
    Node* children[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    V3 origin, half;
    std::vector<int> members;  // particle indices, insertion order
    size_t level;
    Part exp;                  // multExp (particleFMM.cpp:29-58)
    bool has_exp = false;
    ~Node() {
        for (Node* c : children) delete c;
    }
    bool leaf() const {
        for (Node* c : children)
            if (c) return false;
        return true;
    }
};

struct Tree {
    std::vector<Part> parts;
    Node* root = nullptr;
    size_t max_members = 10;
    double theta = 0.5;
    int depth_cap = 60;  // guards coincident points (reference recurses forever)
    ~Tree() { delete root; }
};

// node::getChildContainingPnt (node.h:424-440)
int child_of(const Node* nd, V3 p) {
    int c = 0;
    if (p.x > nd->origin.x) c |= 4;
    if (p.y > nd->origin.y) c |= 2;
    if (p.z > nd->origin.z) c |= 1;
    return c;
}

void add_member(Tree& T, Node* nd, int idx);

// node::createChild + pushMember (node.h:47-76)
void push_member(Tree& T, Node* nd, int idx) {
    int c = child_of(nd, T.parts[idx].pos);
    if (!nd->children[c]) {
        Node* ch = new Node;
        ch->origin = nd->origin;
        ch->origin.x += nd->half.x * (c & 4 ? 0.5 : -0.5);
        ch->origin.y += nd->half.y * (c & 2 ? 0.5 : -0.5);
        ch->origin.z += nd->half.z * (c & 1 ? 0.5 : -0.5);
        ch->half = {0.5 * nd->half.x, 0.5 * nd->half.y, 0.5 * nd->half.z};
        ch->level = nd->level + 1;
        nd->children[c] = ch;
    }
    add_member(T, nd->children[c], idx);
}

// node::addMember (node.h:205-224)
void add_member(Tree& T, Node* nd, int idx) {
    if (nd->leaf()) {
        nd->members.push_back(idx);
        if (nd->members.size() > T.max_members && (int)nd->level < T.depth_cap) {
            std::vector<int> m;
            m.swap(nd->members);
            for (int i : m) push_member(T, nd, i);
        }
    } else {
        push_member(T, nd, idx);
    }
}

// particleFMM::findExpPos / findExpStrength (particleFMM.cpp:120-170)
Part make_exp(const std::vector<const Part*>& ps, const Node* nd) {
    V3 wp = {0, 0, 0};
    double tot = 0;
    for (const Part* p : ps) {
        wp = wp + norm(p->str) * p->pos;
        tot += norm(p->str);
    }
    Part e;
    e.pos = {wp.x / tot, wp.y / tot, wp.z / tot};
    V3 s = {0, 0, 0};
    for (const Part* p : ps) s = s + p->str;
    e.str = s;
    e.radius = norm(nd->half) * 2 * 1.5;
    return e;
}

// particleFMM::computeMultExp (particleFMM.cpp:29-58): every node but the
// root; leaves from members, others from the children's expansions in child
// order.  Post-order gives the same values as the reference's level sweep.
void compute_exp(Tree& T, Node* nd, bool is_root) {
    for (Node* c : nd->children)
        if (c) compute_exp(T, c, false);
    if (is_root) return;
    std::vector<const Part*> ps;
    if (nd->leaf()) {
        for (int i : nd->members) ps.push_back(&T.parts[i]);
    } else {
        for (Node* c : nd->children)
            if (c) ps.push_back(&c->exp);
    }
    nd->exp = make_exp(ps, nd);
    nd->has_exp = true;
}

// node::isFarField(POI) (node.h:395-411)
bool far_field(const Tree& T, const Node* nd, V3 poi) {
    double s = 2 * nd->half.x;
    double d = norm(nd->exp.pos - poi);
    return s / d < T.theta;
}

// node::calcVel(POI) (node.h:236-262)
V3 calc_vel_point(const Tree& T, const Node* nd, V3 poi) {
    V3 v = {0, 0, 0};
    if (nd->leaf()) {
        for (int i : nd->members) v = v + vel_alg(T.parts[i], poi);
    } else if (far_field(T, nd, poi)) {
        v = vel_alg(nd->exp, poi);
    } else {
        for (Node* c : nd->children)
            if (c) v = v + calc_vel_point(T, c, poi);
    }
    return v;
}

// node::calcVel / calcStretch / calcDiff(particle*) (node.h:264-366)
template <class F>
V3 calc_part(const Tree& T, const Node* nd, int t, F kern) {
    V3 v = {0, 0, 0};
    const Part& tp = T.parts[t];
    if (nd->leaf()) {
        for (int i : nd->members)
            if (i != t) v = v + kern(T.parts[i], tp);
    } else if (far_field(T, nd, tp.pos)) {
        v = kern(nd->exp, tp);
    } else {
        for (Node* c : nd->children)
            if (c) v = v + calc_part(T, c, t, kern);
    }
    return v;
}

int count_nodes(const Node* nd, size_t& maxlev) {
    if (nd->level > maxlev) maxlev = nd->level;
    int n = 1;
    for (Node* c : nd->children)
        if (c) n += count_nodes(c, maxlev);
    return n;
}

}  // namespace

extern "C" {

void* or_bh_build(int n, const double* xyz, const double* alpha, const double* radius, int max_members,
                  double theta) {
    Tree* T = new Tree;
    T->max_members = (size_t)max_members;
    T->theta = theta;
    T->parts.resize(n);
    for (int i = 0; i < n; i++)
        T->parts[i] = {{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]},
                       {alpha[3 * i], alpha[3 * i + 1], alpha[3 * i + 2]},
                       radius[i]};
    if (n == 0) return T;
    // octree::setDimensions (octree.h:70-80)
    V3 mn = T->parts[0].pos, mx = T->parts[0].pos;
    for (int i = 0; i < n; i++) {
        V3 p = T->parts[i].pos;
        if (p.x < mn.x) mn.x = p.x;
        if (p.y < mn.y) mn.y = p.y;
        if (p.z < mn.z) mn.z = p.z;
        if (p.x > mx.x) mx.x = p.x;
        if (p.y > mx.y) mx.y = p.y;
        if (p.z > mx.z) mx.z = p.z;
    }
    V3 c = 0.5 * (mn + mx);
    V3 ext = mx - mn;
    double maxD = std::fmax(std::fmax(ext.x, ext.y), ext.z);
    T->root = new Node;
    T->root->origin = c;
    T->root->half = {0.51 * maxD, 0.51 * maxD, 0.51 * maxD};
    T->root->level = 1;  // root built with parent_level 0 (octree.h:181)
    for (int i = 0; i < n; i++) add_member(*T, T->root, i);
    compute_exp(*T, T->root, true);
    return T;
}

void or_bh_destroy(void* h) { delete (Tree*)h; }

/* n_nodes, levels, root origin[3], root half[3] */
void or_bh_info(void* h, int* n_nodes, int* levels, double* origin, double* half) {
    Tree* T = (Tree*)h;
    if (!T->root) {
        *n_nodes = 0;
        *levels = 0;
        return;
    }
    size_t ml = 0;
    *n_nodes = count_nodes(T->root, ml);
    *levels = (int)ml;
    origin[0] = T->root->origin.x;
    origin[1] = T->root->origin.y;
    origin[2] = T->root->origin.z;
    half[0] = T->root->half.x;
    half[1] = T->root->half.y;
    half[2] = T->root->half.z;
}

/* Barnes-Hut sums, particleFMM::barnesHut* (particleFMM.cpp:65-116): the
 * root's children in child order (root excluded).
 * mode 0: velocity at points pts[3*m] (algebraic kernel)
 * mode 1: velocity on the tree's particles (Gaussian), out[3*n]
 * mode 2: vortex stretching on the particles, out[3*n]
 * mode 3: viscous diffusion on the particles, out[3*n] */
void or_bh_eval(void* h, int mode, int m, const double* pts, double* out) {
    Tree* T = (Tree*)h;
    int nt = mode == 0 ? m : (int)T->parts.size();
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < nt; t++) {
        V3 v = {0, 0, 0};
        if (T->root) {
            for (Node* c : T->root->children) {
                if (!c) continue;
                if (mode == 0)
                    v = v + calc_vel_point(*T, c, {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]});
                else if (mode == 1)
                    v = v + calc_part(*T, c, t, vel_gauss);
                else if (mode == 2)
                    v = v + calc_part(*T, c, t, stretch);
                else
                    v = v + calc_part(*T, c, t, diffusion);
            }
        }
        out[3 * t] = v.x;
        out[3 * t + 1] = v.y;
        out[3 * t + 2] = v.z;
    }
}

/* Direct sums with the traversal's leaf kernels (what Barnes-Hut reduces to
 * when no cell is far): mode as in or_bh_eval; mode 4: Gaussian velocity at
 * points (particle::velInfl(POI), the non-accelerated path of cpCaseVP.cpp:611). */
void or_vp_direct(int mode, int n, const double* xyz, const double* alpha, const double* radius, int m,
                  const double* pts, double* out) {
    std::vector<Part> P(n);
    for (int i = 0; i < n; i++)
        P[i] = {{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, {alpha[3 * i], alpha[3 * i + 1], alpha[3 * i + 2]},
                radius[i]};
    int nt = (mode == 0 || mode == 4) ? m : n;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; t++) {
        V3 v = {0, 0, 0};
        for (int j = 0; j < n; j++) {
            if (mode == 0)
                v = v + vel_alg(P[j], {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]});
            else if (mode == 4)
                v = v + vel_gauss_point(P[j], {pts[3 * t], pts[3 * t + 1], pts[3 * t + 2]});
            else if (j != t) {
                if (mode == 1)
                    v = v + vel_gauss(P[j], P[t]);
                else if (mode == 2)
                    v = v + stretch(P[j], P[t]);
                else
                    v = v + diffusion(P[j], P[t]);
            }
        }
        out[3 * t] = v.x;
        out[3 * t + 1] = v.y;
        out[3 * t + 2] = v.z;
    }
}

/* Direct sums of the configured case's kernel (vel_alg_core,
 * stretch_alg_core; one core radius): mode 0 velocity at m points, 1
 * velocity on the n particles, 2 stretching on the n particles (a particle
 * excludes itself). */
void or_vp_direct_core(int mode, double core, int n, const double* xyz, const double* alpha, int m, const double* pts,
                       double* out) {
    int nt = mode == 0 ? m : n;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; t++) {
        V3 v = {0, 0, 0};
        const double* tp = mode == 0 ? pts + 3 * t : xyz + 3 * t;
        V3 tpos = {tp[0], tp[1], tp[2]};
        V3 tstr = mode == 2 ? V3{alpha[3 * t], alpha[3 * t + 1], alpha[3 * t + 2]} : V3{0, 0, 0};
        for (int j = 0; j < n; j++) {
            if (mode != 0 && j == t) continue;
            V3 sp = {xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2]};
            V3 sa = {alpha[3 * j], alpha[3 * j + 1], alpha[3 * j + 2]};
            v = v + (mode == 2 ? stretch_alg_core(sp, sa, tpos, tstr, core) : vel_alg_core(sp, sa, tpos, core));
        }
        out[3 * t] = v.x;
        out[3 * t + 1] = v.y;
        out[3 * t + 2] = v.z;
    }
}

/* One pair of the configured case's kernel: mode 0 velocity, 2 stretching. */
void or_vp_pair_core(int mode, double core, const double* xs, const double* as, const double* xt, const double* at,
                     double* out) {
    V3 sp = {xs[0], xs[1], xs[2]}, sa = {as[0], as[1], as[2]}, tp = {xt[0], xt[1], xt[2]},
       ta = {at[0], at[1], at[2]};
    V3 v = mode == 2 ? stretch_alg_core(sp, sa, tp, ta, core) : vel_alg_core(sp, sa, tp, core);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

/* One pair: kernel `mode` of source (xs, as, rs) at target (xt, at, rt);
 * mode 0 ignores the target's strength and radius. */
void or_vp_pair(int mode, const double* xs, const double* as, double rs, const double* xt, const double* at,
                double rt, double* out) {
    Part s = {{xs[0], xs[1], xs[2]}, {as[0], as[1], as[2]}, rs};
    Part t = {{xt[0], xt[1], xt[2]}, {at[0], at[1], at[2]}, rt};
    V3 v;
    if (mode == 0)
        v = vel_alg(s, t.pos);
    else if (mode == 1)
        v = vel_gauss(s, t);
    else if (mode == 2)
        v = stretch(s, t);
    else if (mode == 3)
        v = diffusion(s, t);
    else
        v = vel_gauss_point(s, t.pos);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

/* Velocity induced at points by body panels (source sigma, doublet mu:
 * bodyPanel::panelV, bodyPanel.cpp:122-130) and wake panels (doublet mu_w:
 * wakePanel::panelV, wakePanel.cpp:89-92).  Panels are given by vertex
 * arrays verts[12*k] with nv[k] vertices. */
void or_panels_to_points(int nb, const int* nvb, const double* vb, const double* sigma, const double* mu, int nw,
                         const int* nvw, const double* vw, const double* mu_w, int m, const double* pts,
                         double* out) {
#pragma omp parallel for schedule(static)
    for (int t = 0; t < m; t++) {
        double acc[3] = {0, 0, 0};
        for (int j = 0; j < nb; j++) {
            double vs[3], vd[3];
            or_panel_v_inf(nvb[j], vb + 12 * j, pts + 3 * t, vs, vd);
            for (int k = 0; k < 3; k++) acc[k] += sigma[j] * vs[k] + mu[j] * vd[k];
        }
        for (int j = 0; j < nw; j++) {
            double vd[3];
            or_dub_v_inf(nvw[j], vw + 12 * j, pts + 3 * t, vd);
            for (int k = 0; k < 3; k++) acc[k] += mu_w[j] * vd[k];
        }
        out[3 * t] = acc[0];
        out[3 * t + 1] = acc[1];
        out[3 * t + 2] = acc[2];
    }
}

}  // extern "C"
