// Oracle: lagged, C1-smoothed friction of the IPC contact model.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// friction_basis restates Compute_Friction_Basis (FEM/FRICTION.h:17-112): the friction
// stencils are the constraints without a mollifier (FRICTION.h:37-41), each with its
// closest-point coordinates (FEM/FRICTION_UTILS.h:55-66 point-triangle, :123-144
// edge-edge, :204-212 point-edge), a 3x2 tangent basis (:43-53, :111-121, :192-202,
// :242-259) and the normal force lambda = -b'(d^2 - xi^2) 2 d (FRICTION.h:104-107).
// friction_energy / _grad / _hess restate Compute_Friction_Potential / _Gradient /
// _Hessian (FRICTION.h:173-243, 245-394, 396-661) with the smoothing f0, f1/x, f2 of
// FRICTION_UTILS.h:9-37 and eps_v h = sqrt(epsv2) h (IMPLICIT_EULER.h:436).
//
// Formulation used here: every stencil's relative tangential displacement is
// u = B^T sum_a c_a (x_a - xn_a) with per-vertex weights c_a (point-triangle
// (1, b1 + b2 - 1, -b1, -b2), edge-edge (1 - g1, g1, g2 - 1, -g2), point-edge
// (1, e - 1, -e), point-point (1, -1)); the gradient block of vertex a is c_a B w and
// the Hessian block (a, b) is c_a c_b B M B^T with the 2x2 matrix M of the reference.
// Parity pins (tests/test_oracle_friction.py): gradient = central difference of the
// energy, Hessian = central difference of the gradient (the 2x2 matrix is the exact
// Hessian of the smoothed norm, so the projection never acts away from rounding),
// closest points orthogonality, normal force = norm of the barrier gradient on the
// point, and the momentum balance of a sheet sliding on the ground.
#include "oracle.h"
#include <algorithm>

namespace orc {

static inline V3 X3(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
static inline V3 normalized(V3 a) { return (1.0 / std::sqrt(sq(a))) * a; }

bool friction_by_component(const Scene& S) {
    const size_t nc = S.P.compNodeRange.size();
    return nc > 0 && S.P.muComp.size() == nc * nc;
}
// getCompI (FRICTION.h:114-124): first component whose range end exceeds the vertex
static int comp_of(const Scene& S, int v) {
    const std::vector<int>& r = S.P.compNodeRange;
    for (size_t i = 0; i < r.size(); ++i)
        if (v < r[i]) return (int)i;
    return (int)r.size() - 1;
}

// vertex ids, count and weights of a friction stencil (info: cp0, cp1, basis cols, lambda)
static int fric_stencil(const Constraint& cc, const double* info, int* vid, double* coef, double* mult) {
    const int* c = cc.c;
    *mult = 1;
    if (c[0] >= 0) {   // edge-edge
        for (int i = 0; i < 4; ++i) vid[i] = c[i];
        const double g1 = info[0], g2 = info[1];
        coef[0] = 1.0 - g1; coef[1] = g1; coef[2] = g2 - 1.0; coef[3] = -g2;
        return 4;
    }
    vid[0] = -c[0] - 1;
    vid[1] = c[1];
    if (c[2] < 0) {    // point-point
        *mult = -c[3];
        coef[0] = 1; coef[1] = -1;
        return 2;
    }
    vid[2] = c[2];
    if (c[3] < 0) {    // point-edge
        *mult = -c[3];
        const double e = info[0];
        coef[0] = 1; coef[1] = e - 1.0; coef[2] = -e;
        return 3;
    }
    vid[3] = c[3];     // point-triangle
    const double b1 = info[0], b2 = info[1];
    coef[0] = 1; coef[1] = -1.0 + b1 + b2; coef[2] = -b1; coef[3] = -b2;
    return 4;
}

void friction_basis(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa, FricSet& F) {
    const double xi2 = S.P.xi * S.P.xi;
    const double shat = S.dHat2 + 2 * std::sqrt(S.dHat2) * S.P.xi;
    F.C.clear();
    F.info.clear();
    for (const Constraint& cc : C) {
        const int* c = cc.c;
        if (c[0] >= 0 && (c[2] < 0 || c[3] < 0)) continue;   // mollified: no friction
        double inf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        V3 b0, b1;
        double d2;
        if (c[0] >= 0) {
            V3 a0 = X3(x, c[0]), a1 = X3(x, c[1]), q0 = X3(x, c[2]), q1 = X3(x, c[3]);
            // closest points of the two lines: minimise |a0 + g1 (a1 - a0) - q0 - g2 (q1 - q0)|^2
            V3 ea = a1 - a0, eb = q1 - q0, w = a0 - q0;
            double m00 = sq(ea), m01 = -dot(eb, ea), m11 = sq(eb);
            double r0 = -dot(w, ea), r1 = dot(w, eb);
            double det = m00 * m11 - m01 * m01;
            inf[0] = (m11 * r0 - m01 * r1) / det;
            inf[1] = (m00 * r1 - m01 * r0) / det;
            b0 = normalized(ea);
            b1 = normalized(cross(cross(ea, eb), ea));
            d2 = dist2_ee(a0, a1, q0, q1);
        } else {
            V3 p = X3(x, -c[0] - 1);
            if (c[2] < 0) {
                V3 p1 = X3(x, c[1]);
                V3 d = p1 - p;
                V3 xc = cross(V3{1, 0, 0}, d), yc = cross(V3{0, 1, 0}, d);
                V3 t = sq(xc) > sq(yc) ? xc : yc;
                b0 = normalized(t);
                b1 = normalized(cross(d, t));
                d2 = dist2_pp(p, p1);
            } else if (c[3] < 0) {
                V3 e0 = X3(x, c[1]), e1 = X3(x, c[2]);
                V3 e = e1 - e0;
                inf[0] = dot(p - e0, e) / sq(e);
                b0 = normalized(e);
                b1 = normalized(cross(e, p - e0));
                d2 = dist2_pe(p, e0, e1);
            } else {
                V3 t0 = X3(x, c[1]), t1 = X3(x, c[2]), t2 = X3(x, c[3]);
                V3 e1 = t1 - t0, e2 = t2 - t0, w = p - t0;
                double m00 = sq(e1), m01 = dot(e1, e2), m11 = sq(e2);
                double r0 = dot(e1, w), r1 = dot(e2, w);
                double det = m00 * m11 - m01 * m01;
                inf[0] = (m11 * r0 - m01 * r1) / det;
                inf[1] = (m00 * r1 - m01 * r0) / det;
                b0 = normalized(e1);
                b1 = normalized(cross(cross(e1, e2), e1));
                d2 = dist2_pt(p, t0, t1, t2);
            }
        }
        inf[2] = b0.x; inf[3] = b0.y; inf[4] = b0.z;
        inf[5] = b1.x; inf[6] = b1.y; inf[7] = b1.z;
        inf[8] = -barrier_g(d2 - xi2, shat, kappa) * 2 * std::sqrt(d2);   // stencil weight 1
        if (friction_by_component(S)) {   // Compute_Friction_Coef, FRICTION.h:127-171
            const int a = c[0] >= 0 ? c[0] : -c[0] - 1, b = c[0] >= 0 ? c[2] : c[1];
            const int nc = (int)S.P.compNodeRange.size();
            inf[8] *= S.P.muComp[comp_of(S, a) + comp_of(S, b) * nc];
        }
        F.C.push_back(cc);
        F.info.insert(F.info.end(), inf, inf + 9);
    }
}

// per-stencil evaluation: u, |u|^2, weights; returns nv
struct FricLocal {
    int nv, vid[4];
    double coef[4], mult, u[2], q;
    V3 b0, b1;
};
static void fric_local(const double* x, const double* xn, const Constraint& cc, const double* inf, FricLocal& L) {
    L.nv = fric_stencil(cc, inf, L.vid, L.coef, &L.mult);
    V3 rel{0, 0, 0};
    for (int a = 0; a < L.nv; ++a) rel = rel + L.coef[a] * (X3(x, L.vid[a]) - X3(xn, L.vid[a]));
    L.b0 = {inf[2], inf[3], inf[4]};
    L.b1 = {inf[5], inf[6], inf[7]};
    L.u[0] = dot(L.b0, rel);
    L.u[1] = dot(L.b1, rel);
    L.q = L.u[0] * L.u[0] + L.u[1] * L.u[1];
}

// FRICTION_UTILS.h:9-37
static double f0_smooth(double q, double ev) {
    if (q >= ev * ev) return std::sqrt(q);
    return q * (-std::sqrt(q) / 3.0 + ev) / (ev * ev) + ev / 3.0;
}
static double f1_div_norm(double q, double ev) {
    if (q >= ev * ev) return 1.0 / std::sqrt(q);
    return (-std::sqrt(q) + 2.0 * ev) / (ev * ev);
}

static double fric_eps(const Scene& S) { return std::sqrt(S.P.epsv2 * S.P.dt * S.P.dt); }
// the coefficient multiplying every stencil: mu, or 1 when the coefficients come per component
static double fric_mu(const Scene& S) { return friction_by_component(S) ? 1.0 : S.P.fricMu; }

double friction_energy(const Scene& S, const double* x, const double* xn, const FricSet& F) {
    const double ev = fric_eps(S);
    double E = 0;
    for (size_t k = 0; k < F.C.size(); ++k) {
        FricLocal L;
        const double* inf = &F.info[9 * k];
        fric_local(x, xn, F.C[k], inf, L);
        E += L.mult * inf[8] * f0_smooth(L.q, ev);
    }
    return fric_mu(S) * E;
}

void friction_grad(const Scene& S, const double* x, const double* xn, const FricSet& F, double* g) {
    const double ev = fric_eps(S);
    for (size_t k = 0; k < F.C.size(); ++k) {
        FricLocal L;
        const double* inf = &F.info[9 * k];
        fric_local(x, xn, F.C[k], inf, L);
        const double s = f1_div_norm(L.q, ev) * fric_mu(S) * L.mult * inf[8];
        V3 w = (s * L.u[0]) * L.b0 + (s * L.u[1]) * L.b1;
        for (int a = 0; a < L.nv; ++a) {
            g[3 * L.vid[a]] += L.coef[a] * w.x;
            g[3 * L.vid[a] + 1] += L.coef[a] * w.y;
            g[3 * L.vid[a] + 2] += L.coef[a] * w.z;
        }
    }
}

// 2x2 symmetric PSD clamp (Math/UTILS.h:10-27 for n = 2), closed-form eigenpairs
static void make_pd2(double* M) {
    const double a = M[0], b = M[1], d = M[3];
    const double h = 0.5 * (a - d), r = std::sqrt(h * h + b * b);
    const double l0 = 0.5 * (a + d) - r, l1 = 0.5 * (a + d) + r;
    if (l0 >= 0) return;
    // eigenvector of l1
    double vx, vy;
    if (r == 0) { vx = 1; vy = 0; }
    else if (h >= 0) { vx = h + r; vy = b; }
    else { vx = b; vy = r - h; }
    const double n = std::sqrt(vx * vx + vy * vy);
    vx /= n; vy /= n;
    const double l = std::max(l1, 0.0);
    M[0] = l * vx * vx; M[1] = M[2] = l * vx * vy; M[3] = l * vy * vy;
}

void friction_hess(const Scene& S, const double* x, const double* xn, const FricSet& F, std::vector<Block>& out) {
    const double ev = fric_eps(S);
    for (size_t k = 0; k < F.C.size(); ++k) {
        FricLocal L;
        const double* inf = &F.info[9 * k];
        fric_local(x, xn, F.C[k], inf, L);
        const double sc = fric_mu(S) * L.mult * inf[8];
        const double f1d = f1_div_norm(L.q, ev);
        double M[4];
        if (L.q >= ev * ev) {
            // sliding: (f1 / |u|) (I - u u^T / |u|^2) = f1/|u| / |u|^2 * ubar ubar^T
            const double c = sc * f1d / L.q, ub0 = -L.u[1], ub1 = L.u[0];
            M[0] = c * ub0 * ub0; M[1] = M[2] = c * ub0 * ub1; M[3] = c * ub1 * ub1;
        } else if (L.q == 0) {
            M[0] = M[3] = sc * f1d; M[1] = M[2] = 0;
        } else {
            const double un = std::sqrt(L.q), f2 = -1.0 / (ev * ev);
            M[0] = f2 / un * L.u[0] * L.u[0] + f1d;
            M[1] = M[2] = f2 / un * L.u[0] * L.u[1];
            M[3] = f2 / un * L.u[1] * L.u[1] + f1d;
            make_pd2(M);
            for (double& m : M) m *= sc;
        }
        // K = B M B^T
        const double B[3][2] = {{L.b0.x, L.b1.x}, {L.b0.y, L.b1.y}, {L.b0.z, L.b1.z}};
        double K[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                K[i * 3 + j] = B[i][0] * (M[0] * B[j][0] + M[1] * B[j][1]) + B[i][1] * (M[2] * B[j][0] + M[3] * B[j][1]);
        for (int a = 0; a < L.nv; ++a)
            for (int b = 0; b < L.nv; ++b) {
                Block blk;
                blk.r = L.vid[a];
                blk.c = L.vid[b];
                for (int e = 0; e < 9; ++e) blk.b[e] = L.coef[a] * L.coef[b] * K[e];
                out.push_back(blk);
            }
    }
}

}  // namespace orc
