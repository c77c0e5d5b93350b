// Oracle: codimensional primitives -- discrete rods (segment springs and hinge
// bending) and particles.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates FEM/Shell/Rod/DISCRETE_ROD.h:87-165 (Initialize_Discrete_Rod: per-segment
// E, rest length, thickness; lumped mass l pi t^2/4 rho / 2 per end; one rod hinge
// per pair of segments meeting at a node, in std::map node order), DISCRETE_PARTICLE.h:
// 52-75 (particle mass 4/3 pi t^3/8 rho), FEM/Shell/Rod/ROD_STRETCHING.h:10-107
// (spring energy h^2 pi t^2/4 * lbar * E/2 * (l/lbar - 1)^2 and its derivatives,
// MASS_SPRING_DERIVATIVES.h) and FEM/Shell/Rod/ROD_BENDING.h:10-117 (bending energy
// h^2 * E t^4 pi/64 * |kappa|^2 / lsum, kappa = 2 e0 x e1 / (|e0||e1| + e0.e1)).
//
// The derivatives are derived here, not transcribed: the spring's from l = |x0 - x1|,
// and the hinge's from the identity |kappa|^2 = 4 (ab - c) / (ab + c) with a = |e0|,
// b = |e1|, c = e0.e1 (Lagrange: |e0 x e1|^2 = a^2 b^2 - c^2), i.e. the chain rule
// through u = ab and v = c.  The energy itself is evaluated with the reference's
// cross-product form.  Pins (tests/test_oracle_codim.py): torch autograd of the
// reference's literal formulas and central differences.
#include "oracle.h"
#include <map>

namespace orc {

static inline V3 X3(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

double rod_spring_energy(V3 x0, V3 x1, double Er, double lbar) {
    double l = std::sqrt(sq(x0 - x1));
    double r = l / lbar - 1;
    return lbar * Er / 2 * r * r;
}
void rod_spring_grad(V3 x0, V3 x1, double Er, double lbar, double* g) {
    V3 d = x0 - x1;
    double l = std::sqrt(sq(d));
    double s = Er / lbar * (l - lbar) / l;
    g[0] = s * d.x; g[1] = s * d.y; g[2] = s * d.z;
    g[3] = -g[0]; g[4] = -g[1]; g[5] = -g[2];
}
void rod_spring_hess(V3 x0, V3 x1, double Er, double lbar, double* H) {
    V3 d = x0 - x1;
    double l = std::sqrt(sq(d));
    double n[3] = {d.x / l, d.y / l, d.z / l};
    double k = Er / lbar, perp = Er / lbar * (l - lbar) / l;
    double K[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K[3 * i + j] = (k - perp) * n[i] * n[j] + (i == j ? perp : 0.0);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            double s = (a == b) ? 1.0 : -1.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) H[(3 * a + i) * 6 + 3 * b + j] = s * K[3 * i + j];
        }
}

double rod_kappa2(V3 x0, V3 x1, V3 x2) {
    V3 e0 = x1 - x0, e1 = x2 - x1;
    double den = std::sqrt(sq(e0) * sq(e1)) + dot(e0, e1);
    V3 k = (2 / den) * cross(e0, e1);
    return sq(k);
}

// f(u, v) = 4 (u - v) / (u + v) and its partial derivatives
struct RodF { double fu, fv, fuu, fvv, fuv; };
static RodF rod_f(double u, double v) {
    double s = u + v, s2 = s * s, s3 = s2 * s;
    return {8 * v / s2, -8 * u / s2, -16 * v / s3, 16 * u / s3, 8 * (u - v) / s3};
}

void rod_kappa2_grad(V3 x0, V3 x1, V3 x2, double* g) {
    V3 e0 = x1 - x0, e1 = x2 - x1;
    double a = std::sqrt(sq(e0)), b = std::sqrt(sq(e1)), c = dot(e0, e1);
    RodF F = rod_f(a * b, c);
    V3 g0 = (F.fu * b / a) * e0 + F.fv * e1;   // d f / d e0
    V3 g1 = (F.fu * a / b) * e1 + F.fv * e0;   // d f / d e1
    V3 r0 = -1.0 * g0, r1 = g0 - g1, r2 = g1;
    double o[9] = {r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r2.x, r2.y, r2.z};
    for (int i = 0; i < 9; ++i) g[i] = o[i];
}

void rod_kappa2_hess(V3 x0, V3 x1, V3 x2, double* H) {
    V3 e0v = x1 - x0, e1v = x2 - x1;
    double e0[3] = {e0v.x, e0v.y, e0v.z}, e1[3] = {e1v.x, e1v.y, e1v.z};
    double a = std::sqrt(sq(e0v)), b = std::sqrt(sq(e1v)), c = dot(e0v, e1v);
    RodF F = rod_f(a * b, c);
    // gradients of u = ab and v = c over (e0, e1)
    double gu[6], gv[6];
    for (int i = 0; i < 3; ++i) {
        gu[i] = b / a * e0[i]; gu[3 + i] = a / b * e1[i];
        gv[i] = e1[i]; gv[3 + i] = e0[i];
    }
    double He[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            // second derivatives of u and v
            int bi = i / 3, bj = j / 3, k = i % 3, l = j % 3;
            double d2u, d2v;
            if (bi == 0 && bj == 0) d2u = b * ((k == l ? 1.0 : 0.0) / a - e0[k] * e0[l] / (a * a * a));
            else if (bi == 1 && bj == 1) d2u = a * ((k == l ? 1.0 : 0.0) / b - e1[k] * e1[l] / (b * b * b));
            else if (bi == 0) d2u = e0[k] / a * e1[l] / b;
            else d2u = e1[k] / b * e0[l] / a;
            d2v = (bi != bj && k == l) ? 1.0 : 0.0;
            He[6 * i + j] = F.fu * d2u + F.fv * d2v + F.fuu * gu[i] * gu[j] + F.fvv * gv[i] * gv[j] +
                            F.fuv * (gu[i] * gv[j] + gv[i] * gu[j]);
        }
    // x-space: e0 = x1 - x0, e1 = x2 - x1
    auto blk = [&](int p, int q, int k, int l) { return He[6 * (3 * p + k) + 3 * q + l]; };
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
            double H00 = blk(0, 0, k, l), H01 = blk(0, 1, k, l), H10 = blk(1, 0, k, l), H11 = blk(1, 1, k, l);
            double X[3][3] = {{H00, -H00 + H01, -H01},
                              {-H00 + H10, H00 - H01 - H10 + H11, H01 - H11},
                              {-H10, H10 - H11, H11}};
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) H[9 * (3 * p + k) + 3 * q + l] = X[p][q];
        }
}

// ---------------------------------------------------------------- set-up
void codim_init(Scene& S) {
    const Params& P = S.P;
    const int nR = (int)S.rod.size() / 2;
    // DISCRETE_ROD.h:106-134: per-segment data, lumped mass; neighbour lists per node
    S.rodInfo.assign(3 * (size_t)nR, 0.0);
    std::map<int, std::vector<int>> nbr;
    for (int r = 0; r < nR; ++r) {
        int a = S.rod[2 * r], b = S.rod[2 * r + 1];
        nbr[a].push_back(b);
        nbr[b].push_back(a);
        double l = std::sqrt(sq(X3(S.X0.data(), a) - X3(S.X0.data(), b)));
        S.rodInfo[3 * r] = P.rodE;
        S.rodInfo[3 * r + 1] = l;
        S.rodInfo[3 * r + 2] = P.rodThick;
        double mp = l * M_PI * P.rodThick * P.rodThick / 4 * P.rodRho / 2;
        S.mass[a] += mp;
        S.mass[b] += mp;
    }
    // DISCRETE_ROD.h:136-154: a hinge per pair of segments at a node
    S.rodHinge.clear(); S.rodHingeInfo.clear();
    for (auto& kv : nbr) {
        const std::vector<int>& n = kv.second;
        if (n.size() < 2) continue;
        for (size_t i = 0; i < n.size(); ++i)
            for (size_t j = i + 1; j < n.size(); ++j) {
                S.rodHinge.push_back(n[i]); S.rodHinge.push_back(kv.first); S.rodHinge.push_back(n[j]);
                V3 p0 = X3(S.X0.data(), n[i]), p1 = X3(S.X0.data(), kv.first), p2 = X3(S.X0.data(), n[j]);
                S.rodHingeInfo.push_back(P.rodE * P.rodBendMult);
                S.rodHingeInfo.push_back(std::sqrt(sq(p1 - p0)) + std::sqrt(sq(p2 - p1)));
                S.rodHingeInfo.push_back(P.rodThick);
            }
    }
    // DISCRETE_PARTICLE.h:60-69
    const double pm = 4.0 / 3.0 * M_PI * P.partThick * P.partThick * P.partThick / 8 * P.partRho;
    for (int v : S.particle) S.mass[v] += pm;
    // boundary primitives, IMPLICIT_EULER.h:250-277: rod segments after the surface
    // edges, rod nodes (ascending) then particles after the surface nodes
    for (int r = 0; r < nR; ++r) { S.bEdge.push_back(S.rod[2 * r]); S.bEdge.push_back(S.rod[2 * r + 1]); }
    S.codimBN0 = (int)S.bNode.size();
    for (auto& kv : nbr) S.bNode.push_back(kv.first);
    S.codimBN1 = (int)S.bNode.size();
    for (int v : S.particle) S.bNode.push_back(v);
}

// ---------------------------------------------------------------- energies
// weights of ROD_STRETCHING.h:22-24, 50, 99 and ROD_BENDING.h:30, 55-56, 106-107
static inline double spring_w(const Scene& S, int r) {
    double t = S.rodInfo[3 * r + 2];
    return S.P.dt * S.P.dt * M_PI * t * t / 4;
}
static inline double bend_w(const Scene& S, int h) {
    const double* I = &S.rodHingeInfo[3 * h];
    return S.P.dt * S.P.dt / I[1] * I[0] * std::pow(I[2], 4) * M_PI / 64;
}

void codim_energy(const Scene& S, const double* x, double* stretch, double* bend) {
    const int nR = (int)S.rod.size() / 2, nRH = (int)S.rodHinge.size() / 3;
    for (int r = 0; r < nR; ++r)
        *stretch += spring_w(S, r) * rod_spring_energy(X3(x, S.rod[2 * r]), X3(x, S.rod[2 * r + 1]), S.rodInfo[3 * r],
                                                       S.rodInfo[3 * r + 1]);
    for (int h = 0; h < nRH; ++h) {
        const int* v = &S.rodHinge[3 * h];
        *bend += bend_w(S, h) * rod_kappa2(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]));
    }
}

void codim_grad(const Scene& S, const double* x, double* g) {
    const int nR = (int)S.rod.size() / 2, nRH = (int)S.rodHinge.size() / 3;
    for (int r = 0; r < nR; ++r) {
        int v[2] = {S.rod[2 * r], S.rod[2 * r + 1]};
        double gl[6];
        rod_spring_grad(X3(x, v[0]), X3(x, v[1]), S.rodInfo[3 * r], S.rodInfo[3 * r + 1], gl);
        double w = spring_w(S, r);
        for (int a = 0; a < 2; ++a)
            for (int k = 0; k < 3; ++k) g[3 * v[a] + k] += w * gl[3 * a + k];
    }
    for (int h = 0; h < nRH; ++h) {
        const int* v = &S.rodHinge[3 * h];
        double gl[9];
        rod_kappa2_grad(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), gl);
        double w = bend_w(S, h);
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) g[3 * v[a] + k] += w * gl[3 * a + k];
    }
}

// 36 / 81 triplets per segment / hinge (ROD_STRETCHING.h:66-106, ROD_BENDING.h:82-116)
void codim_hess(const Scene& S, const double* x, std::vector<Block>& out) {
    const int nR = (int)S.rod.size() / 2, nRH = (int)S.rodHinge.size() / 3;
    const long base = (long)out.size();
    out.resize(base + 4l * nR + 9l * nRH);
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (int r = 0; r < nR; ++r) {
        int v[2] = {S.rod[2 * r], S.rod[2 * r + 1]};
        double H[36];
        rod_spring_hess(X3(x, v[0]), X3(x, v[1]), S.rodInfo[3 * r], S.rodInfo[3 * r + 1], H);
        if (S.P.projectPD) make_pd(6, H);
        double w = spring_w(S, r);
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                Block& B = out[base + 4l * r + 2 * a + b];
                B.r = v[a]; B.c = v[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) B.b[3 * k + l] = w * H[(3 * a + k) * 6 + 3 * b + l];
            }
    }
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (int h = 0; h < nRH; ++h) {
        const int* v = &S.rodHinge[3 * h];
        double H[81];
        rod_kappa2_hess(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), H);
        if (S.P.projectPD) make_pd(9, H);
        double w = bend_w(S, h);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                Block& B = out[base + 4l * nR + 9l * h + 3 * a + b];
                B.r = v[a]; B.c = v[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) B.b[3 * k + l] = w * H[(3 * a + k) * 9 + 3 * b + l];
            }
    }
}

}  // namespace orc
