// Oracle: per-stencil geometry -- squared distances, region classifiers, their
// derivatives, the edge-edge mollifier, the log barrier, the symmetric eigensolver
// behind the PSD projection, the Neo-Hookean membrane and the hinge dihedral angle.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Parity pins (tests/test_oracle_geom.py):
//  * distances and region types: closed forms + brute-force minimisation over the
//    barycentric / edge parameters on random and degenerate stencils;
//  * every gradient/Hessian: torch float64 autograd of the plain energy formula and
//    central finite differences;
//  * sym_eig/make_pd: numpy.linalg.eigh fixtures (tests/golden) and PSD invariants;
//  * barrier derivatives: closed forms against autograd.
// Bit-level behaviour of Eigen's 2x2 LDLT inside Point_Triangle_Distance_Type and of
// Eigen's SelfAdjointEigenSolver is "parity unpinned": no reference test fixes it.
#include "oracle.h"
#include <algorithm>
#include <cstring>

namespace orc {

// ------------------------------------------------------------ distances
// POINT_POINT.h:12-17
double dist2_pp(V3 a, V3 b) { return sq(a - b); }
// POINT_EDGE.h:12-25 (3D branch)
double dist2_pe(V3 p, V3 e0, V3 e1) { return sq(cross(e0 - p, e1 - p)) / sq(e1 - e0); }
// POINT_TRIANGLE.h:12-22
double dist2_pt(V3 p, V3 t0, V3 t1, V3 t2) {
    V3 b = cross(t1 - t0, t2 - t0);
    double aTb = dot(p - t0, b);
    return aTb * aTb / sq(b);
}
// EDGE_EDGE.h:12-22
double dist2_ee(V3 a0, V3 a1, V3 b0, V3 b1) {
    V3 b = cross(a1 - a0, b1 - b0);
    double aTb = dot(b0 - a0, b);
    return aTb * aTb / sq(b);
}

// DISTANCE_TYPE.h:12-28
int pe_type(V3 p, V3 e0, V3 e1) {
    V3 e = e1 - e0;
    double ratio = dot(e, p - e0) / sq(e);
    if (ratio < 0) return 0;
    if (ratio > 1) return 1;
    return 2;
}

// 2x2 symmetric solve by LDL^T with diagonal pivoting (the operation order is shared
// with the HIP kernel so that region predicates round identically).
static void ldlt2_solve(double a, double b, double c, double r0, double r1, double* x0, double* x1) {
    bool swp = std::fabs(c) > std::fabs(a);
    if (swp) { double t = a; a = c; c = t; t = r0; r0 = r1; r1 = t; }
    double l = b / a;
    double d1 = c - l * b;
    double y1 = r1 - l * r0;
    double z0 = r0 / a;
    double z1 = y1 / d1;
    double s0 = z0 - l * z1;
    if (swp) { *x0 = z1; *x1 = s0; } else { *x0 = s0; *x1 = z1; }
}

// param of p - o in the basis (r0, r0 x n): DISTANCE_TYPE.h:40-47
static void pt_edge_param(V3 o, V3 r0, V3 nv, V3 p, double* s, double* t) {
    V3 r1 = cross(r0, nv);
    V3 w = p - o;
    ldlt2_solve(dot(r0, r0), dot(r0, r1), dot(r1, r1), dot(r0, w), dot(r1, w), s, t);
}

// DISTANCE_TYPE.h:31-82
int pt_type(V3 p, V3 t0, V3 t1, V3 t2) {
    V3 nv = cross(t1 - t0, t2 - t0);
    double s0, h0, s1, h1, s2, h2;
    pt_edge_param(t0, t1 - t0, nv, p, &s0, &h0);
    if (s0 > 0.0 && s0 < 1.0 && h0 >= 0.0) return 3;
    pt_edge_param(t1, t2 - t1, nv, p, &s1, &h1);
    if (s1 > 0.0 && s1 < 1.0 && h1 >= 0.0) return 4;
    pt_edge_param(t2, t0 - t2, nv, p, &s2, &h2);
    if (s2 > 0.0 && s2 < 1.0 && h2 >= 0.0) return 5;
    if (s0 <= 0.0 && s2 >= 1.0) return 0;
    if (s1 <= 0.0 && s0 >= 1.0) return 1;
    if (s2 <= 0.0 && s1 >= 1.0) return 2;
    return 6;
}

// DISTANCE_TYPE.h:86-164
int ee_type(V3 a0, V3 a1, V3 b0, V3 b1) {
    V3 u = a1 - a0, v = b1 - b0, w = a0 - b0;
    double a = sq(u), b = dot(u, v), c = sq(v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b;
    double tD = D;
    double sN, tN;
    int defaultCase = 8;
    sN = b * e - c * d;
    if (sN <= 0.0) {
        tN = e; tD = c; defaultCase = 2;
    } else if (sN >= D) {
        tN = e + b; tD = c; defaultCase = 5;
    } else {
        tN = a * e - b * d;
        V3 uv = cross(u, v);
        if (tN > 0.0 && tN < tD && (dot(uv, w) == 0.0 || sq(uv) < 1.0e-20 * a * c)) {
            if (sN < D / 2) { tN = e; tD = c; defaultCase = 2; }
            else { tN = e + b; tD = c; defaultCase = 5; }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return 0;
        else if (-d >= a) return 3;
        else return 6;
    } else if (tN >= tD) {
        if ((-d + b) <= 0.0) return 1;
        else if ((-d + b) >= a) return 4;
        else return 7;
    }
    return defaultCase;
}

// DISTANCE_UNCLASSIFIED.h:16-62
double dist2_pt_unclassified(V3 p, V3 t0, V3 t1, V3 t2) {
    switch (pt_type(p, t0, t1, t2)) {
        case 0: return dist2_pp(p, t0);
        case 1: return dist2_pp(p, t1);
        case 2: return dist2_pp(p, t2);
        case 3: return dist2_pe(p, t0, t1);
        case 4: return dist2_pe(p, t1, t2);
        case 5: return dist2_pe(p, t2, t0);
        default: return dist2_pt(p, t0, t1, t2);
    }
}
// DISTANCE_UNCLASSIFIED.h:65-121
double dist2_ee_unclassified(V3 a0, V3 a1, V3 b0, V3 b1) {
    switch (ee_type(a0, a1, b0, b1)) {
        case 0: return dist2_pp(a0, b0);
        case 1: return dist2_pp(a0, b1);
        case 2: return dist2_pe(a0, b0, b1);
        case 3: return dist2_pp(a1, b0);
        case 4: return dist2_pp(a1, b1);
        case 5: return dist2_pe(a1, b0, b1);
        case 6: return dist2_pe(b0, a0, a1);
        case 7: return dist2_pe(b1, a0, a1);
        default: return dist2_ee(a0, a1, b0, b1);
    }
}
// DISTANCE_UNCLASSIFIED.h:125-148
double dist2_pe_unclassified(V3 p, V3 e0, V3 e1) {
    V3 v = e1 - e0, w = p - e0;
    double c1 = dot(w, v);
    if (c1 <= 0.0) return dist2_pp(p, e0);
    double c2 = sq(v);
    if (c2 <= c1) return dist2_pp(p, e1);
    double b = c1 / c2;
    return dist2_pp(p, e0 + b * v);
}

// ------------------------------------------------------------ small matrix helpers
static void cpm(V3 v, double* M) {  // [v]x
    M[0] = 0; M[1] = -v.z; M[2] = v.y;
    M[3] = v.z; M[4] = 0; M[5] = -v.x;
    M[6] = -v.y; M[7] = v.x; M[8] = 0;
}
static void mm3(const double* A, const double* B, double* C) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[i * 3 + j] = A[i * 3 + 0] * B[0 * 3 + j] + A[i * 3 + 1] * B[1 * 3 + j] + A[i * 3 + 2] * B[2 * 3 + j];
}
// write 3x3 block M (scaled) into n x n matrix H at block (bi, bj)
static void put_block(double* H, int n, int bi, int bj, const double* M, double s) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) H[(bi * 3 + i) * n + bj * 3 + j] += s * M[i * 3 + j];
}
static void put_vec(double* g, int bi, V3 v) { g[bi * 3] = v.x; g[bi * 3 + 1] = v.y; g[bi * 3 + 2] = v.z; }

// map derivatives from "difference" variables to vertex variables:
// vertex k depends on variable slots coef[k][*] with sign; nv variables, nx vertices
struct Lin { int nx; int cnt[4]; int slot[4][3]; double sgn[4][3]; };
static void map_grad(const Lin& L, int nvar, const double* gv, double* gx) {
    (void)nvar;
    for (int k = 0; k < L.nx; ++k)
        for (int d = 0; d < 3; ++d) {
            double s = 0;
            for (int t = 0; t < L.cnt[k]; ++t) s += L.sgn[k][t] * gv[L.slot[k][t] * 3 + d];
            gx[k * 3 + d] = s;
        }
}
static void map_hess(const Lin& L, int nvar, const double* Hv, double* Hx) {
    int nv = nvar * 3, nx = L.nx * 3;
    for (int i = 0; i < nx * nx; ++i) Hx[i] = 0;
    for (int ki = 0; ki < L.nx; ++ki)
        for (int kj = 0; kj < L.nx; ++kj)
            for (int ti = 0; ti < L.cnt[ki]; ++ti)
                for (int tj = 0; tj < L.cnt[kj]; ++tj) {
                    double s = L.sgn[ki][ti] * L.sgn[kj][tj];
                    int si = L.slot[ki][ti], sj = L.slot[kj][tj];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            Hx[(ki * 3 + a) * nx + kj * 3 + b] += s * Hv[(si * 3 + a) * nv + sj * 3 + b];
                }
}

// second-order data of q = |a x b|^2 over (a, b): gradient (6) and Hessian (6x6)
static void crossnorm2_derivs(V3 a, V3 b, double* g6, double* H6) {
    V3 n = cross(a, b);
    put_vec(g6, 0, 2.0 * cross(b, n));
    put_vec(g6, 1, 2.0 * cross(n, a));
    double Ba[9], Bb[9], Bn[9], T[9];
    cpm(a, Ba); cpm(b, Bb); cpm(n, Bn);
    for (int i = 0; i < 36; ++i) H6[i] = 0;
    mm3(Bb, Bb, T); put_block(H6, 6, 0, 0, T, -2.0);
    mm3(Ba, Ba, T); put_block(H6, 6, 1, 1, T, -2.0);
    mm3(Bb, Ba, T);
    double M[9];
    for (int i = 0; i < 9; ++i) M[i] = 2.0 * T[i] - 2.0 * Bn[i];
    put_block(H6, 6, 0, 1, M, 1.0);
    double Mt[9];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Mt[i * 3 + j] = M[j * 3 + i];
    put_block(H6, 6, 1, 0, Mt, 1.0);
}

// d = s^2 / q with s = w.(a x b), q = |a x b|^2, over variables (w, a, b)
static void plane_dist_derivs(V3 w, V3 a, V3 b, double* g9, double* H9) {
    V3 n = cross(a, b);
    double s = dot(w, n), q = sq(n);
    double gs[9], gq[9], Hs[81], Hq[81];
    put_vec(gs, 0, n); put_vec(gs, 1, cross(b, w)); put_vec(gs, 2, cross(w, a));
    double g6[6], H6[36];
    crossnorm2_derivs(a, b, g6, H6);
    for (int i = 0; i < 3; ++i) gq[i] = 0;
    for (int i = 0; i < 6; ++i) gq[3 + i] = g6[i];
    for (int i = 0; i < 81; ++i) { Hs[i] = 0; Hq[i] = 0; }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) Hq[(3 + i) * 9 + 3 + j] = H6[i * 6 + j];
    double Ba[9], Bb[9], Bw[9];
    cpm(a, Ba); cpm(b, Bb); cpm(w, Bw);
    put_block(Hs, 9, 0, 1, Bb, -1.0); put_block(Hs, 9, 1, 0, Bb, 1.0);
    put_block(Hs, 9, 0, 2, Ba, 1.0);  put_block(Hs, 9, 2, 0, Ba, -1.0);
    put_block(Hs, 9, 1, 2, Bw, -1.0); put_block(Hs, 9, 2, 1, Bw, 1.0);
    double q2 = q * q, q3 = q2 * q;
    for (int i = 0; i < 9; ++i) g9[i] = 2.0 * s / q * gs[i] - s * s / q2 * gq[i];
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j)
            H9[i * 9 + j] = 2.0 / q * gs[i] * gs[j] + 2.0 * s / q * Hs[i * 9 + j]
                          - 2.0 * s / q2 * (gs[i] * gq[j] + gq[i] * gs[j])
                          + 2.0 * s * s / q3 * gq[i] * gq[j] - s * s / q2 * Hq[i * 9 + j];
}

void dist2_pp_gh(V3 a, V3 b, double* g, double* H) {
    V3 d = a - b;
    put_vec(g, 0, 2.0 * d); put_vec(g, 1, -2.0 * d);
    if (H) {
        for (int i = 0; i < 36; ++i) H[i] = 0;
        for (int k = 0; k < 3; ++k) {
            H[k * 6 + k] = 2; H[(3 + k) * 6 + 3 + k] = 2;
            H[k * 6 + 3 + k] = -2; H[(3 + k) * 6 + k] = -2;
        }
    }
}

// d = |c|^2/|e|^2 with u = e0 - p, v = e1 - p, c = u x v, e = v - u
void dist2_pe_gh(V3 p, V3 e0, V3 e1, double* g, double* H) {
    V3 u = e0 - p, v = e1 - p, c = cross(u, v), e = v - u;
    double N = sq(c), D = sq(e);
    double gN[6], HN[36], gD[6], HD[36];
    crossnorm2_derivs(u, v, gN, HN);
    put_vec(gD, 0, -2.0 * e); put_vec(gD, 1, 2.0 * e);
    for (int i = 0; i < 36; ++i) HD[i] = 0;
    for (int k = 0; k < 3; ++k) {
        HD[k * 6 + k] = 2; HD[(3 + k) * 6 + 3 + k] = 2;
        HD[k * 6 + 3 + k] = -2; HD[(3 + k) * 6 + k] = -2;
    }
    double D2 = D * D, D3 = D2 * D;
    double gv[6], Hv[36];
    for (int i = 0; i < 6; ++i) gv[i] = gN[i] / D - N / D2 * gD[i];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
            Hv[i * 6 + j] = HN[i * 6 + j] / D - (gN[i] * gD[j] + gD[i] * gN[j]) / D2
                          + 2.0 * N / D3 * gD[i] * gD[j] - N / D2 * HD[i * 6 + j];
    Lin L{3, {2, 1, 1}, {{0, 1}, {0}, {1}}, {{-1, -1}, {1}, {1}}};
    map_grad(L, 2, gv, g);
    if (H) map_hess(L, 2, Hv, H);
}

void dist2_pt_gh(V3 p, V3 t0, V3 t1, V3 t2, double* g, double* H) {
    double g9[9], H9[81];
    plane_dist_derivs(p - t0, t1 - t0, t2 - t0, g9, H9);
    Lin L{4, {1, 3, 1, 1}, {{0}, {0, 1, 2}, {1}, {2}}, {{1}, {-1, -1, -1}, {1}, {1}}};
    map_grad(L, 3, g9, g);
    if (H) map_hess(L, 3, H9, H);
}

void dist2_ee_gh(V3 a0, V3 a1, V3 b0, V3 b1, double* g, double* H) {
    double g9[9], H9[81];
    plane_dist_derivs(b0 - a0, a1 - a0, b1 - b0, g9, H9);
    Lin L{4, {2, 1, 2, 1}, {{0, 1}, {1}, {0, 2}, {2}}, {{-1, -1}, {1}, {1, -1}, {1}}};
    map_grad(L, 3, g9, g);
    if (H) map_hess(L, 3, H9, H);
}

// ------------------------------------------------------------ mollifier
// EDGE_EDGE_MOLLIFIER.h:9-17
double ee_cross_norm2(V3 a0, V3 a1, V3 b0, V3 b1) { return sq(cross(a1 - a0, b1 - b0)); }
// EDGE_EDGE_MOLLIFIER.h:583-591
double ee_mollifier_threshold(V3 a0r, V3 a1r, V3 b0r, V3 b1r) {
    return 1.0e-3 * sq(a0r - a1r) * sq(b0r - b1r);
}
// EDGE_EDGE_MOLLIFIER.h:441-525: e = (-c/eps + 2) c/eps for c < eps, else 1
void ee_mollifier(V3 a0, V3 a1, V3 b0, V3 b1, double eps, double* e, double* g, double* H) {
    double c = ee_cross_norm2(a0, a1, b0, b1);
    if (!(c < eps)) {
        *e = 1.0;
        if (g) for (int i = 0; i < 12; ++i) g[i] = 0;
        if (H) for (int i = 0; i < 144; ++i) H[i] = 0;
        return;
    }
    double r = c / eps;
    *e = (-r + 2.0) * r;
    if (!g && !H) return;
    double e1 = 2.0 * (1.0 / eps) * (-(1.0 / eps) * c + 1.0);
    double e2 = -2.0 / (eps * eps);
    double g6[6], H6[36], gv[6], Hv[36];
    crossnorm2_derivs(a1 - a0, b1 - b0, g6, H6);
    for (int i = 0; i < 6; ++i) gv[i] = e1 * g6[i];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) Hv[i * 6 + j] = e1 * H6[i * 6 + j] + e2 * g6[i] * g6[j];
    Lin L{4, {1, 1, 1, 1}, {{0}, {0}, {1}, {1}}, {{-1}, {1}, {-1}, {1}}};
    if (g) map_grad(L, 2, gv, g);
    if (H) map_hess(L, 2, Hv, H);
}

// ------------------------------------------------------------ barrier (BARRIER.h)
double barrier(double d, double dhat, double kappa) {
    return -kappa * (d - dhat) * (d - dhat) * std::log(d / dhat);
}
double barrier_g(double d, double dhat, double kappa) {
    double t2 = d - dhat;
    return kappa * (t2 * std::log(d / dhat) * -2.0 - (t2 * t2) / d);
}
double barrier_h(double d, double dhat, double kappa) {
    double t2 = d - dhat;
    return kappa * ((std::log(d / dhat) * -2.0 - t2 * 4.0 / d) + 1.0 / (d * d) * (t2 * t2));
}

// ------------------------------------------------------------ eigen + PSD
// Householder reduction to tridiagonal form followed by the QL algorithm with
// implicit shifts (Wilkinson & Reinsch, Handbook for Automatic Computation vol. II,
// routines tred2/tql2), the same family of method as Eigen's SelfAdjointEigenSolver.
void sym_eig(int n, const double* A, double* w, double* Vout) {
    double V[12][12], d[12], e[12];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i][j] = A[i * n + j];
    for (int j = 0; j < n; ++j) d[j] = V[n - 1][j];
    for (int i = n - 1; i > 0; --i) {
        double scale = 0, h = 0;
        for (int k = 0; k < i; ++k) scale += std::fabs(d[k]);
        if (scale == 0.0) {
            e[i] = d[i - 1];
            for (int j = 0; j < i; ++j) { d[j] = V[i - 1][j]; V[i][j] = 0; V[j][i] = 0; }
        } else {
            for (int k = 0; k < i; ++k) { d[k] /= scale; h += d[k] * d[k]; }
            double f = d[i - 1], g = std::sqrt(h);
            if (f > 0) g = -g;
            e[i] = scale * g;
            h -= f * g;
            d[i - 1] = f - g;
            for (int j = 0; j < i; ++j) e[j] = 0;
            for (int j = 0; j < i; ++j) {
                f = d[j]; V[j][i] = f; g = e[j] + V[j][j] * f;
                for (int k = j + 1; k <= i - 1; ++k) { g += V[k][j] * d[k]; e[k] += V[k][j] * f; }
                e[j] = g;
            }
            f = 0;
            for (int j = 0; j < i; ++j) { e[j] /= h; f += e[j] * d[j]; }
            double hh = f / (h + h);
            for (int j = 0; j < i; ++j) e[j] -= hh * d[j];
            for (int j = 0; j < i; ++j) {
                f = d[j]; g = e[j];
                for (int k = j; k <= i - 1; ++k) V[k][j] -= (f * e[k] + g * d[k]);
                d[j] = V[i - 1][j]; V[i][j] = 0;
            }
        }
        d[i] = h;
    }
    for (int i = 0; i < n - 1; ++i) {
        V[n - 1][i] = V[i][i]; V[i][i] = 1.0;
        double h = d[i + 1];
        if (h != 0.0) {
            for (int k = 0; k <= i; ++k) d[k] = V[k][i + 1] / h;
            for (int j = 0; j <= i; ++j) {
                double g = 0;
                for (int k = 0; k <= i; ++k) g += V[k][i + 1] * V[k][j];
                for (int k = 0; k <= i; ++k) V[k][j] -= g * d[k];
            }
        }
        for (int k = 0; k <= i; ++k) V[k][i + 1] = 0;
    }
    for (int j = 0; j < n; ++j) { d[j] = V[n - 1][j]; V[n - 1][j] = 0; }
    V[n - 1][n - 1] = 1.0;
    e[0] = 0;
    // QL with implicit shifts
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0;
    double f = 0, tst1 = 0;
    const double eps = std::ldexp(1.0, -52);
    for (int l = 0; l < n; ++l) {
        tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
        int m = l;
        while (m < n) { if (std::fabs(e[m]) <= eps * tst1) break; ++m; }
        if (m == n) m = n - 1;
        if (m > l) {
            int iter = 0;
            do {
                if (++iter > 100) break;
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = std::hypot(p, 1.0);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                double dl1 = d[l + 1];
                double h = g - d[l];
                for (int i = l + 2; i < n; ++i) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1, c2 = 1, c3 = 1, el1 = e[l + 1], s = 0, s2 = 0;
                for (int i = m - 1; i >= l; --i) {
                    c3 = c2; c2 = c; s2 = s;
                    g = c * e[i]; h = c * p;
                    r = std::hypot(p, e[i]);
                    e[i + 1] = s * r; s = e[i] / r; c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    for (int k = 0; k < n; ++k) {
                        h = V[k][i + 1];
                        V[k][i + 1] = s * V[k][i] + c * h;
                        V[k][i] = c * V[k][i] - s * h;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p; d[l] = c * p;
            } while (std::fabs(e[l]) > eps * tst1);
        }
        d[l] = d[l] + f; e[l] = 0;
    }
    for (int i = 0; i < n - 1; ++i) {
        int k = i; double p = d[i];
        for (int j = i + 1; j < n; ++j) if (d[j] < p) { k = j; p = d[j]; }
        if (k != i) {
            d[k] = d[i]; d[i] = p;
            for (int j = 0; j < n; ++j) std::swap(V[j][i], V[j][k]);
        }
    }
    for (int i = 0; i < n; ++i) w[i] = d[i];
    if (Vout) for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) Vout[i * n + j] = V[i][j];
}

// Math/UTILS.h:10-27
void make_pd(int n, double* A) {
    double w[12], V[144];
    sym_eig(n, A, w, V);
    if (w[0] >= 0) return;
    for (int i = 0; i < n; ++i) if (w[i] < 0) w[i] = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += V[i * n + k] * w[k] * V[j * n + k];
            A[i * n + j] = s;
        }
}

// ------------------------------------------------------------ membrane
// FEM/Shell/MEMBRANE.h:7-50 (energy), 52-119 (gradient), 121-290 (Hessian); B is the
// inverse of the rest first fundamental form (IB.invert() in the reference).
struct MembData { V3 D0, D1; double A00, A01, A11, detA, p, q, r, lnJ; };
static MembData memb_setup(V3 x0, V3 x1, V3 x2, const double* B) {
    MembData m;
    m.D0 = x1 - x0; m.D1 = x2 - x0;
    m.A00 = sq(m.D0); m.A01 = dot(m.D0, m.D1); m.A11 = sq(m.D1);
    m.detA = m.A00 * m.A11 - m.A01 * m.A01;
    m.p = m.A11 / m.detA; m.q = -m.A01 / m.detA; m.r = m.A00 / m.detA;
    double detB = B[0] * B[2] - B[1] * B[1];
    m.lnJ = std::log(m.detA * detB) / 2;
    return m;
}
double membrane_energy(V3 x0, V3 x1, V3 x2, const double* B, double vol, double lam, double mu) {
    MembData m = memb_setup(x0, x1, x2, B);
    double tr = B[0] * m.A00 + 2.0 * B[1] * m.A01 + B[2] * m.A11;
    return vol * (mu / 2 * (tr - 2 - 2 * m.lnJ) + lam / 2 * m.lnJ * m.lnJ);
}
void membrane_grad(V3 x0, V3 x1, V3 x2, const double* B, double vol, double lam, double mu, double* g) {
    MembData m = memb_setup(x0, x1, x2, B);
    double c1 = (-mu + lam * m.lnJ) / 2;
    double G00 = vol * (mu / 2 * B[0] + c1 * m.p);
    double G01 = vol * (mu / 2 * B[1] + c1 * m.q);
    double G11 = vol * (mu / 2 * B[2] + c1 * m.r);
    V3 gD0 = 2.0 * (V3{G00 * m.D0.x + G01 * m.D1.x, G00 * m.D0.y + G01 * m.D1.y, G00 * m.D0.z + G01 * m.D1.z});
    V3 gD1 = 2.0 * (V3{G01 * m.D0.x + G11 * m.D1.x, G01 * m.D0.y + G11 * m.D1.y, G01 * m.D0.z + G11 * m.D1.z});
    put_vec(g, 1, gD0); put_vec(g, 2, gD1);
    put_vec(g, 0, -1.0 * (gD0 + gD1));
}
void membrane_hess(V3 x0, V3 x1, V3 x2, const double* B, double vol, double lam, double mu, double* H) {
    MembData m = memb_setup(x0, x1, x2, B);
    double c1 = (-mu + lam * m.lnJ) / 2;
    double G[2][2];
    G[0][0] = vol * (mu / 2 * B[0] + c1 * m.p);
    G[0][1] = G[1][0] = vol * (mu / 2 * B[1] + c1 * m.q);
    G[1][1] = vol * (mu / 2 * B[2] + c1 * m.r);
    // per-slot vectors of alpha = 2 D0.dD0, beta = D1.dD0 + D0.dD1, gamma = 2 D1.dD1
    V3 vec[3][2] = {{2.0 * m.D0, {0, 0, 0}}, {m.D1, m.D0}, {{0, 0, 0}, 2.0 * m.D1}};
    double K[3][3] = {{m.p * m.p, 2 * m.p * m.q, m.q * m.q},
                      {2 * m.p * m.q, 2 * m.q * m.q + 2 * m.p * m.r, 2 * m.q * m.r},
                      {m.q * m.q, 2 * m.q * m.r, m.r * m.r}};
    double kt = vol * (mu - lam * m.lnJ) / 2;
    double kw = vol * lam / 4;
    V3 w[2] = {2.0 * m.p * m.D0 + 2.0 * m.q * m.D1, 2.0 * m.q * m.D0 + 2.0 * m.r * m.D1};
    double HD[36];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            double blk[9];
            double wa[3] = {w[a].x, w[a].y, w[a].z}, wb[3] = {w[b].x, w[b].y, w[b].z};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = (i == j ? 2.0 * G[a][b] : 0.0) + kw * wa[i] * wb[j];
                    for (int mm = 0; mm < 3; ++mm)
                        for (int nn = 0; nn < 3; ++nn) {
                            double va[3] = {vec[mm][a].x, vec[mm][a].y, vec[mm][a].z};
                            double vb[3] = {vec[nn][b].x, vec[nn][b].y, vec[nn][b].z};
                            s += kt * K[mm][nn] * va[i] * vb[j];
                        }
                    blk[i * 3 + j] = s;
                }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) HD[(a * 3 + i) * 6 + b * 3 + j] = blk[i * 3 + j];
        }
    Lin L{3, {2, 1, 1}, {{0, 1}, {0}, {1}}, {{-1, -1}, {1}, {1}}};
    map_hess(L, 2, HD, H);
}

// ------------------------------------------------------------ dihedral angle
// Math/DIHEDRAL_ANGLE.h:10-24: n1 = (x1-x0)x(x2-x0), n2 = (x2-x3)x(x1-x3),
// theta = acos(n1.n2/|n1||n2|), negated when (n2 x n1).(x1 - x2) < 0.
double dihedral_angle(V3 x0, V3 x1, V3 x2, V3 x3) {
    V3 n1 = cross(x1 - x0, x2 - x0);
    V3 n2 = cross(x2 - x3, x1 - x3);
    double c = dot(n1, n2) / std::sqrt(sq(n1) * sq(n2));
    c = std::max(-1.0, std::min(1.0, c));
    double th = std::acos(c);
    if (dot(cross(n2, n1), x1 - x2) < 0) th = -th;
    return th;
}
// gradient: BENDING.h:177-208 uses DIHEDRAL_ANGLE.h:176-200 (closed form in terms of
// the face normals and the hinge edge)
void dihedral_angle_grad(V3 x0, V3 x1, V3 x2, V3 x3, double* g) {
    V3 e = x2 - x1;
    double L = std::sqrt(sq(e));
    V3 nA = cross(x1 - x0, x2 - x0), nB = cross(x2 - x3, x1 - x3);
    double AA = sq(nA), AB = sq(nB);
    put_vec(g, 0, (-L / AA) * nA);
    put_vec(g, 3, (-L / AB) * nB);
    put_vec(g, 1, (-dot(e, x0 - x2) / (L * AA)) * nA + (-dot(e, x3 - x2) / (L * AB)) * nB);
    put_vec(g, 2, (dot(e, x0 - x1) / (L * AA)) * nA + (dot(e, x3 - x1) / (L * AB)) * nB);
}
// Hessian by differentiating the four gradient blocks g_i = sA_i nA + sB_i nB
void dihedral_angle_hess(V3 x0, V3 x1, V3 x2, V3 x3, double* H) {
    V3 X[4] = {x0, x1, x2, x3};
    V3 e = x2 - x1;
    double L = std::sqrt(sq(e));
    V3 nA = cross(x1 - x0, x2 - x0), nB = cross(x2 - x3, x1 - x3);
    double AA = sq(nA), AB = sq(nB);
    // dn/dx_j = [vA_j]x, [vB_j]x
    V3 zero{0, 0, 0};
    V3 vA[4] = {x2 - x1, x0 - x2, x1 - x0, zero};
    V3 vB[4] = {zero, x2 - x3, x3 - x1, x1 - x2};
    V3 gL[4] = {zero, (-1.0 / L) * e, (1.0 / L) * e, zero};
    V3 gAA[4], gAB[4];
    for (int j = 0; j < 4; ++j) { gAA[j] = 2.0 * cross(nA, vA[j]); gAB[j] = 2.0 * cross(nB, vB[j]); }
    // P terms: PA1 = e.(x0-x2), PB1 = e.(x3-x2), PA2 = e.(x0-x1), PB2 = e.(x3-x1)
    double PA1 = dot(e, x0 - x2), PB1 = dot(e, x3 - x2), PA2 = dot(e, x0 - x1), PB2 = dot(e, x3 - x1);
    V3 gPA1[4] = {e, -1.0 * (x0 - x2), (x0 - x2) - e, zero};
    V3 gPB1[4] = {zero, -1.0 * (x3 - x2), (x3 - x2) - e, e};
    V3 gPA2[4] = {e, -1.0 * (x0 - x1) - e, x0 - x1, zero};
    V3 gPB2[4] = {zero, -1.0 * (x3 - x1) - e, x3 - x1, e};
    // coefficient scalars and gradients: s = sign * P / (L A) or s = -L/A
    double sA[4], sB[4];
    V3 gsA[4][4], gsB[4][4];
    auto ratio = [&](double sgn, double P, const V3* gP, double A, const V3* gA, double* s, V3* gs) {
        double LA = L * A;
        *s = sgn * P / LA;
        for (int j = 0; j < 4; ++j) {
            V3 num = (A * gL[j] + L * gA[j]);
            gs[j] = sgn * ((1.0 / LA) * gP[j] - (P / (LA * LA)) * num);
        }
    };
    // g0 = -(L/AA) nA
    sA[0] = -L / AA; sB[0] = 0;
    for (int j = 0; j < 4; ++j) { gsA[0][j] = (-1.0 / AA) * gL[j] + (L / (AA * AA)) * gAA[j]; gsB[0][j] = zero; }
    // g3 = -(L/AB) nB
    sA[3] = 0; sB[3] = -L / AB;
    for (int j = 0; j < 4; ++j) { gsB[3][j] = (-1.0 / AB) * gL[j] + (L / (AB * AB)) * gAB[j]; gsA[3][j] = zero; }
    ratio(-1.0, PA1, gPA1, AA, gAA, &sA[1], gsA[1]);
    ratio(-1.0, PB1, gPB1, AB, gAB, &sB[1], gsB[1]);
    ratio(1.0, PA2, gPA2, AA, gAA, &sA[2], gsA[2]);
    ratio(1.0, PB2, gPB2, AB, gAB, &sB[2], gsB[2]);
    (void)X;
    double Hr[144];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double CA[9], CB[9];
            cpm(vA[j], CA); cpm(vB[j], CB);
            double nAv[3] = {nA.x, nA.y, nA.z}, nBv[3] = {nB.x, nB.y, nB.z};
            double ga[3] = {gsA[i][j].x, gsA[i][j].y, gsA[i][j].z};
            double gb[3] = {gsB[i][j].x, gsB[i][j].y, gsB[i][j].z};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    Hr[(i * 3 + a) * 12 + j * 3 + b] = nAv[a] * ga[b] + sA[i] * CA[a * 3 + b]
                                                      + nBv[a] * gb[b] + sB[i] * CB[a * 3 + b];
        }
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) H[i * 12 + j] = 0.5 * (Hr[i * 12 + j] + Hr[j * 12 + i]);
}

}  // namespace orc
