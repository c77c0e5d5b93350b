// Device math for the contact-and-Newton step: squared distances and their region
// classifiers, closed-form gradients/Hessians, the edge-edge mollifier, the log
// barrier, per-stencil PSD projection (cyclic Jacobi), the Neo-Hookean membrane and
// the hinge dihedral angle.
//
// Reference behaviour: Math/Distance/{DISTANCE_TYPE,POINT_POINT,POINT_EDGE,
// POINT_TRIANGLE,EDGE_EDGE,EDGE_EDGE_MOLLIFIER,DISTANCE_UNCLASSIFIED}.h,
// Math/BARRIER.h, Math/UTILS.h (makePD), FEM/Shell/MEMBRANE.h, Math/DIHEDRAL_ANGLE.h.
//
// Classification predicates and distances (everything that decides membership of
// the active set or a CCD time of impact) are evaluated without fused multiply-adds
// (fp contract off) so they round exactly like an IEEE double evaluation of the same
// expression order on the host.  Derivative blocks allow contraction.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cipc {

#define CIPC_DEV __device__ __forceinline__

struct d3 { double x, y, z; };

#pragma clang fp contract(off)

CIPC_DEV d3 mk(double x, double y, double z) { return {x, y, z}; }
CIPC_DEV d3 ld3(const double* __restrict__ p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
CIPC_DEV d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
CIPC_DEV d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
CIPC_DEV d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
CIPC_DEV double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CIPC_DEV double sq(d3 a) { return dot(a, a); }
CIPC_DEV d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
CIPC_DEV double dmin(double a, double b) { return a < b ? a : b; }
CIPC_DEV double dmax(double a, double b) { return a > b ? a : b; }

// ------------------------------------------------------------- squared distances
CIPC_DEV double dist2_pp(d3 a, d3 b) { return sq(a - b); }
CIPC_DEV double dist2_pe(d3 p, d3 e0, d3 e1) { return sq(cross(e0 - p, e1 - p)) / sq(e1 - e0); }
CIPC_DEV double dist2_pt(d3 p, d3 t0, d3 t1, d3 t2) {
    d3 b = cross(t1 - t0, t2 - t0);
    double aTb = dot(p - t0, b);
    return aTb * aTb / sq(b);
}
CIPC_DEV double dist2_ee(d3 a0, d3 a1, d3 b0, d3 b1) {
    d3 b = cross(a1 - a0, b1 - b0);
    double aTb = dot(b0 - a0, b);
    return aTb * aTb / sq(b);
}

// ------------------------------------------------------------- region classifiers
CIPC_DEV int pe_type(d3 p, d3 e0, d3 e1) {
    d3 e = e1 - e0;
    double ratio = dot(e, p - e0) / sq(e);
    return ratio < 0 ? 0 : (ratio > 1 ? 1 : 2);
}

// 2x2 symmetric LDL^T solve with diagonal pivoting (the Eigen LDLT of
// DISTANCE_TYPE.h:46,54,62 on a 2x2 system)
CIPC_DEV void ldlt2_solve(double a, double b, double c, double r0, double r1, double& x0, double& x1) {
    bool swp = fabs(c) > fabs(a);
    if (swp) { double t = a; a = c; c = t; t = r0; r0 = r1; r1 = t; }
    double l = b / a;
    double d1 = c - l * b;
    double y1 = r1 - l * r0;
    double z0 = r0 / a;
    double z1 = y1 / d1;
    double s0 = z0 - l * z1;
    if (swp) { x0 = z1; x1 = s0; } else { x0 = s0; x1 = z1; }
}
CIPC_DEV void pt_edge_param(d3 o, d3 r0, d3 nv, d3 p, double& s, double& t) {
    d3 r1 = cross(r0, nv);
    d3 w = p - o;
    ldlt2_solve(dot(r0, r0), dot(r0, r1), dot(r1, r1), dot(r0, w), dot(r1, w), s, t);
}
// DISTANCE_TYPE.h:31-82
CIPC_DEV int pt_type(d3 p, d3 t0, d3 t1, d3 t2) {
    d3 nv = cross(t1 - t0, t2 - t0);
    double s0, h0, s1, h1, s2, h2;
    pt_edge_param(t0, t1 - t0, nv, p, s0, h0);
    if (s0 > 0.0 && s0 < 1.0 && h0 >= 0.0) return 3;
    pt_edge_param(t1, t2 - t1, nv, p, s1, h1);
    if (s1 > 0.0 && s1 < 1.0 && h1 >= 0.0) return 4;
    pt_edge_param(t2, t0 - t2, nv, p, s2, h2);
    if (s2 > 0.0 && s2 < 1.0 && h2 >= 0.0) return 5;
    if (s0 <= 0.0 && s2 >= 1.0) return 0;
    if (s1 <= 0.0 && s0 >= 1.0) return 1;
    if (s2 <= 0.0 && s1 >= 1.0) return 2;
    return 6;
}
// DISTANCE_TYPE.h:86-164
CIPC_DEV int ee_type(d3 a0, d3 a1, d3 b0, d3 b1) {
    d3 u = a1 - a0, v = b1 - b0, w = a0 - b0;
    double a = sq(u), b = dot(u, v), c = sq(v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b;
    double tD = D, sN, tN;
    int defaultCase = 8;
    sN = b * e - c * d;
    if (sN <= 0.0) {
        tN = e; tD = c; defaultCase = 2;
    } else if (sN >= D) {
        tN = e + b; tD = c; defaultCase = 5;
    } else {
        tN = a * e - b * d;
        d3 uv = cross(u, v);
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
// DISTANCE_UNCLASSIFIED.h:125-148
CIPC_DEV double dist2_pe_unclassified(d3 p, d3 e0, d3 e1) {
    d3 v = e1 - e0, w = p - e0;
    double c1 = dot(w, v);
    if (c1 <= 0.0) return dist2_pp(p, e0);
    double c2 = sq(v);
    if (c2 <= c1) return dist2_pp(p, e1);
    double b = c1 / c2;
    return dist2_pp(p, e0 + b * v);
}
CIPC_DEV double dist2_pt_unclassified(d3 p, d3 t0, d3 t1, d3 t2) {
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
CIPC_DEV double dist2_ee_unclassified(d3 a0, d3 a1, d3 b0, d3 b1) {
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
CIPC_DEV double ee_mollifier_threshold(d3 a0r, d3 a1r, d3 b0r, d3 b1r) {
    return 1.0e-3 * sq(a0r - a1r) * sq(b0r - b1r);
}
CIPC_DEV double ee_cross_norm2(d3 a0, d3 a1, d3 b0, d3 b1) { return sq(cross(a1 - a0, b1 - b0)); }

// ------------------------------------------------------------- barrier (BARRIER.h)
CIPC_DEV double barrier_e(double d, double dhat, double kappa) {
    return -kappa * (d - dhat) * (d - dhat) * log(d / dhat);
}
CIPC_DEV double barrier_g(double d, double dhat, double kappa) {
    double t2 = d - dhat;
    return kappa * (t2 * log(d / dhat) * -2.0 - (t2 * t2) / d);
}
CIPC_DEV double barrier_h(double d, double dhat, double kappa) {
    double t2 = d - dhat;
    return kappa * ((log(d / dhat) * -2.0 - t2 * 4.0 / d) + 1.0 / (d * d) * (t2 * t2));
}

#pragma clang fp contract(fast)

// ------------------------------------------------------------- small matrices
// n x n row-major local matrices live in thread-private arrays; indices are
// compile-time after unrolling so they stay in registers.
CIPC_DEV void cpm(d3 v, double* M) {
    M[0] = 0; M[1] = -v.z; M[2] = v.y;
    M[3] = v.z; M[4] = 0; M[5] = -v.x;
    M[6] = -v.y; M[7] = v.x; M[8] = 0;
}
CIPC_DEV void mm3(const double* A, const double* B, double* C) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i * 3 + j] = A[i * 3 + 0] * B[0 * 3 + j] + A[i * 3 + 1] * B[1 * 3 + j] + A[i * 3 + 2] * B[2 * 3 + j];
}
template <int N>
CIPC_DEV void put_block(double* H, int bi, int bj, const double* M, double s) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) H[(bi * 3 + i) * N + bj * 3 + j] += s * M[i * 3 + j];
}
CIPC_DEV void put_vec(double* g, int bi, d3 v) { g[bi * 3] = v.x; g[bi * 3 + 1] = v.y; g[bi * 3 + 2] = v.z; }

// gradient (6) and Hessian (6x6) of |a x b|^2 over (a, b)
CIPC_DEV void crossnorm2_derivs(d3 a, d3 b, double* g6, double* H6) {
    d3 n = cross(a, b);
    put_vec(g6, 0, 2.0 * cross(b, n));
    put_vec(g6, 1, 2.0 * cross(n, a));
    if (!H6) return;
    double Ba[9], Bb[9], Bn[9], T[9];
    cpm(a, Ba); cpm(b, Bb); cpm(n, Bn);
#pragma unroll
    for (int i = 0; i < 36; ++i) H6[i] = 0;
    mm3(Bb, Bb, T); put_block<6>(H6, 0, 0, T, -2.0);
    mm3(Ba, Ba, T); put_block<6>(H6, 1, 1, T, -2.0);
    mm3(Bb, Ba, T);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double m = 2.0 * T[i * 3 + j] - 2.0 * Bn[i * 3 + j];
            H6[i * 6 + 3 + j] = m;
            H6[(3 + j) * 6 + i] = m;
        }
}

// d = s^2/q, s = w.(a x b), q = |a x b|^2 over (w, a, b)
CIPC_DEV void plane_dist_derivs(d3 w, d3 a, d3 b, double* g9, double* H9) {
    d3 n = cross(a, b);
    double s = dot(w, n), q = sq(n);
    double gs[9], gq[9];
    put_vec(gs, 0, n); put_vec(gs, 1, cross(b, w)); put_vec(gs, 2, cross(w, a));
    double g6[6], H6[36];
    crossnorm2_derivs(a, b, g6, H9 ? H6 : nullptr);
    gq[0] = gq[1] = gq[2] = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) gq[3 + i] = g6[i];
    double iq = 1.0 / q;
    double c1 = 2.0 * s * iq, c2 = s * s * iq * iq;
#pragma unroll
    for (int i = 0; i < 9; ++i) g9[i] = c1 * gs[i] - c2 * gq[i];
    if (!H9) return;
    double Ba[9], Bb[9], Bw[9];
    cpm(a, Ba); cpm(b, Bb); cpm(w, Bw);
    // second derivative of s (blocks (w,a) = -[b]x, (w,b) = [a]x, (a,b) = -[w]x)
    double c3 = 2.0 * iq, c4 = 2.0 * s * iq * iq, c5 = 2.0 * s * s * iq * iq * iq;
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            int bi = i / 3, bj = j / 3, r = i % 3, c = j % 3;
            double hs = 0;
            if (bi == 0 && bj == 1) hs = -Bb[r * 3 + c];
            if (bi == 1 && bj == 0) hs = Bb[r * 3 + c];
            if (bi == 0 && bj == 2) hs = Ba[r * 3 + c];
            if (bi == 2 && bj == 0) hs = -Ba[r * 3 + c];
            if (bi == 1 && bj == 2) hs = -Bw[r * 3 + c];
            if (bi == 2 && bj == 1) hs = Bw[r * 3 + c];
            double hq = (bi > 0 && bj > 0) ? H6[(i - 3) * 6 + (j - 3)] : 0.0;
            H9[i * 9 + j] = c3 * gs[i] * gs[j] + c1 * hs - c4 * (gs[i] * gq[j] + gq[i] * gs[j]) + c5 * gq[i] * gq[j] - c2 * hq;
        }
}

// vertex-space maps: each vertex k of the stencil is a signed sum of difference slots
struct SlotMap { int cnt[4]; int slot[4][3]; double sgn[4][3]; };
template <int NX, int NV>
CIPC_DEV void map_grad(const SlotMap& L, const double* gv, double* gx) {
#pragma unroll
    for (int k = 0; k < NX; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double s = 0;
#pragma unroll
            for (int t = 0; t < 3; ++t)
                if (t < L.cnt[k]) s += L.sgn[k][t] * gv[L.slot[k][t] * 3 + d];
            gx[k * 3 + d] = s;
        }
}
template <int NX, int NV>
CIPC_DEV void map_hess(const SlotMap& L, const double* Hv, double* Hx) {
    constexpr int nv = NV * 3, nx = NX * 3;
#pragma unroll
    for (int ki = 0; ki < NX; ++ki)
#pragma unroll
        for (int kj = 0; kj < NX; ++kj)
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    double s = 0;
#pragma unroll
                    for (int ti = 0; ti < 3; ++ti)
#pragma unroll
                        for (int tj = 0; tj < 3; ++tj)
                            if (ti < L.cnt[ki] && tj < L.cnt[kj])
                                s += L.sgn[ki][ti] * L.sgn[kj][tj] * Hv[(L.slot[ki][ti] * 3 + a) * nv + L.slot[kj][tj] * 3 + b];
                    Hx[(ki * 3 + a) * nx + kj * 3 + b] = s;
                }
}

CIPC_DEV void dist2_pp_gh(d3 a, d3 b, double* g, double* H) {
    d3 d = a - b;
    put_vec(g, 0, 2.0 * d); put_vec(g, 1, -2.0 * d);
    if (H) {
#pragma unroll
        for (int i = 0; i < 36; ++i) H[i] = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            H[k * 6 + k] = 2; H[(3 + k) * 6 + 3 + k] = 2;
            H[k * 6 + 3 + k] = -2; H[(3 + k) * 6 + k] = -2;
        }
    }
}
CIPC_DEV void dist2_pe_gh(d3 p, d3 e0, d3 e1, double* g, double* H) {
    d3 u = e0 - p, v = e1 - p, c = cross(u, v), e = v - u;
    double N = sq(c), D = sq(e);
    double gN[6], HN[36], gD[6];
    crossnorm2_derivs(u, v, gN, H ? HN : nullptr);
    put_vec(gD, 0, -2.0 * e); put_vec(gD, 1, 2.0 * e);
    double iD = 1.0 / D, ND2 = N * iD * iD;
    double gv[6], Hv[36];
#pragma unroll
    for (int i = 0; i < 6; ++i) gv[i] = gN[i] * iD - ND2 * gD[i];
    const SlotMap L{{2, 1, 1, 0}, {{0, 1, 0}, {0, 0, 0}, {1, 0, 0}, {0, 0, 0}}, {{-1, -1, 0}, {1, 0, 0}, {1, 0, 0}, {0, 0, 0}}};
    map_grad<3, 2>(L, gv, g);
    if (!H) return;
    double c3 = 2.0 * N * iD * iD * iD;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double hd = (i % 3 == j % 3) ? ((i / 3 == j / 3) ? 2.0 : -2.0) : 0.0;
            Hv[i * 6 + j] = HN[i * 6 + j] * iD - (gN[i] * gD[j] + gD[i] * gN[j]) * iD * iD + c3 * gD[i] * gD[j] - ND2 * hd;
        }
    map_hess<3, 2>(L, Hv, H);
}
CIPC_DEV void dist2_pt_gh(d3 p, d3 t0, d3 t1, d3 t2, double* g, double* H) {
    double g9[9], H9[81];
    plane_dist_derivs(p - t0, t1 - t0, t2 - t0, g9, H ? H9 : nullptr);
    const SlotMap L{{1, 3, 1, 1}, {{0, 0, 0}, {0, 1, 2}, {1, 0, 0}, {2, 0, 0}}, {{1, 0, 0}, {-1, -1, -1}, {1, 0, 0}, {1, 0, 0}}};
    map_grad<4, 3>(L, g9, g);
    if (H) map_hess<4, 3>(L, H9, H);
}
CIPC_DEV void dist2_ee_gh(d3 a0, d3 a1, d3 b0, d3 b1, double* g, double* H) {
    double g9[9], H9[81];
    plane_dist_derivs(b0 - a0, a1 - a0, b1 - b0, g9, H ? H9 : nullptr);
    const SlotMap L{{2, 1, 2, 1}, {{0, 1, 0}, {1, 0, 0}, {0, 2, 0}, {2, 0, 0}}, {{-1, -1, 0}, {1, 0, 0}, {1, -1, 0}, {1, 0, 0}}};
    map_grad<4, 3>(L, g9, g);
    if (H) map_hess<4, 3>(L, H9, H);
}

// EDGE_EDGE_MOLLIFIER.h:441-525: value, 12-gradient, 12x12 Hessian
CIPC_DEV double ee_mollifier(d3 a0, d3 a1, d3 b0, d3 b1, double eps, double* g, double* H) {
    double c;
    {
#pragma clang fp contract(off)
        c = ee_cross_norm2(a0, a1, b0, b1);
    }
    if (!(c < eps)) {
        if (g) {
#pragma unroll
            for (int i = 0; i < 12; ++i) g[i] = 0;
        }
        if (H) {
#pragma unroll
            for (int i = 0; i < 144; ++i) H[i] = 0;
        }
        return 1.0;
    }
    double e;
    {
#pragma clang fp contract(off)
        double r = c / eps;
        e = (-r + 2.0) * r;
    }
    if (!g && !H) return e;
    double e1 = 2.0 * (1.0 / eps) * (-(1.0 / eps) * c + 1.0);
    double e2 = -2.0 / (eps * eps);
    double g6[6], H6[36], gv[6], Hv[36];
    crossnorm2_derivs(a1 - a0, b1 - b0, g6, H ? H6 : nullptr);
#pragma unroll
    for (int i = 0; i < 6; ++i) gv[i] = e1 * g6[i];
    const SlotMap L{{1, 1, 1, 1}, {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {1, 0, 0}}, {{-1, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {1, 0, 0}}};
    if (g) map_grad<4, 2>(L, gv, g);
    if (H) {
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) Hv[i * 6 + j] = e1 * H6[i * 6 + j] + e2 * g6[i] * g6[j];
        map_hess<4, 2>(L, Hv, H);
    }
    return e;
}

// ------------------------------------------------------------- translation-reduced PSD projection
// For the Hessian of a translation-invariant function of NV points (hinge bending,
// every barrier: distances and the edge-edge mollifier depend only on differences),
// H t = 0 for every rigid translation t, so with an orthonormal basis Q = Wq (x) I3 of the
// translation complement (Wq: NV x (NV-1), columns orthogonal to (1, ..., 1)),
// makePD(H) = Q makePD(Q^T H Q) Q^T: the eigenproblem shrinks from 3 NV to 3 (NV - 1).
// Wq: NV = 2: (1, -1)/sqrt2; NV = 3: (1, -1, 0)/sqrt2, (1, 1, -2)/sqrt6; NV = 4: the
// +-1/2 patterns (+,-,+,-), (+,+,-,-), (+,-,-,+).
template <int NV>
CIPC_DEV constexpr double tr_basis(int k, int a) {
    if (NV == 2) return k == 0 ? 0.70710678118654752440 : -0.70710678118654752440;
    if (NV == 3) {
        if (a == 0) return k == 0 ? 0.70710678118654752440 : (k == 1 ? -0.70710678118654752440 : 0.0);
        return k == 2 ? -0.81649658092772603273 : 0.40824829046386301637;
    }
    const int sg = a == 0 ? ((k & 1) ? -1 : 1) : a == 1 ? ((k >> 1) ? -1 : 1) : (((k & 1) ^ (k >> 1)) ? -1 : 1);
    return 0.5 * sg;
}
CIPC_DEV constexpr int upk(int n, int i, int j) { return i <= j ? i * n - i * (i - 1) / 2 + (j - i) : j * n - j * (j - 1) / 2 + (i - j); }

// H (row stride `ld`, leading 3 NV x 3 NV block) -> PSD projection into out(r, c) via
// store(r, c, value).  Positive-definite reduced matrices are returned unchanged; the rest
// go through cyclic Jacobi on the reduced matrix (off^2 <= 1e-24 of ||.||_F^2).
template <int NV, class Store>
CIPC_DEV void make_pd_translation(const double* __restrict__ Hm, int ld, Store store) {
    constexpr int R = 3 * (NV - 1), NU = R * (R + 1) / 2;
    double a[NU];
#pragma unroll
    for (int e = 0; e < NU; ++e) a[e] = 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int l = 0; l < NV; ++l)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double v = Hm[(3 * k + i) * ld + 3 * l + j];
#pragma unroll
                    for (int p = 0; p < NV - 1; ++p)
#pragma unroll
                        for (int q = 0; q < NV - 1; ++q) {
                            const int rr = 3 * p + i, cc = 3 * q + j;
                            if (rr <= cc) a[upk(R, rr, cc)] += tr_basis<NV>(k, p) * tr_basis<NV>(l, q) * v;
                        }
                }
    // positive definite: unchanged
    bool pd = true;
    {
        double L[NU];
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                double sum = a[upk(R, j, i)];
#pragma unroll
                for (int k = 0; k < j; ++k) sum -= L[upk(R, k, i)] * L[upk(R, k, j)];
                if (i == j) {
                    pd = pd && sum > 0;
                    L[upk(R, i, i)] = sqrt(sum > 0 ? sum : 1.0);
                } else {
                    L[upk(R, j, i)] = sum / L[upk(R, j, j)];
                }
            }
    }
    if (pd) {
#pragma unroll
        for (int r = 0; r < 3 * NV; ++r)
#pragma unroll
            for (int c = 0; c < 3 * NV; ++c) store(r, c, Hm[r * ld + c]);
        return;
    }
    double V[R * R];
#pragma unroll
    for (int e = 0; e < R * R; ++e) V[e] = (e % (R + 1) == 0) ? 1.0 : 0.0;
    double fro = 0;
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = i; j < R; ++j) fro += (i == j ? 1.0 : 2.0) * a[upk(R, i, j)] * a[upk(R, i, j)];
    const double tol = 1e-24 * fro;
    for (int sweep = 0; sweep < 16; ++sweep) {
        double offd = 0;
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = i + 1; j < R; ++j) offd += a[upk(R, i, j)] * a[upk(R, i, j)];
        if (offd <= 0.5 * tol) break;
#pragma unroll
        for (int p = 0; p < R - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < R; ++q) {
                const double apq = a[upk(R, p, q)];
                if (apq == 0.0) continue;
                const double app = a[upk(R, p, p)], aqq = a[upk(R, q, q)];
                const double theta = (aqq - app) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                a[upk(R, p, p)] = app - t * apq;
                a[upk(R, q, q)] = aqq + t * apq;
                a[upk(R, p, q)] = 0.0;
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (k == p || k == q) continue;
                    const double akp = a[upk(R, k, p)], akq = a[upk(R, k, q)];
                    a[upk(R, k, p)] = c * akp - s * akq;
                    a[upk(R, k, q)] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const double vkp = V[R * k + p], vkq = V[R * k + q];
                    V[R * k + p] = c * vkp - s * vkq;
                    V[R * k + q] = s * vkp + c * vkq;
                }
            }
    }
    double lam[R];
#pragma unroll
    for (int i = 0; i < R; ++i) lam[i] = a[upk(R, i, i)] > 0 ? a[upk(R, i, i)] : 0.0;
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = i; j < R; ++j) {
            double sum = 0;
#pragma unroll
            for (int m = 0; m < R; ++m) sum += V[R * i + m] * lam[m] * V[R * j + m];
            a[upk(R, i, j)] = sum;
        }
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int l = 0; l < NV; ++l)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double sum = 0;
#pragma unroll
                    for (int p = 0; p < NV - 1; ++p)
#pragma unroll
                        for (int q = 0; q < NV - 1; ++q) sum += tr_basis<NV>(k, p) * tr_basis<NV>(l, q) * a[upk(R, 3 * p + i, 3 * q + j)];
                    store(3 * k + i, 3 * l + j, sum);
                }
}


// ------------------------------------------------------------- membrane
// FEM/Shell/MEMBRANE.h (useNH): E = vol (mu/2 (tr(B A) - 2 - 2 lnJ) + lam/2 lnJ^2),
// A = D^T D, D = [x1 - x0, x2 - x0], B = rest A^{-1}, lnJ = ln(det A det B)/2.
struct Memb { d3 D0, D1; double A00, A01, A11, detA, p, q, r, lnJ; };
CIPC_DEV Memb memb_setup(d3 x0, d3 x1, d3 x2, double B0, double B1, double B2) {
    Memb m;
    m.D0 = x1 - x0; m.D1 = x2 - x0;
    m.A00 = sq(m.D0); m.A01 = dot(m.D0, m.D1); m.A11 = sq(m.D1);
    m.detA = m.A00 * m.A11 - m.A01 * m.A01;
    double id = 1.0 / m.detA;
    m.p = m.A11 * id; m.q = -m.A01 * id; m.r = m.A00 * id;
    m.lnJ = log(m.detA * (B0 * B2 - B1 * B1)) / 2;
    return m;
}
CIPC_DEV double membrane_energy(d3 x0, d3 x1, d3 x2, double B0, double B1, double B2, double vol, double lam, double mu) {
    Memb m = memb_setup(x0, x1, x2, B0, B1, B2);
    double tr = B0 * m.A00 + 2.0 * B1 * m.A01 + B2 * m.A11;
    return vol * (mu / 2 * (tr - 2 - 2 * m.lnJ) + lam / 2 * m.lnJ * m.lnJ);
}
CIPC_DEV void membrane_grad(d3 x0, d3 x1, d3 x2, double B0, double B1, double B2, double vol, double lam, double mu, double* g) {
    Memb m = memb_setup(x0, x1, x2, B0, B1, B2);
    double c1 = (-mu + lam * m.lnJ) / 2;
    double G00 = vol * (mu / 2 * B0 + c1 * m.p), G01 = vol * (mu / 2 * B1 + c1 * m.q), G11 = vol * (mu / 2 * B2 + c1 * m.r);
    d3 gD0 = 2.0 * (G00 * m.D0 + G01 * m.D1);
    d3 gD1 = 2.0 * (G01 * m.D0 + G11 * m.D1);
    put_vec(g, 1, gD0); put_vec(g, 2, gD1); put_vec(g, 0, -1.0 * (gD0 + gD1));
}
CIPC_DEV void membrane_hess(d3 x0, d3 x1, d3 x2, double B0, double B1, double B2, double vol, double lam, double mu, double* H) {
    Memb m = memb_setup(x0, x1, x2, B0, B1, B2);
    double c1 = (-mu + lam * m.lnJ) / 2;
    double G[2][2];
    G[0][0] = vol * (mu / 2 * B0 + c1 * m.p);
    G[0][1] = G[1][0] = vol * (mu / 2 * B1 + c1 * m.q);
    G[1][1] = vol * (mu / 2 * B2 + c1 * m.r);
    d3 zero = {0, 0, 0};
    d3 vec[3][2] = {{2.0 * m.D0, zero}, {m.D1, m.D0}, {zero, 2.0 * m.D1}};
    double K[3][3] = {{m.p * m.p, 2 * m.p * m.q, m.q * m.q},
                      {2 * m.p * m.q, 2 * m.q * m.q + 2 * m.p * m.r, 2 * m.q * m.r},
                      {m.q * m.q, 2 * m.q * m.r, m.r * m.r}};
    double kt = vol * (mu - lam * m.lnJ) / 2, kw = vol * lam / 4;
    d3 w[2] = {2.0 * m.p * m.D0 + 2.0 * m.q * m.D1, 2.0 * m.q * m.D0 + 2.0 * m.r * m.D1};
    double HD[36];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            double wa[3] = {w[a].x, w[a].y, w[a].z}, wb[3] = {w[b].x, w[b].y, w[b].z};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double s = (i == j ? 2.0 * G[a][b] : 0.0) + kw * wa[i] * wb[j];
#pragma unroll
                    for (int mm = 0; mm < 3; ++mm)
#pragma unroll
                        for (int nn = 0; nn < 3; ++nn) {
                            double va[3] = {vec[mm][a].x, vec[mm][a].y, vec[mm][a].z};
                            double vb[3] = {vec[nn][b].x, vec[nn][b].y, vec[nn][b].z};
                            s += kt * K[mm][nn] * va[i] * vb[j];
                        }
                    HD[(a * 3 + i) * 6 + b * 3 + j] = s;
                }
        }
    const SlotMap L{{2, 1, 1, 0}, {{0, 1, 0}, {0, 0, 0}, {1, 0, 0}, {0, 0, 0}}, {{-1, -1, 0}, {1, 0, 0}, {1, 0, 0}, {0, 0, 0}}};
    map_hess<3, 2>(L, HD, H);
}

// ------------------------------------------------------------- dihedral angle
// Math/DIHEDRAL_ANGLE.h:10-24 convention: n1 = (x1-x0)x(x2-x0), n2 = (x2-x3)x(x1-x3),
// theta = acos(n1.n2/|n1||n2|) negated when (n2 x n1).(x1 - x2) < 0.
CIPC_DEV double dihedral_angle(d3 x0, d3 x1, d3 x2, d3 x3) {
    d3 n1 = cross(x1 - x0, x2 - x0), n2 = cross(x2 - x3, x1 - x3);
    double c = dot(n1, n2) / sqrt(sq(n1) * sq(n2));
    c = dmax(-1.0, dmin(1.0, c));
    double th = acos(c);
    if (dot(cross(n2, n1), x1 - x2) < 0) th = -th;
    return th;
}
CIPC_DEV void dihedral_angle_grad(d3 x0, d3 x1, d3 x2, d3 x3, double* g) {
    d3 e = x2 - x1;
    double L = sqrt(sq(e));
    d3 nA = cross(x1 - x0, x2 - x0), nB = cross(x2 - x3, x1 - x3);
    double AA = sq(nA), AB = sq(nB);
    put_vec(g, 0, (-L / AA) * nA);
    put_vec(g, 3, (-L / AB) * nB);
    put_vec(g, 1, (-dot(e, x0 - x2) / (L * AA)) * nA + (-dot(e, x3 - x2) / (L * AB)) * nB);
    put_vec(g, 2, (dot(e, x0 - x1) / (L * AA)) * nA + (dot(e, x3 - x1) / (L * AB)) * nB);
}
// Hessian of theta: derivative of the four gradient blocks g_i = sA_i nA + sB_i nB
CIPC_DEV void dihedral_angle_hess(d3 x0, d3 x1, d3 x2, d3 x3, double* H) {
    d3 e = x2 - x1;
    double L = sqrt(sq(e)), iL = 1.0 / L;
    d3 nA = cross(x1 - x0, x2 - x0), nB = cross(x2 - x3, x1 - x3);
    double AA = sq(nA), AB = sq(nB);
    d3 zero = {0, 0, 0};
    d3 vA[4] = {x2 - x1, x0 - x2, x1 - x0, zero};
    d3 vB[4] = {zero, x2 - x3, x3 - x1, x1 - x2};
    d3 gL[4] = {zero, (-iL) * e, iL * e, zero};
    d3 gAA[4], gAB[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { gAA[j] = 2.0 * cross(nA, vA[j]); gAB[j] = 2.0 * cross(nB, vB[j]); }
    double PA1 = dot(e, x0 - x2), PB1 = dot(e, x3 - x2), PA2 = dot(e, x0 - x1), PB2 = dot(e, x3 - x1);
    d3 gPA1[4] = {e, -1.0 * (x0 - x2), (x0 - x2) - e, zero};
    d3 gPB1[4] = {zero, -1.0 * (x3 - x2), (x3 - x2) - e, e};
    d3 gPA2[4] = {e, -1.0 * (x0 - x1) - e, x0 - x1, zero};
    d3 gPB2[4] = {zero, -1.0 * (x3 - x1) - e, x3 - x1, e};
    double sA[4], sB[4];
    d3 gsA[4][4], gsB[4][4];
    sA[0] = -L / AA; sB[0] = 0; sA[3] = 0; sB[3] = -L / AB;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        gsA[0][j] = (-1.0 / AA) * gL[j] + (L / (AA * AA)) * gAA[j]; gsB[0][j] = zero;
        gsB[3][j] = (-1.0 / AB) * gL[j] + (L / (AB * AB)) * gAB[j]; gsA[3][j] = zero;
    }
    // s = sgn P / (L A): grad = sgn (gP / (L A) - P (A gL + L gA) / (L A)^2)
    {
        double LA = L * AA, iLA = 1.0 / LA, c = PA1 * iLA * iLA;
        sA[1] = -PA1 * iLA;
#pragma unroll
        for (int j = 0; j < 4; ++j) gsA[1][j] = -1.0 * (iLA * gPA1[j] - c * (AA * gL[j] + L * gAA[j]));
        c = PA2 * iLA * iLA;
        sA[2] = PA2 * iLA;
#pragma unroll
        for (int j = 0; j < 4; ++j) gsA[2][j] = iLA * gPA2[j] - c * (AA * gL[j] + L * gAA[j]);
    }
    {
        double LB = L * AB, iLB = 1.0 / LB, c = PB1 * iLB * iLB;
        sB[1] = -PB1 * iLB;
#pragma unroll
        for (int j = 0; j < 4; ++j) gsB[1][j] = -1.0 * (iLB * gPB1[j] - c * (AB * gL[j] + L * gAB[j]));
        c = PB2 * iLB * iLB;
        sB[2] = PB2 * iLB;
#pragma unroll
        for (int j = 0; j < 4; ++j) gsB[2][j] = iLB * gPB2[j] - c * (AB * gL[j] + L * gAB[j]);
    }
    double nAv[3] = {nA.x, nA.y, nA.z}, nBv[3] = {nB.x, nB.y, nB.z};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double CA[9], CB[9];
            cpm(vA[j], CA); cpm(vB[j], CB);
            double ga[3] = {gsA[i][j].x, gsA[i][j].y, gsA[i][j].z};
            double gb[3] = {gsB[i][j].x, gsB[i][j].y, gsB[i][j].z};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    H[(i * 3 + a) * 12 + j * 3 + b] = nAv[a] * ga[b] + sA[i] * CA[a * 3 + b] + nBv[a] * gb[b] + sB[i] * CB[a * 3 + b];
        }
#pragma unroll
    for (int i = 0; i < 12; ++i)
#pragma unroll
        for (int j = i + 1; j < 12; ++j) {
            double s = 0.5 * (H[i * 12 + j] + H[j * 12 + i]);
            H[i * 12 + j] = s; H[j * 12 + i] = s;
        }
}

// Row r (vertex r / 3, component r % 3) of the dihedral angle's Hessian before the
// symmetrisation done by dihedral_angle_hess: the same closed form, evaluated for one
// vertex of the stencil (used by the lane-cooperative hinge kernel, one row per lane).
CIPC_DEV void dihedral_hess_row(d3 x0, d3 x1, d3 x2, d3 x3, int r, double* row) {
    const int i = r / 3, a = r % 3;
    d3 e = x2 - x1;
    double L = sqrt(sq(e)), iL = 1.0 / L;
    d3 nA = cross(x1 - x0, x2 - x0), nB = cross(x2 - x3, x1 - x3);
    double AA = sq(nA), AB = sq(nB);
    d3 zero = {0, 0, 0};
    d3 vA[4] = {x2 - x1, x0 - x2, x1 - x0, zero};
    d3 vB[4] = {zero, x2 - x3, x3 - x1, x1 - x2};
    d3 gL[4] = {zero, (-iL) * e, iL * e, zero};
    d3 gAA[4], gAB[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { gAA[j] = 2.0 * cross(nA, vA[j]); gAB[j] = 2.0 * cross(nB, vB[j]); }
    double sA = 0, sB = 0;
    d3 gsA[4] = {zero, zero, zero, zero}, gsB[4] = {zero, zero, zero, zero};
    if (i == 0) {
        sA = -L / AA;
#pragma unroll
        for (int j = 0; j < 4; ++j) gsA[j] = (-1.0 / AA) * gL[j] + (L / (AA * AA)) * gAA[j];
    } else if (i == 3) {
        sB = -L / AB;
#pragma unroll
        for (int j = 0; j < 4; ++j) gsB[j] = (-1.0 / AB) * gL[j] + (L / (AB * AB)) * gAB[j];
    } else {
        const double LA = L * AA, iLA = 1.0 / LA, LB = L * AB, iLB = 1.0 / LB;
        if (i == 1) {
            double PA1 = dot(e, x0 - x2), PB1 = dot(e, x3 - x2);
            d3 gPA1[4] = {e, -1.0 * (x0 - x2), (x0 - x2) - e, zero};
            d3 gPB1[4] = {zero, -1.0 * (x3 - x2), (x3 - x2) - e, e};
            double cA = PA1 * iLA * iLA, cB = PB1 * iLB * iLB;
            sA = -PA1 * iLA;
            sB = -PB1 * iLB;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                gsA[j] = -1.0 * (iLA * gPA1[j] - cA * (AA * gL[j] + L * gAA[j]));
                gsB[j] = -1.0 * (iLB * gPB1[j] - cB * (AB * gL[j] + L * gAB[j]));
            }
        } else {
            double PA2 = dot(e, x0 - x1), PB2 = dot(e, x3 - x1);
            d3 gPA2[4] = {e, -1.0 * (x0 - x1) - e, x0 - x1, zero};
            d3 gPB2[4] = {zero, -1.0 * (x3 - x1) - e, x3 - x1, e};
            double cA = PA2 * iLA * iLA, cB = PB2 * iLB * iLB;
            sA = PA2 * iLA;
            sB = PB2 * iLB;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                gsA[j] = iLA * gPA2[j] - cA * (AA * gL[j] + L * gAA[j]);
                gsB[j] = iLB * gPB2[j] - cB * (AB * gL[j] + L * gAB[j]);
            }
        }
    }
    const double na = a == 0 ? nA.x : (a == 1 ? nA.y : nA.z);
    const double nb = a == 0 ? nB.x : (a == 1 ? nB.y : nB.z);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double CA[9], CB[9];
        cpm(vA[j], CA); cpm(vB[j], CB);
        const double ga[3] = {gsA[j].x, gsA[j].y, gsA[j].z}, gb[3] = {gsB[j].x, gsB[j].y, gsB[j].z};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double ca = a == 0 ? CA[b] : (a == 1 ? CA[3 + b] : CA[6 + b]);
            const double cb = a == 0 ? CB[b] : (a == 1 ? CB[3 + b] : CB[6 + b]);
            row[j * 3 + b] = na * ga[b] + sA * ca + nb * gb[b] + sB * cb;
        }
    }
}

// ------------------------------------------------------------- rods
// Segment spring, Rod/ROD_STRETCHING.h:10-107: E = Er / (2 lbar) (l - lbar)^2 with
// l = |x0 - x1| (the caller multiplies by h^2 pi t^2 / 4).  Gradient, and the Hessian
// already projected to PSD in closed form: [[K, -K], [-K, K]] with
// K = Er / lbar (n n^T + max(0, 1 - lbar / l) (I - n n^T)) -- the block structure has
// eigenvalues 2 Er / lbar, 2 Er / lbar (1 - lbar / l) (twice) and 0 (three times), so
// clamping the negative ones (makePD, Math/UTILS.h:10-27) is exactly this.
CIPC_DEV double rod_spring_energy(d3 x0, d3 x1, double Er, double lbar) {
    double l = sqrt(sq(x0 - x1));
    double r = l / lbar - 1;
    return lbar * Er / 2 * r * r;
}
CIPC_DEV void rod_spring_grad(d3 x0, d3 x1, double Er, double lbar, double* g) {
    d3 d = x0 - x1;
    double l = sqrt(sq(d));
    double s = Er / lbar * (l - lbar) / l;
    g[0] = s * d.x; g[1] = s * d.y; g[2] = s * d.z;
    g[3] = -g[0]; g[4] = -g[1]; g[5] = -g[2];
}
// K (3x3, row-major) of the projected Hessian
CIPC_DEV void rod_spring_hess_pd(d3 x0, d3 x1, double Er, double lbar, double* K) {
    d3 d = x0 - x1;
    double l = sqrt(sq(d));
    double n[3] = {d.x / l, d.y / l, d.z / l};
    double k = Er / lbar, perp = dmax(Er / lbar * (l - lbar) / l, 0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) K[3 * i + j] = (k - perp) * n[i] * n[j] + (i == j ? perp : 0.0);
}

// Rod hinge, Rod/ROD_BENDING.h:10-117: |kappa|^2 with kappa = 2 e0 x e1 / (|e0||e1| + e0.e1),
// e0 = x1 - x0, e1 = x2 - x1.  Derivatives through |kappa|^2 = 4 (u - v) / (u + v),
// u = |e0||e1|, v = e0.e1.
CIPC_DEV double rod_kappa2(d3 x0, d3 x1, d3 x2) {
    d3 e0 = x1 - x0, e1 = x2 - x1;
    double den = sqrt(sq(e0) * sq(e1)) + dot(e0, e1);
    d3 k = (2 / den) * cross(e0, e1);
    return sq(k);
}
CIPC_DEV void rod_kappa2_grad(d3 x0, d3 x1, d3 x2, double* g) {
    d3 e0 = x1 - x0, e1 = x2 - x1;
    double a = sqrt(sq(e0)), b = sqrt(sq(e1)), c = dot(e0, e1);
    double s = a * b + c, s2 = s * s;
    double fu = 8 * c / s2, fv = -8 * (a * b) / s2;
    d3 g0 = (fu * b / a) * e0 + fv * e1;
    d3 g1 = (fu * a / b) * e1 + fv * e0;
    g[0] = -g0.x; g[1] = -g0.y; g[2] = -g0.z;
    g[3] = g0.x - g1.x; g[4] = g0.y - g1.y; g[5] = g0.z - g1.z;
    g[6] = g1.x; g[7] = g1.y; g[8] = g1.z;
}
CIPC_DEV void rod_kappa2_hess(d3 x0, d3 x1, d3 x2, double* H) {
    d3 e0v = x1 - x0, e1v = x2 - x1;
    double e[2][3] = {{e0v.x, e0v.y, e0v.z}, {e1v.x, e1v.y, e1v.z}};
    double a = sqrt(sq(e0v)), b = sqrt(sq(e1v)), c = dot(e0v, e1v);
    double u = a * b, s = u + c, s2 = s * s, s3 = s2 * s;
    double fu = 8 * c / s2, fv = -8 * u / s2, fuu = -16 * c / s3, fvv = 16 * u / s3, fuv = 8 * (u - c) / s3;
    double len[2] = {a, b};
    double gu[6], gv[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        gu[i] = b / a * e[0][i]; gu[3 + i] = a / b * e[1][i];
        gv[i] = e[1][i]; gv[3 + i] = e[0][i];
    }
    double He[36];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int bi = i / 3, bj = j / 3, k = i % 3, l = j % 3;
            double d2u;
            if (bi == bj) {
                double lo = len[1 - bi], li = len[bi];
                d2u = lo * ((k == l ? 1.0 : 0.0) / li - e[bi][k] * e[bi][l] / (li * li * li));
            } else {
                d2u = e[bi][k] / len[bi] * e[bj][l] / len[bj];
            }
            double d2v = (bi != bj && k == l) ? 1.0 : 0.0;
            He[6 * i + j] = fu * d2u + fv * d2v + fuu * gu[i] * gu[j] + fvv * gv[i] * gv[j] + fuv * (gu[i] * gv[j] + gv[i] * gu[j]);
        }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            double H00 = He[6 * k + l], H01 = He[6 * k + 3 + l], H10 = He[6 * (3 + k) + l], H11 = He[6 * (3 + k) + 3 + l];
            H[9 * k + l] = H00;                  H[9 * k + 3 + l] = -H00 + H01;                   H[9 * k + 6 + l] = -H01;
            H[9 * (3 + k) + l] = -H00 + H10;     H[9 * (3 + k) + 3 + l] = H00 - H01 - H10 + H11;  H[9 * (3 + k) + 6 + l] = H01 - H11;
            H[9 * (6 + k) + l] = -H10;           H[9 * (6 + k) + 3 + l] = H10 - H11;              H[9 * (6 + k) + 6 + l] = H11;
        }
}

}  // namespace cipc
