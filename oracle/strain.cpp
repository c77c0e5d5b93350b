// Oracle: isotropic strain limiting of the shell (barrier on the singular values of the
// in-plane deformation gradient).
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates Compute_Inextensibility_Energy / _Gradient / _Hessian (FEM/Shell/INEXT.h:
// 120-467, the default C2 barrier f_sl of INEXT.h:5-69 without C1SL) and its use in the
// step: the potential is infinite (invalid) once a singular value reaches the limit
// s[0] (or falls to s[1] when the lower limit is on), the line search first halves the
// step until the potential is valid (IMPLICIT_EULER.h:74-92) and rejects invalid trials
// (:131-146).  F = A B^-1 with A = [x1 - x0, x2 - x0] and B the upper-triangular factor
// of the rest first fundamental form (B^T B = I_rest, INEXT.h:145-151); triangles whose
// three nodes are Dirichlet are skipped.  All terms carry h^2 vol kappa_s.
//
// Formulation used here: E(F) = sum_i phi(sigma_i); with F = U S V^T (thin SVD from
// the 2x2 eigenproblem of F^T F, u2 = u0 x u1) the Hessian in F is diagonal in the
// orthonormal directions u0 v0^T, u1 v1^T (phi''), twist/flip (u0 v1^T -+ u1 v0^T)/sqrt2
// ((phi'0 + phi'1)/(s0 + s1), (phi'0 - phi'1)/(s0 - s1)) and u2 v_b^T (phi'_b / s_b).
// The projected Hessian (Math/UTILS.h makePD on the 6x6 dE/dF^2, INEXT.h:437-439) is
// formed here by the general symmetric eigensolver of geom.cpp.
// Parity pins (tests/test_oracle_strain.py): energy of uniform / uniaxial stretch in
// closed form, gradient = central difference of the energy, unprojected Hessian =
// central difference of the gradient, PSD after projection, validity at the limit.
// Parity unpinned: the reference's regularised singular-vector derivative when the two
// singular values are within 2e-6 (INEXT.h:313-319); here the exact limit is used.
#include "oracle.h"

namespace orc {

static inline V3 X3(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

void strain_limit_setup(Scene& S) {
    S.slIB.resize(3 * S.nT);
    for (int t = 0; t < S.nT; ++t) {
        const int* v = &S.tri[3 * t];
        V3 a = X3(S.X0.data(), v[0]), b = X3(S.X0.data(), v[1]), c = X3(S.X0.data(), v[2]);
        V3 e1 = b - a, e2 = c - a;
        double I00 = sq(e1), I01 = dot(e1, e2), I11 = sq(e2);
        double B00 = std::sqrt(I00), B01 = I01 / B00, B11 = std::sqrt(I11 - I01 * I01 / I00);
        double inv = 1.0 / (B00 * B11);
        S.slIB[3 * t] = B11 * inv;        // (B^-1)_00
        S.slIB[3 * t + 1] = -B01 * inv;   // (B^-1)_01
        S.slIB[3 * t + 2] = B00 * inv;    // (B^-1)_11
    }
}

// barrier of INEXT.h:5-38 (C2 form): value, first and second derivative in sigma
static void f_sl(double sg, double slim, double shat, double* f, double* g, double* hh) {
    const double a = shat - sg, b = slim - sg, D = (slim - shat) * (slim - shat);
    const double L = std::log(b / (slim - shat));
    if (f) *f = -a * a * L / D;
    if (g) *g = (2 * a * L + a * a / b) / D;
    if (hh) { const double c = a / b; *hh = (-2 * L + c * (c - 4)) / D; }
}

struct SLTri {
    double F[3][2];
    double sig[2], V[2][2];   // V columns = right singular vectors
    V3 u[3];
};

static void sl_svd(const Scene& S, const double* x, int t, SLTri& T) {
    const int* v = &S.tri[3 * t];
    V3 x0 = X3(x, v[0]), x1 = X3(x, v[1]), x2 = X3(x, v[2]);
    V3 A0 = x1 - x0, A1 = x2 - x0;
    const double i00 = S.slIB[3 * t], i01 = S.slIB[3 * t + 1], i11 = S.slIB[3 * t + 2];
    V3 F0 = i00 * A0, F1 = i01 * A0 + i11 * A1;
    T.F[0][0] = F0.x; T.F[1][0] = F0.y; T.F[2][0] = F0.z;
    T.F[0][1] = F1.x; T.F[1][1] = F1.y; T.F[2][1] = F1.z;
    // F^T F = [[a, b], [b, d]]
    const double a = sq(F0), b = dot(F0, F1), d = sq(F1);
    const double hm = 0.5 * (a - d), r = std::sqrt(hm * hm + b * b);
    const double l0 = 0.5 * (a + d) + r, l1 = std::max(0.5 * (a + d) - r, 0.0);
    T.sig[0] = std::sqrt(l0);
    T.sig[1] = std::sqrt(l1);
    double vx, vy;   // eigenvector of l0
    if (r == 0) { vx = 1; vy = 0; }
    else if (hm >= 0) { vx = hm + r; vy = b; }
    else { vx = b; vy = r - hm; }
    const double n = std::sqrt(vx * vx + vy * vy);
    vx /= n; vy /= n;
    T.V[0][0] = vx; T.V[1][0] = vy;     // v0
    T.V[0][1] = -vy; T.V[1][1] = vx;    // v1
    T.u[0] = (1.0 / T.sig[0]) * (vx * F0 + vy * F1);
    V3 w = (-vy) * F0 + vx * F1;
    T.u[1] = T.sig[1] > 0 ? (1.0 / T.sig[1]) * w : V3{0, 0, 0};
    V3 nrm = cross(T.u[0], T.u[1]);
    T.u[2] = (1.0 / std::sqrt(sq(nrm))) * nrm;
}

// phi and derivatives of one singular value (upper and, if on, lower barrier)
static bool sl_phi(const Scene& S, double sg, double* f, double* g, double* hh) {
    const Params& P = S.P;
    *f = *g = *hh = 0;
    if (sg >= P.slS[0]) return false;
    if (sg > P.slSHat[0]) {
        double a, b, c;
        f_sl(sg, P.slS[0], P.slSHat[0], &a, &b, &c);
        *f += P.slKappa[0] * a; *g += P.slKappa[0] * b; *hh += P.slKappa[0] * c;
    }
    if (P.slKappa[1] > 0) {
        if (sg <= P.slS[1]) return false;
        if (sg < P.slSHat[1]) {
            double a, b, c;
            f_sl(sg, P.slS[1], P.slSHat[1], &a, &b, &c);
            *f += P.slKappa[1] * a; *g += P.slKappa[1] * b; *hh += P.slKappa[1] * c;
        }
    }
    return true;
}

static inline bool tri_all_dbc(const Scene& S, int t) {
    const int* v = &S.tri[3 * t];
    return S.isDBC[v[0]] && S.isDBC[v[1]] && S.isDBC[v[2]];
}

// one triangle: energy, gradient (9) and Hessian (81, projected if proj) in x; false if invalid
bool strain_limit_tri(const Scene& S, const double* x, int t, double* E, double* g9, double* H81, bool proj, double* sigOut) {
    SLTri T;
    sl_svd(S, x, t, T);
    if (sigOut) { sigOut[0] = T.sig[0]; sigOut[1] = T.sig[1]; }
    const double w = shell_h2(S) * S.vol[t];
    double f[2], g[2], hh[2];
    for (int i = 0; i < 2; ++i)
        if (!sl_phi(S, T.sig[i], &f[i], &g[i], &hh[i])) return false;
    if (E) *E = w * (f[0] + f[1]);
    const double i00 = S.slIB[3 * t], i01 = S.slIB[3 * t + 1], i11 = S.slIB[3 * t + 2];
    // dvec(F)/dx per vertex: column 0 weight c0[a], column 1 weight c1[a]
    const double c0[3] = {-i00, i00, 0.0}, c1[3] = {-(i01 + i11), i01, i11};
    auto dirF = [&](int ua, int vb, double* out6) {   // vec(u_a v_b^T), columns stacked
        const V3 u = T.u[ua];
        const double vb0 = T.V[0][vb], vb1 = T.V[1][vb];
        out6[0] = u.x * vb0; out6[1] = u.y * vb0; out6[2] = u.z * vb0;
        out6[3] = u.x * vb1; out6[4] = u.y * vb1; out6[5] = u.z * vb1;
    };
    if (g9) {
        double G[6] = {0, 0, 0, 0, 0, 0}, d[6];
        for (int i = 0; i < 2; ++i) {
            dirF(i, i, d);
            for (int k = 0; k < 6; ++k) G[k] += w * g[i] * d[k];
        }
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) g9[3 * a + k] = c0[a] * G[k] + c1[a] * G[3 + k];
    }
    if (H81) {
        double HF[36] = {0};
        auto add = [&](const double* d, double lam) {
            for (int r = 0; r < 6; ++r)
                for (int q = 0; q < 6; ++q) HF[r * 6 + q] += lam * d[r] * d[q];
        };
        double d00[6], d11[6], d01[6], d10[6], d20[6], d21[6], tw[6], fl[6];
        dirF(0, 0, d00); dirF(1, 1, d11); dirF(0, 1, d01); dirF(1, 0, d10); dirF(2, 0, d20); dirF(2, 1, d21);
        const double s0 = T.sig[0], s1 = T.sig[1];
        for (int k = 0; k < 6; ++k) { tw[k] = (d01[k] - d10[k]) / std::sqrt(2.0); fl[k] = (d01[k] + d10[k]) / std::sqrt(2.0); }
        const double lt = (g[0] + g[1]) / (s0 + s1);
        const double lf = std::fabs(s0 - s1) > 1e-12 * s0 ? (g[0] - g[1]) / (s0 - s1) : 0.5 * (hh[0] + hh[1]);
        add(d00, hh[0]); add(d11, hh[1]); add(tw, lt); add(fl, lf); add(d20, g[0] / s0); add(d21, g[1] / s1);
        for (double& v : HF) v *= w;
        if (proj) make_pd(6, HF);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l)
                        H81[(3 * a + k) * 9 + 3 * b + l] = c0[a] * c0[b] * HF[k * 6 + l] + c0[a] * c1[b] * HF[k * 6 + 3 + l] +
                                                           c1[a] * c0[b] * HF[(3 + k) * 6 + l] + c1[a] * c1[b] * HF[(3 + k) * 6 + 3 + l];
    }
    return true;
}

static inline bool sl_on(const Scene& S) { return S.P.slKappa[0] > 0; }

bool strain_limit_energy(const Scene& S, const double* x, double* E) {
    if (!sl_on(S)) return true;
    bool valid = true;
    double sum = 0;
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t)) continue;
        double e = 0;
        if (!strain_limit_tri(S, x, t, &e, nullptr, nullptr, false, nullptr)) { valid = false; continue; }
        sum += e;
    }
    *E += sum;
    return valid;
}

// largest singular value of each triangle's F (1 for a triangle whose three nodes are
// Dirichlet), Compute_Inextensibility (INEXT.h:72-114)
void strain_max_sv(const Scene& S, const double* x, std::vector<double>& smax) {
    smax.assign(S.nT, 1.0);
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t)) continue;
        SLTri T;
        sl_svd(S, x, t, T);
        smax[t] = T.sig[0];
    }
}

void strain_limit_grad(const Scene& S, const double* x, double* g) {
    if (!sl_on(S)) return;
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t)) continue;
        double gl[9];
        if (!strain_limit_tri(S, x, t, nullptr, gl, nullptr, false, nullptr)) continue;
        const int* v = &S.tri[3 * t];
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) g[3 * v[a] + k] += gl[3 * a + k];
    }
}

void strain_limit_hess(const Scene& S, const double* x, std::vector<Block>& out) {
    if (!sl_on(S)) return;
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t)) continue;
        double H[81];
        if (!strain_limit_tri(S, x, t, nullptr, nullptr, H, true, nullptr)) continue;
        const int* v = &S.tri[3 * t];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                Block B;
                B.r = v[a]; B.c = v[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) B.b[3 * k + l] = H[(3 * a + k) * 9 + 3 * b + l];
                out.push_back(B);
            }
    }
}

// ---------------------------------------------------------------- anisotropic (fibre) model
// Restates Check_Fiber_Feasibility / Compute_Fiber_Energy / _Gradient / _Hessian
// (FEM/Shell/ANISO_INEXT.h:33-440; eta of :5-31 without ORTHO_STVK).  It replaces the
// membrane when fiberStiffMult[0] > 0 or [1] > 0 (INC_POTENTIAL.h:81-92, 213-219,
// 340-346).  Green strain E = (F^T F - I)/2 of F = A B^-1, B from the rest positions'
// x and z coordinates (the sheet's rest plane, ANISO_INEXT.h:64-68);
// psi = h^2 vol [k0/2 eta(E00^2, L0^2) + k1 eta(E00 E11, L0 L1) + k2/2 eta(E11^2, L1^2)
//                + k3 eta(E01^2, L2^2)],  eta(s, m) = -m ln((m - s)/m);
// infeasible when |E00| >= L0, |E11| >= L1 or |E01| >= L2.  The F-space Hessian is
// sum_ab d2psi/dEa dEb dEa dEb^T + [[S00 I, S01/2 I], [S01/2 I, S11 I]] (own derivation),
// projected to PSD in F space (ANISO_INEXT.h:322-324) and mapped to x.
// Parity pins (tests/test_oracle_strain.py): closed-form energies of uniaxial stretch and
// shear, central differences of energy and gradient, feasibility at the limits.

static inline bool fiber_on(const Scene& S) { return S.P.fiberK[0] > 0 || S.P.fiberK[1] > 0; }

void fiber_setup(Scene& S) {
    S.fiberIB.resize(4 * S.nT);
    for (int t = 0; t < S.nT; ++t) {
        const int* v = &S.tri[3 * t];
        const double* X = S.X0.data();
        const double b00 = X[3 * v[1]] - X[3 * v[0]], b01 = X[3 * v[2]] - X[3 * v[0]];
        const double b10 = X[3 * v[1] + 2] - X[3 * v[0] + 2], b11 = X[3 * v[2] + 2] - X[3 * v[0] + 2];
        const double inv = 1.0 / (b00 * b11 - b01 * b10);
        S.fiberIB[4 * t] = b11 * inv; S.fiberIB[4 * t + 1] = -b01 * inv;     // (B^-1)_00, (B^-1)_01
        S.fiberIB[4 * t + 2] = -b10 * inv; S.fiberIB[4 * t + 3] = b00 * inv; // (B^-1)_10, (B^-1)_11
    }
}

static inline double eta0(double s, double m) { return -m * std::log((m - s) / m); }
static inline double eta1(double s, double m) { return m / (m - s); }
static inline double eta2(double s, double m) { return m / ((m - s) * (m - s)); }

// one triangle: false when infeasible; E, g (9), H (81, projected if proj)
bool fiber_tri(const Scene& S, const double* x, int t, double* E, double* g9, double* H81, bool proj, double* Eout3) {
    const Params& P = S.P;
    const int* v = &S.tri[3 * t];
    const double* ib = &S.fiberIB[4 * t];
    V3 A0 = X3(x, v[1]) - X3(x, v[0]), A1 = X3(x, v[2]) - X3(x, v[0]);
    V3 F0 = ib[0] * A0 + ib[2] * A1, F1 = ib[1] * A0 + ib[3] * A1;
    const double E00 = 0.5 * (sq(F0) - 1), E11 = 0.5 * (sq(F1) - 1), E01 = 0.5 * dot(F0, F1);
    if (Eout3) { Eout3[0] = E00; Eout3[1] = E11; Eout3[2] = E01; }
    const double L0 = P.fiberLimit[0], L1 = P.fiberLimit[1], L2 = P.fiberLimit[2];
    if (std::fabs(E00) >= L0 || std::fabs(E11) >= L1 || std::fabs(E01) >= L2) return false;
    const double* k = P.fiberK;
    const double w = shell_h2(S) * S.vol[t];
    const double m00 = L0 * L0, m11 = L1 * L1, m01 = L0 * L1, m2 = L2 * L2;
    const double s00 = E00 * E00, s11 = E11 * E11, sx = E00 * E11, s2 = E01 * E01;
    if (E) *E = w * (k[0] / 2 * eta0(s00, m00) + k[1] * eta0(sx, m01) + k[2] / 2 * eta0(s11, m11) + k[3] * eta0(s2, m2));
    const double S00 = w * (k[0] * eta1(s00, m00) * E00 + k[1] * eta1(sx, m01) * E11);
    const double S11 = w * (k[2] * eta1(s11, m11) * E11 + k[1] * eta1(sx, m01) * E00);
    const double S01 = w * 2 * k[3] * eta1(s2, m2) * E01;
    const double c0[3] = {-(ib[0] + ib[2]), ib[0], ib[2]}, c1[3] = {-(ib[1] + ib[3]), ib[1], ib[3]};
    if (g9) {
        V3 G0 = S00 * F0 + (0.5 * S01) * F1, G1 = S11 * F1 + (0.5 * S01) * F0;
        for (int a = 0; a < 3; ++a) {
            g9[3 * a] = c0[a] * G0.x + c1[a] * G1.x;
            g9[3 * a + 1] = c0[a] * G0.y + c1[a] * G1.y;
            g9[3 * a + 2] = c0[a] * G0.z + c1[a] * G1.z;
        }
    }
    if (H81) {
        const double H00 = w * (k[0] * (2 * eta2(s00, m00) * s00 + eta1(s00, m00)) + k[1] * eta2(sx, m01) * s11);
        const double H11 = w * (k[2] * (2 * eta2(s11, m11) * s11 + eta1(s11, m11)) + k[1] * eta2(sx, m01) * s00);
        const double H0x = w * k[1] * (eta2(sx, m01) * sx + eta1(sx, m01));
        const double H22 = w * 2 * k[3] * (2 * eta2(s2, m2) * s2 + eta1(s2, m2));
        double dA[6] = {F0.x, F0.y, F0.z, 0, 0, 0}, dB[6] = {0, 0, 0, F1.x, F1.y, F1.z};
        double dC[6] = {0.5 * F1.x, 0.5 * F1.y, 0.5 * F1.z, 0.5 * F0.x, 0.5 * F0.y, 0.5 * F0.z};
        double HF[36];
        for (int r = 0; r < 6; ++r)
            for (int q = 0; q < 6; ++q)
                HF[r * 6 + q] = H00 * dA[r] * dA[q] + H11 * dB[r] * dB[q] + H0x * (dA[r] * dB[q] + dB[r] * dA[q]) + H22 * dC[r] * dC[q];
        for (int i = 0; i < 3; ++i) {
            HF[i * 6 + i] += S00;
            HF[(3 + i) * 6 + 3 + i] += S11;
            HF[i * 6 + 3 + i] += 0.5 * S01;
            HF[(3 + i) * 6 + i] += 0.5 * S01;
        }
        if (proj) make_pd(6, HF);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int kk = 0; kk < 3; ++kk)
                    for (int l = 0; l < 3; ++l)
                        H81[(3 * a + kk) * 9 + 3 * b + l] = c0[a] * c0[b] * HF[kk * 6 + l] + c0[a] * c1[b] * HF[kk * 6 + 3 + l] +
                                                            c1[a] * c0[b] * HF[(3 + kk) * 6 + l] + c1[a] * c1[b] * HF[(3 + kk) * 6 + 3 + l];
    }
    return true;
}

bool fiber_active(const Scene& S) { return fiber_on(S); }

bool fiber_energy(const Scene& S, const double* x, double* E) {
    bool ok = true;
    double sum = 0;
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t)) continue;
        double e = 0;
        if (!fiber_tri(S, x, t, &e, nullptr, nullptr, false, nullptr)) { ok = false; continue; }
        sum += e;
    }
    *E += sum;
    return ok;
}
void fiber_grad(const Scene& S, const double* x, double* g) {
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t)) continue;
        double gl[9];
        if (!fiber_tri(S, x, t, nullptr, gl, nullptr, false, nullptr)) continue;
        const int* v = &S.tri[3 * t];
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) g[3 * v[a] + k] += gl[3 * a + k];
    }
}
void fiber_hess(const Scene& S, const double* x, std::vector<Block>& out) {
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t)) continue;
        double H[81];
        if (!fiber_tri(S, x, t, nullptr, nullptr, H, true, nullptr)) continue;
        const int* v = &S.tri[3 * t];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                Block B;
                B.r = v[a]; B.c = v[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) B.b[3 * k + l] = H[(3 * a + k) * 9 + 3 * b + l];
                out.push_back(B);
            }
    }
}

}  // namespace orc
