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
    const double w = S.P.dt * S.P.dt * S.vol[t];
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

}  // namespace orc
