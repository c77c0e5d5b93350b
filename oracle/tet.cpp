// Oracle: volumetric tetrahedra inside the shell step.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// tet_setup restates Compute_Vol_And_Inv_Basis (FEM/DEFORMATION_GRADIENT.h:39-77: rest
// basis Dm = [X1 - X0, X2 - X0, X3 - X0], vol = det(Dm) / 6, Dm^-1 kept), the lumped mass
// of Compute_Mass_And_Init_Velocity (FEM/ELEM_TO_NODE.h:11-46: vol rho / 4 per node) and
// the Lame parameters of NEOHOOKEAN_FUNCTOR (Physics/NEOHOOKEAN.h:46-47).
// tet_surface restates Find_Surface_TriMesh (Utils/MESHIO.h:486-544): every face of every
// tet keyed as written (v0 v2 v1, v0 v3 v2, v0 v1 v3, v1 v2 v3), ordered as a std::map of
// the three indices; a face is on the surface when no rotation of its reverse is a key.
// The Neo-Hookean energy (NEOHOOKEAN.h:78-93, 136-142), first Piola-Kirchhoff stress
// (NEOHOOKEAN.h:100-109) and the PSD-projected stress derivative (NEOHOOKEAN.h:116-300:
// d2E/dsigma2 and the three 2x2 twist/flip blocks clamped separately) are weighted by
// h^2 vol (INC_POTENTIAL.h:98-109, 226-229, 354-364); a tet with det F <= 0 makes the
// potential invalid (INC_POTENTIAL.h:99-107).
// inversion_free_stepsize restates Compute_Inversion_Free_StepSize (FEM/IPC.h:1785-1877)
// with the cubic root of IPC.h:1752-1782 (Cardano with complex arithmetic).
// Parity pins (tests/test_oracle_tet.py): energy of uniaxial / uniform stretch in closed
// form; gradient and unprojected Hessian by central differences; the projected stress
// derivative equals numpy's eigh projection of the finite-difference one; surfaces of
// small meshes against a face count; step sizes against brute-force determinant roots.
#include "oracle.h"
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <map>

namespace orc {

static inline V3 X3(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

static double det3(const double* M) {   // row-major 3x3
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}
static void inv3(const double* M, double* I) {
    const double d = det3(M);
    I[0] = (M[4] * M[8] - M[5] * M[7]) / d;
    I[1] = (M[2] * M[7] - M[1] * M[8]) / d;
    I[2] = (M[1] * M[5] - M[2] * M[4]) / d;
    I[3] = (M[5] * M[6] - M[3] * M[8]) / d;
    I[4] = (M[0] * M[8] - M[2] * M[6]) / d;
    I[5] = (M[2] * M[3] - M[0] * M[5]) / d;
    I[6] = (M[3] * M[7] - M[4] * M[6]) / d;
    I[7] = (M[1] * M[6] - M[0] * M[7]) / d;
    I[8] = (M[0] * M[4] - M[1] * M[3]) / d;
}

bool tet_setup(Scene& S) {
    const int nTet = (int)S.tet.size() / 4;
    const Params& P = S.P;
    S.tetIB.resize(9 * (size_t)nTet);
    S.tetVol.resize(nTet);
    S.tetLam = P.tetE * P.tetNu / ((1.0 + P.tetNu) * (1.0 - 2.0 * P.tetNu));
    S.tetMu = P.tetE / (2.0 * (1.0 + P.tetNu));
    for (int t = 0; t < nTet; ++t) {
        const int* v = &S.tet[4 * t];
        V3 X0 = X3(S.X0.data(), v[0]);
        V3 e[3] = {X3(S.X0.data(), v[1]) - X0, X3(S.X0.data(), v[2]) - X0, X3(S.X0.data(), v[3]) - X0};
        double Dm[9];
        for (int c = 0; c < 3; ++c) { Dm[c] = e[c].x; Dm[3 + c] = e[c].y; Dm[6 + c] = e[c].z; }
        const double vol = det3(Dm) / 6.0;
        if (!(vol > 0)) return false;   // rest tets must be positively oriented
        S.tetVol[t] = vol;
        inv3(Dm, &S.tetIB[9 * (size_t)t]);
        const double m = vol * P.tetRho / 4;
        for (int k = 0; k < 4; ++k) S.mass[v[k]] += m;
    }
    return true;
}

void tet_surface(const Scene& S, std::vector<int>& tri) {
    tri.clear();
    std::map<std::array<int, 3>, int> faces;
    const int nTet = (int)S.tet.size() / 4;
    for (int t = 0; t < nTet; ++t) {
        const int* v = &S.tet[4 * t];
        faces[{v[0], v[2], v[1]}] = t;
        faces[{v[0], v[3], v[2]}] = t;
        faces[{v[0], v[1], v[3]}] = t;
        faces[{v[1], v[2], v[3]}] = t;
    }
    for (const auto& f : faces) {
        const auto& k = f.first;
        if (faces.count({k[2], k[1], k[0]}) || faces.count({k[1], k[0], k[2]}) || faces.count({k[0], k[2], k[1]})) continue;
        tri.push_back(k[0]);
        tri.push_back(k[1]);
        tri.push_back(k[2]);
    }
}

// F = Ds Dm^-1 (row-major), Ds columns x1 - x0, x2 - x0, x3 - x0
static void tet_F(const Scene& S, const double* x, int t, double* F) {
    const int* v = &S.tet[4 * t];
    V3 x0 = X3(x, v[0]);
    V3 e[3] = {X3(x, v[1]) - x0, X3(x, v[2]) - x0, X3(x, v[3]) - x0};
    const double* B = &S.tetIB[9 * (size_t)t];
    const double Ds[9] = {e[0].x, e[1].x, e[2].x, e[0].y, e[1].y, e[2].y, e[0].z, e[1].z, e[2].z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[3 * i + j] = Ds[3 * i] * B[j] + Ds[3 * i + 1] * B[3 + j] + Ds[3 * i + 2] * B[6 + j];
}

// psi(F) = mu/2 (|F|^2 - 3) - mu ln J + lambda/2 ln^2 J
static double nh_psi(const double* F, double lam, double mu) {
    double f2 = 0;
    for (int k = 0; k < 9; ++k) f2 += F[k] * F[k];
    const double lj = std::log(det3(F));
    return mu / 2 * (f2 - 3) - (mu - lam / 2 * lj) * lj;
}

bool tets_valid(const Scene& S, const double* x) {
    const int nTet = (int)S.tet.size() / 4;
    for (int t = 0; t < nTet; ++t) {
        double F[9];
        tet_F(S, x, t, F);
        if (det3(F) <= 0) return false;
    }
    return true;
}

double tet_energy(const Scene& S, const double* x) {
    const double h2 = S.P.dt * S.P.dt;
    const int nTet = (int)S.tet.size() / 4;
    double E = 0;
    for (int t = 0; t < nTet; ++t) {
        double F[9];
        tet_F(S, x, t, F);
        E += h2 * S.tetVol[t] * nh_psi(F, S.tetLam, S.tetMu);
    }
    return E;
}

// dPsi/dx of one tet (12), weight w: P = w (mu (F - F^-T) + lambda ln J F^-T), g_a = P Dm^-T e_a
static void tet_grad_local(const Scene& S, const double* x, int t, double w, double* g) {
    double F[9], Fi[9];
    tet_F(S, x, t, F);
    inv3(F, Fi);
    const double lj = std::log(det3(F));
    double Pk[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double fit = Fi[3 * j + i];   // (F^-T)_ij
            Pk[3 * i + j] = w * (S.tetMu * (F[3 * i + j] - fit) + S.tetLam * lj * fit);
        }
    const double* B = &S.tetIB[9 * (size_t)t];
    for (int k = 0; k < 3; ++k) g[k] = 0;
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 3; ++i) {
            double s = 0;
            for (int j = 0; j < 3; ++j) s += Pk[3 * i + j] * B[3 * a + j];
            g[3 * (a + 1) + i] = s;
            g[i] -= s;
        }
}

void tet_grad(const Scene& S, const double* x, double* g) {
    const double h2 = S.P.dt * S.P.dt;
    const int nTet = (int)S.tet.size() / 4;
    for (int t = 0; t < nTet; ++t) {
        double gl[12];
        tet_grad_local(S, x, t, h2 * S.tetVol[t], gl);
        const int* v = &S.tet[4 * t];
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) g[3 * v[a] + k] += gl[3 * a + k];
    }
}

// SVD F = U diag(s) V^T from the eigen-decomposition of F^T F (V a rotation, s descending;
// U = F V diag(s)^-1 for the first two columns, u3 = u1 x u2, s3 = u3 . F v3 (negative
// for an inverted F))
static void svd3(const double* F, double* U, double* s, double* V) {
    double C[9], w[3], Q[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    sym_eig(3, C, w, Q);   // ascending
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) V[3 * r + c] = Q[3 * r + (2 - c)];
    if (det3(V) < 0)
        for (int r = 0; r < 3; ++r) V[3 * r + 2] = -V[3 * r + 2];
    V3 u[3];
    for (int c = 0; c < 2; ++c) {
        const V3 vc{V[c], V[3 + c], V[6 + c]};
        const V3 fv{F[0] * vc.x + F[1] * vc.y + F[2] * vc.z, F[3] * vc.x + F[4] * vc.y + F[5] * vc.z,
                    F[6] * vc.x + F[7] * vc.y + F[8] * vc.z};
        s[c] = std::sqrt(sq(fv));
        u[c] = (1.0 / s[c]) * fv;
    }
    u[2] = cross(u[0], u[1]);
    const V3 v2{V[2], V[5], V[8]};
    const V3 fv2{F[0] * v2.x + F[1] * v2.y + F[2] * v2.z, F[3] * v2.x + F[4] * v2.y + F[5] * v2.z, F[6] * v2.x + F[7] * v2.y + F[8] * v2.z};
    s[2] = dot(u[2], fv2);
    for (int c = 0; c < 3; ++c) { U[c] = u[c].x; U[3 + c] = u[c].y; U[6 + c] = u[c].z; }
}

// d2 psi / dF2 (9x9, vec(F) column-major: index i + 3 j), projected (or not) block-wise in
// the singular basis as NEOHOOKEAN.h:192-300 does; scaled by w
void nh_dPdF(const double* F, double lam, double mu, double w, bool project, double* D) {
    double U[9], s[3], V[9];
    svd3(F, U, s, V);
    const double ls = std::log(s[0] * s[1] * s[2]);
    double dE[3], A[9];
    for (int i = 0; i < 3; ++i) {
        const double inv = 1.0 / s[i];
        dE[i] = mu * (s[i] - inv) + lam * inv * ls;
        A[4 * i] = mu * (1.0 + inv * inv) - lam * inv * inv * (ls - 1.0);
    }
    A[1] = A[3] = lam / s[0] / s[1];
    A[5] = A[7] = lam / s[1] / s[2];
    A[2] = A[6] = lam / s[2] / s[0];
    if (project) make_pd(3, A);
    // twist / flip pairs (i, i+1 mod 3): B = [[l + r, l - r], [l - r, l + r]], eigenvalues
    // 2l on (1, 1) and 2r on (1, -1)
    const double mid = mu - lam * ls;
    double Bd[3], Bo[3];
    for (int c = 0; c < 3; ++c) {
        const int d = (c + 1) % 3;
        double l = (mu + mid / s[c] / s[d]) / 2.0;
        double r = dE[c] + dE[d];
        const double sum = s[c] + s[d];
        r /= (sum < 1.0e-6 ? 2 * 1.0e-6 : 2 * sum);
        if (project) {
            const double e1 = std::max(2 * l, 0.0), e2 = std::max(2 * r, 0.0);
            l = e1 / 2;
            r = e2 / 2;
        }
        Bd[c] = l + r;
        Bo[c] = l - r;
    }
    // M in the basis E_ab = u_a v_b^T (vec index a + 3 b)
    double M[81] = {0};
    auto at = [&](int a, int b, int c, int d) -> double& { return M[(a + 3 * b) * 9 + (c + 3 * d)]; };
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) at(a, a, c, c) = A[3 * a + c];
    for (int c = 0; c < 3; ++c) {
        const int d = (c + 1) % 3;
        at(c, d, c, d) = Bd[c];
        at(d, c, d, c) = Bd[c];
        at(c, d, d, c) = Bo[c];
        at(d, c, c, d) = Bo[c];
    }
    // D(ij, rs) = sum M(ab, cd) U_ia V_jb U_rc V_sd
    double Q[81];   // Q[(i + 3 j) * 9 + (a + 3 b)] = U_ia V_jb
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) Q[(i + 3 * j) * 9 + (a + 3 * b)] = U[3 * i + a] * V[3 * j + b];
    double T[81];
    for (int p = 0; p < 9; ++p)
        for (int q = 0; q < 9; ++q) {
            double acc = 0;
            for (int k = 0; k < 9; ++k) acc += Q[p * 9 + k] * M[k * 9 + q];
            T[p * 9 + q] = acc;
        }
    for (int p = 0; p < 9; ++p)
        for (int q = 0; q < 9; ++q) {
            double acc = 0;
            for (int k = 0; k < 9; ++k) acc += T[p * 9 + k] * Q[q * 9 + k];
            D[p * 9 + q] = w * acc;
        }
}

// 12x12 Hessian of one tet from the 9x9 stress derivative: H = D^T dPdF D, with
// dF_ij / dx_a,k = delta_ik c_a(j), c_a(j) = Dm^-1(a - 1, j) for a = 1..3, c_0 = -sum
void tet_hess_local(const Scene& S, const double* x, int t, double w, bool project, double* H) {
    double F[9], D[81];
    tet_F(S, x, t, F);
    nh_dPdF(F, S.tetLam, S.tetMu, w, project, D);
    const double* B = &S.tetIB[9 * (size_t)t];
    double c[4][3];
    for (int j = 0; j < 3; ++j) {
        c[1][j] = B[j]; c[2][j] = B[3 + j]; c[3][j] = B[6 + j];
        c[0][j] = -(c[1][j] + c[2][j] + c[3][j]);
    }
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k)
            for (int b = 0; b < 4; ++b)
                for (int l = 0; l < 3; ++l) {
                    double acc = 0;
                    for (int j = 0; j < 3; ++j)
                        for (int s = 0; s < 3; ++s) acc += c[a][j] * c[b][s] * D[(k + 3 * j) * 9 + (l + 3 * s)];
                    H[(3 * a + k) * 12 + 3 * b + l] = acc;
                }
}

void tet_hess(const Scene& S, const double* x, std::vector<Block>& out) {
    const double h2 = S.P.dt * S.P.dt;
    const int nTet = (int)S.tet.size() / 4;
    const long base = (long)out.size();
    out.resize(base + 16l * nTet);
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (int t = 0; t < nTet; ++t) {
        double H[144];
        tet_hess_local(S, x, t, h2 * S.tetVol[t], S.P.projectPD != 0, H);
        const int* v = &S.tet[4 * t];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                Block& Bk = out[base + 16l * t + 4 * a + b];
                Bk.r = v[a]; Bk.c = v[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) Bk.b[3 * k + l] = H[(3 * a + k) * 12 + 3 * b + l];
            }
    }
}

// ---------------------------------------------------------------- inversion-free step size
// smallest positive real root of a t^2 + b t + c, negative if none (IPC.h:1731-1749)
static double quad_root(double a, double b, double c, double tol) {
    double t;
    if (std::abs(a) <= tol) t = -c / b;
    else {
        const double desc = b * b - 4 * a * c;
        if (desc > 0) {
            t = (-b - std::sqrt(desc)) / (2 * a);
            if (t < 0) t = (-b + std::sqrt(desc)) / (2 * a);
        } else t = -1;
    }
    return t;
}
// smallest positive real root of a t^3 + b t^2 + c t + d, negative if none (IPC.h:1751-1782)
double cubic_root(double a, double b, double c, double d, double tol) {
    if (std::abs(a) <= tol) return quad_root(b, c, d, tol);
    typedef std::complex<double> cd;
    const cd i(0, 1);
    const cd d0(b * b - 3 * a * c, 0), d1(2 * b * b * b - 9 * a * b * c + 27 * a * a * d, 0);
    cd C = std::pow((d1 + std::sqrt(d1 * d1 - 4.0 * d0 * d0 * d0)) / 2.0, 1.0 / 3.0);
    if (std::abs(C) == 0.0) C = std::pow((d1 - std::sqrt(d1 * d1 - 4.0 * d0 * d0 * d0)) / 2.0, 1.0 / 3.0);
    const cd u2 = (-1.0 + std::sqrt(3.0) * i) / 2.0, u3 = (-1.0 - std::sqrt(3.0) * i) / 2.0;
    const cd t1 = (b + C + d0 / C) / (-3.0 * a);
    const cd t2 = (b + u2 * C + d0 / (u2 * C)) / (-3.0 * a);
    const cd t3 = (b + u3 * C + d0 / (u3 * C)) / (-3.0 * a);
    double t = -1;
    if (std::abs(t1.imag()) < tol && t1.real() > 0) t = t1.real();
    if (std::abs(t2.imag()) < tol && t2.real() > 0 && (t2.real() < t || t < 0)) t = t2.real();
    if (std::abs(t3.imag()) < tol && t3.real() > 0 && (t3.real() < t || t < 0)) t = t3.real();
    return t;
}

// det(Ds(x + t p)) = a t^3 + b t^2 + c t + det(Ds(x)) with Ds columns x1-x0, x2-x0, x3-x0;
// each tet bounds t by the smallest positive root of det(Ds(x + t p)) = 0.2 det(Ds(x))
// (the 1 - 0.2 factor of IPC.h:1871); min with alpha
double inversion_free_stepsize(const Scene& S, const double* x, const double* p, double alpha) {
    const int nTet = (int)S.tet.size() / 4;
    double out = alpha;
    for (int t = 0; t < nTet; ++t) {
        const int* v = &S.tet[4 * t];
        V3 x0 = X3(x, v[0]), p0 = X3(p, v[0]);
        V3 e[3], d[3];
        for (int k = 0; k < 3; ++k) { e[k] = X3(x, v[k + 1]) - x0; d[k] = X3(p, v[k + 1]) - p0; }
        // det[e0 + t d0, e1 + t d1, e2 + t d2] = (e0 + t d0) . ((e1 + t d1) x (e2 + t d2))
        const double a = dot(d[0], cross(d[1], d[2]));
        const double b = dot(e[0], cross(d[1], d[2])) + dot(d[0], cross(e[1], d[2])) + dot(d[0], cross(d[1], e[2]));
        const double c = dot(d[0], cross(e[1], e[2])) + dot(e[0], cross(d[1], e[2])) + dot(e[0], cross(e[1], d[2]));
        const double dd = (1.0 - 0.2) * dot(e[0], cross(e[1], e[2]));
        const double tt = cubic_root(a, b, c, dd, 1.e-6);
        out = std::min(out, tt >= 0 ? tt : 1e20);
    }
    return out;
}

}  // namespace orc
