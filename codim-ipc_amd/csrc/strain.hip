// Isotropic strain limiting of the shell on the device: a barrier on the singular
// values of each triangle's in-plane deformation gradient.
//
// Restates Compute_Inextensibility_Energy / _Gradient / _Hessian (FEM/Shell/INEXT.h:
// 120-467) with the default C2 barrier f_sl (INEXT.h:5-69, C1SL undefined): per
// singular value s, h^2 vol kappa_s[0] f_sl(s; s[0], sHat[0]) when s > sHat[0] and the
// potential is invalid once s >= s[0]; the lower limit (kappa_s[1] > 0) mirrors it
// below sHat[1] / s[1].  F = A B^-1 with A = [x1 - x0, x2 - x0] and B the
// upper-triangular factor of the rest first fundamental form (INEXT.h:145-151).
//
// One thread per triangle.  The thin SVD comes from the 2x2 eigenproblem of F^T F in
// closed form.  The Hessian in F is diagonal in six orthonormal directions (u0 v0^T,
// u1 v1^T, twist / flip (u0 v1^T -+ u1 v0^T)/sqrt2, u2 v0^T, u2 v1^T with u2 = u0 x u1),
// so the PSD projection (Math/UTILS.h makePD on the 6x6 dE/dF^2, INEXT.h:437-439)
// clamps those six eigenvalues and needs no iterative eigensolver.
// An invalid triangle contributes +inf to the energy: the line search reads an infinite
// potential as "invalid" (IMPLICIT_EULER.h:74-92, 131-146).
// The gradient and Hessian are added to the membrane's local slots of the same
// triangle, so they ride the membrane's gather and assembly plans.
#include "ctx.hpp"
#include "dev_math.hpp"

namespace cipc {

struct SLPar {
    double k0, s0, sh0, k1, s1, sh1;
    double h2;
};

// C2 barrier of INEXT.h:5-38: value, first and second derivative
CIPC_DEV void sl_f(double sg, double slim, double shat, double& f, double& g, double& hh) {
    const double a = shat - sg, b = slim - sg, D = (slim - shat) * (slim - shat);
    const double L = log(b / (slim - shat));
    f = -a * a * L / D;
    g = (2 * a * L + a * a / b) / D;
    const double c = a / b;
    hh = (-2 * L + c * (c - 4)) / D;
}
// phi, phi', phi'' of one singular value; false when a limit is reached
CIPC_DEV bool sl_phi(const SLPar& P, double sg, double& f, double& g, double& hh) {
    f = g = hh = 0;
    if (sg >= P.s0) return false;
    if (sg > P.sh0) {
        double a, b, c;
        sl_f(sg, P.s0, P.sh0, a, b, c);
        f += P.k0 * a; g += P.k0 * b; hh += P.k0 * c;
    }
    if (P.k1 > 0) {
        if (sg <= P.s1) return false;
        if (sg < P.sh1) {
            double a, b, c;
            sl_f(sg, P.s1, P.sh1, a, b, c);
            f += P.k1 * a; g += P.k1 * b; hh += P.k1 * c;
        }
    }
    return true;
}

struct SLSvd {
    d3 F0, F1;          // columns of F
    double sig[2];
    double v0x, v0y;    // v0 = (v0x, v0y), v1 = (-v0y, v0x)
    d3 u[3];
};
CIPC_DEV void sl_svd(d3 x0, d3 x1, d3 x2, double i00, double i01, double i11, SLSvd& S) {
    const d3 A0 = x1 - x0, A1 = x2 - x0;
    S.F0 = i00 * A0;
    S.F1 = i01 * A0 + i11 * A1;
    const double a = sq(S.F0), b = dot(S.F0, S.F1), d = sq(S.F1);
    const double hm = 0.5 * (a - d), r = sqrt(hm * hm + b * b);
    const double l0 = 0.5 * (a + d) + r, l1 = dmax(0.5 * (a + d) - r, 0.0);
    S.sig[0] = sqrt(l0);
    S.sig[1] = sqrt(l1);
    double vx, vy;
    if (r == 0) { vx = 1; vy = 0; }
    else if (hm >= 0) { vx = hm + r; vy = b; }
    else { vx = b; vy = r - hm; }
    const double n = sqrt(vx * vx + vy * vy);
    vx /= n; vy /= n;
    S.v0x = vx; S.v0y = vy;
    S.u[0] = (1.0 / S.sig[0]) * (vx * S.F0 + vy * S.F1);
    const d3 w = (-vy) * S.F0 + vx * S.F1;
    S.u[1] = S.sig[1] > 0 ? (1.0 / S.sig[1]) * w : mk(0, 0, 0);
    const d3 nr = cross(S.u[0], S.u[1]);
    S.u[2] = (1.0 / sqrt(sq(nr))) * nr;
}

__device__ __forceinline__ bool sl_all_dbc(const uint8_t* dbc, int a, int b, int c) { return dbc[a] && dbc[b] && dbc[c]; }

__global__ void k_sl_energy(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ IB,
                            const double* __restrict__ vol, SLPar P, const uint8_t* __restrict__ dbc, double* __restrict__ partial) {
    __shared__ double lds[4];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (t < nT) {
        const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
        if (!sl_all_dbc(dbc, a, b, c)) {
            SLSvd S;
            sl_svd(ld3(x, a), ld3(x, b), ld3(x, c), IB[t], IB[nT + t], IB[2 * nT + t], S);
            double f0, g0, h0, f1, g1, h1;
            const bool ok = sl_phi(P, S.sig[0], f0, g0, h0) & sl_phi(P, S.sig[1], f1, g1, h1);
            e = ok ? P.h2 * vol[t] * (f0 + f1) : __builtin_inf();
        }
    }
    const double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Compute_Inextensibility (INEXT.h:72-114) for the kappa_s rule of IMPLICIT_EULER.h:
// 606-645: near[t] = 1 when triangle t's largest singular value lies within thr of the
// upper limit s0 (a triangle whose three nodes are Dirichlet counts as 1, never near);
// *flag is set when a triangle is near now and was near at the previous iterate
__global__ void k_sl_near_limit(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ IB,
                                const uint8_t* __restrict__ dbc, double s0, double thr, const uint8_t* __restrict__ prev,
                                uint8_t* __restrict__ near, uint32_t* __restrict__ flag) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    double smax = 1.0;
    if (!sl_all_dbc(dbc, a, b, c)) {
        SLSvd S;
        sl_svd(ld3(x, a), ld3(x, b), ld3(x, c), IB[t], IB[nT + t], IB[2 * nT + t], S);
        smax = S.sig[0];
    }
    const uint8_t n = (s0 - smax < thr) ? 1 : 0;
    near[t] = n;
    if (n && prev && prev[t]) atomicOr(flag, 1u);
}

// vec(u_a v_b^T) with columns stacked (F column 0 then column 1)
CIPC_DEV void sl_dir(const SLSvd& S, int ua, int vb, double* o) {
    const d3 u = S.u[ua];
    const double c0 = vb == 0 ? S.v0x : -S.v0y, c1 = vb == 0 ? S.v0y : S.v0x;
    o[0] = u.x * c0; o[1] = u.y * c0; o[2] = u.z * c0;
    o[3] = u.x * c1; o[4] = u.y * c1; o[5] = u.z * c1;
}

// gradient added to the membrane's local slots of the triangle (slot 3 t + q)
__global__ void k_sl_grad_add(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ IB,
                              const double* __restrict__ vol, SLPar P, const uint8_t* __restrict__ dbc, double* __restrict__ gl) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    if (sl_all_dbc(dbc, a, b, c)) return;
    SLSvd S;
    const double i00 = IB[t], i01 = IB[nT + t], i11 = IB[2 * nT + t];
    sl_svd(ld3(x, a), ld3(x, b), ld3(x, c), i00, i01, i11, S);
    double f[2], g[2], hh[2];
    if (!(sl_phi(P, S.sig[0], f[0], g[0], hh[0]) & sl_phi(P, S.sig[1], f[1], g[1], hh[1]))) return;
    if (g[0] == 0 && g[1] == 0) return;
    const double w = P.h2 * vol[t];
    double G[6] = {0, 0, 0, 0, 0, 0}, d[6];
    for (int i = 0; i < 2; ++i) {
        sl_dir(S, i, i, d);
#pragma unroll
        for (int k = 0; k < 6; ++k) G[k] += w * g[i] * d[k];
    }
    const double c0[3] = {-i00, i00, 0.0}, c1[3] = {-(i01 + i11), i01, i11};
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) gl[3 * (3l * t + q) + k] += c0[q] * G[k] + c1[q] * G[3 + k];
}

// projected Hessian added to the membrane's local blocks of the triangle
// (block (p, q) at 81 t + 9 (3 p + q), entry (k, l) at 3 k + l)
__global__ void __launch_bounds__(64) k_sl_hess_add(const double* __restrict__ x, const int* __restrict__ tri, int nT,
                                                    const double* __restrict__ IB, const double* __restrict__ vol, SLPar P,
                                                    const uint8_t* __restrict__ dbc, double* __restrict__ hl) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    if (sl_all_dbc(dbc, a, b, c)) return;
    SLSvd S;
    const double i00 = IB[t], i01 = IB[nT + t], i11 = IB[2 * nT + t];
    sl_svd(ld3(x, a), ld3(x, b), ld3(x, c), i00, i01, i11, S);
    double f[2], g[2], hh[2];
    if (!(sl_phi(P, S.sig[0], f[0], g[0], hh[0]) & sl_phi(P, S.sig[1], f[1], g[1], hh[1]))) return;
    if (g[0] == 0 && g[1] == 0 && hh[0] == 0 && hh[1] == 0) return;
    const double w = P.h2 * vol[t];
    const double s0 = S.sig[0], s1 = S.sig[1];
    const double lt = (g[0] + g[1]) / (s0 + s1);
    const double lf = fabs(s0 - s1) > 1e-12 * s0 ? (g[0] - g[1]) / (s0 - s1) : 0.5 * (hh[0] + hh[1]);
    // six eigenpairs of d^2E/dF^2, eigenvalues clamped at zero
    double d00[6], d11[6], d01[6], d10[6], d20[6], d21[6];
    sl_dir(S, 0, 0, d00); sl_dir(S, 1, 1, d11); sl_dir(S, 0, 1, d01); sl_dir(S, 1, 0, d10); sl_dir(S, 2, 0, d20); sl_dir(S, 2, 1, d21);
    const double r2 = 0.70710678118654752440;
    double tw[6], fl[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) { tw[k] = r2 * (d01[k] - d10[k]); fl[k] = r2 * (d01[k] + d10[k]); }
    const double lam[6] = {dmax(hh[0], 0.0), dmax(hh[1], 0.0), dmax(lt, 0.0), dmax(lf, 0.0), dmax(g[0] / s0, 0.0), dmax(g[1] / s1, 0.0)};
    const double* dirs[6] = {d00, d11, tw, fl, d20, d21};
    double HF[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) HF[k] = 0;
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const double le = w * lam[e];
        if (le == 0) continue;
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int q = 0; q < 6; ++q) HF[r * 6 + q] += le * dirs[e][r] * dirs[e][q];
    }
    const double c0[3] = {-i00, i00, 0.0}, c1[3] = {-(i01 + i11), i01, i11};
    double* dst = hl + 81l * t;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l)
                    dst[9 * (3 * p + q) + 3 * k + l] += c0[p] * c0[q] * HF[k * 6 + l] + c0[p] * c1[q] * HF[k * 6 + 3 + l] +
                                                        c1[p] * c0[q] * HF[(3 + k) * 6 + l] + c1[p] * c1[q] * HF[(3 + k) * 6 + 3 + l];
}

// ------------------------------------------------------------------ anisotropic (fibre) model
// Check_Fiber_Feasibility / Compute_Fiber_Energy / _Gradient / _Hessian (FEM/Shell/
// ANISO_INEXT.h:33-440): Green strain E = (F^T F - I)/2 of F = A B^-1, B from the rest
// x and z coordinates; psi = h^2 vol [k0/2 eta(E00^2, L0^2) + k1 eta(E00 E11, L0 L1) +
// k2/2 eta(E11^2, L1^2) + k3 eta(E01^2, L2^2)], eta(s, m) = -m ln((m - s)/m); +inf when
// |E00| >= L0, |E11| >= L1 or |E01| >= L2.  It replaces the membrane (same local slots).
// The 6x6 F-space Hessian is projected (ANISO_INEXT.h:322-324) by embedding it in the
// translation-free subspace of a 9x9 matrix and running the in-register projection
// make_pd_translation<3> there: for orthonormal W (3x2, columns orthogonal to (1,1,1)),
// (W x I) H_F (W x I)^T has reduced matrix H_F, so the projected 9x9 is
// (W x I) makePD(H_F) (W x I)^T.
struct FBPar {
    double k[4], L[3];
    double h2;
};
CIPC_DEV double fb_eta0(double s, double m) { return -m * log((m - s) / m); }
CIPC_DEV double fb_eta1(double s, double m) { return m / (m - s); }
CIPC_DEV double fb_eta2(double s, double m) { return m / ((m - s) * (m - s)); }

struct FBTri {
    d3 F0, F1;
    double E00, E11, E01, c0[3], c1[3];
};
CIPC_DEV bool fb_eval(const double* __restrict__ x, int a, int b, int c, const double* __restrict__ IB, int nT, int t, const FBPar& P,
                      FBTri& T) {
    const double i00 = IB[t], i01 = IB[nT + t], i10 = IB[2 * nT + t], i11 = IB[3 * nT + t];
    const d3 A0 = ld3(x, b) - ld3(x, a), A1 = ld3(x, c) - ld3(x, a);
    T.F0 = i00 * A0 + i10 * A1;
    T.F1 = i01 * A0 + i11 * A1;
    T.E00 = 0.5 * (sq(T.F0) - 1);
    T.E11 = 0.5 * (sq(T.F1) - 1);
    T.E01 = 0.5 * dot(T.F0, T.F1);
    T.c0[0] = -(i00 + i10); T.c0[1] = i00; T.c0[2] = i10;
    T.c1[0] = -(i01 + i11); T.c1[1] = i01; T.c1[2] = i11;
    return fabs(T.E00) < P.L[0] && fabs(T.E11) < P.L[1] && fabs(T.E01) < P.L[2];
}

__global__ void k_fb_energy(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ IB,
                            const double* __restrict__ vol, FBPar P, const uint8_t* __restrict__ dbc, double* __restrict__ partial) {
    __shared__ double lds[4];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (t < nT) {
        const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
        if (!sl_all_dbc(dbc, a, b, c)) {
            FBTri T;
            if (fb_eval(x, a, b, c, IB, nT, t, P, T)) {
                const double m00 = P.L[0] * P.L[0], m11 = P.L[1] * P.L[1], m01 = P.L[0] * P.L[1], m2 = P.L[2] * P.L[2];
                e = P.h2 * vol[t] * (P.k[0] / 2 * fb_eta0(T.E00 * T.E00, m00) + P.k[1] * fb_eta0(T.E00 * T.E11, m01) +
                                     P.k[2] / 2 * fb_eta0(T.E11 * T.E11, m11) + P.k[3] * fb_eta0(T.E01 * T.E01, m2));
            } else {
                e = __builtin_inf();
            }
        }
    }
    const double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// dpsi/dE (S00, S11, and S01 for the single variable E01)
CIPC_DEV void fb_stress(const FBTri& T, const FBPar& P, double w, double& S00, double& S11, double& S01) {
    const double m00 = P.L[0] * P.L[0], m11 = P.L[1] * P.L[1], m01 = P.L[0] * P.L[1], m2 = P.L[2] * P.L[2];
    const double ex = fb_eta1(T.E00 * T.E11, m01);
    S00 = w * (P.k[0] * fb_eta1(T.E00 * T.E00, m00) * T.E00 + P.k[1] * ex * T.E11);
    S11 = w * (P.k[2] * fb_eta1(T.E11 * T.E11, m11) * T.E11 + P.k[1] * ex * T.E00);
    S01 = w * 2 * P.k[3] * fb_eta1(T.E01 * T.E01, m2) * T.E01;
}

// local gradient into the membrane's triangle slots (overwrites; zero when inactive)
__global__ void k_fb_grad_local(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ IB,
                                const double* __restrict__ vol, FBPar P, const uint8_t* __restrict__ dbc, double* __restrict__ gl) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    FBTri T;
    if (!sl_all_dbc(dbc, a, b, c) && fb_eval(x, a, b, c, IB, nT, t, P, T)) {
        double S00, S11, S01;
        fb_stress(T, P, P.h2 * vol[t], S00, S11, S01);
        const d3 G0 = S00 * T.F0 + (0.5 * S01) * T.F1, G1 = S11 * T.F1 + (0.5 * S01) * T.F0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            g[3 * q] = T.c0[q] * G0.x + T.c1[q] * G1.x;
            g[3 * q + 1] = T.c0[q] * G0.y + T.c1[q] * G1.y;
            g[3 * q + 2] = T.c0[q] * G0.z + T.c1[q] * G1.z;
        }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) gl[9l * t + q] = g[q];
}

// projected local Hessian into the membrane's triangle blocks (overwrites)
__global__ void __launch_bounds__(64) k_fb_hess_local(const double* __restrict__ x, const int* __restrict__ tri, int nT,
                                                      const double* __restrict__ IB, const double* __restrict__ vol, FBPar P,
                                                      const uint8_t* __restrict__ dbc, double* __restrict__ hl) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    double* dst = hl + 81l * t;
    FBTri T;
    if (sl_all_dbc(dbc, a, b, c) || !fb_eval(x, a, b, c, IB, nT, t, P, T)) {
#pragma unroll
        for (int k = 0; k < 81; ++k) dst[k] = 0;
        return;
    }
    const double w = P.h2 * vol[t];
    double S00, S11, S01;
    fb_stress(T, P, w, S00, S11, S01);
    const double m00 = P.L[0] * P.L[0], m11 = P.L[1] * P.L[1], m01 = P.L[0] * P.L[1], m2 = P.L[2] * P.L[2];
    const double s00 = T.E00 * T.E00, s11 = T.E11 * T.E11, sx = T.E00 * T.E11, s2 = T.E01 * T.E01;
    const double H00 = w * (P.k[0] * (2 * fb_eta2(s00, m00) * s00 + fb_eta1(s00, m00)) + P.k[1] * fb_eta2(sx, m01) * s11);
    const double H11 = w * (P.k[2] * (2 * fb_eta2(s11, m11) * s11 + fb_eta1(s11, m11)) + P.k[1] * fb_eta2(sx, m01) * s00);
    const double H0x = w * P.k[1] * (fb_eta2(sx, m01) * sx + fb_eta1(sx, m01));
    const double H22 = w * 2 * P.k[3] * (2 * fb_eta2(s2, m2) * s2 + fb_eta1(s2, m2));
    const double dA[6] = {T.F0.x, T.F0.y, T.F0.z, 0, 0, 0}, dB[6] = {0, 0, 0, T.F1.x, T.F1.y, T.F1.z};
    const double dC[6] = {0.5 * T.F1.x, 0.5 * T.F1.y, 0.5 * T.F1.z, 0.5 * T.F0.x, 0.5 * T.F0.y, 0.5 * T.F0.z};
    double HF[36];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q)
            HF[r * 6 + q] = H00 * dA[r] * dA[q] + H11 * dB[r] * dB[q] + H0x * (dA[r] * dB[q] + dB[r] * dA[q]) + H22 * dC[r] * dC[q];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        HF[i * 6 + i] += S00;
        HF[(3 + i) * 6 + 3 + i] += S11;
        HF[i * 6 + 3 + i] += 0.5 * S01;
        HF[(3 + i) * 6 + i] += 0.5 * S01;
    }
    // embed: H9 = (W x I) H_F (W x I)^T, then project in the translation-free subspace
    double H9[81];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double v = 0;
#pragma unroll
                    for (int p = 0; p < 2; ++p)
#pragma unroll
                        for (int q = 0; q < 2; ++q) v += tr_basis<3>(k, p) * tr_basis<3>(l, q) * HF[(3 * p + i) * 6 + 3 * q + j];
                    H9[(3 * k + i) * 9 + 3 * l + j] = v;
                }
    double P9[81];
    make_pd_translation<3>(H9, 9, [&](int r, int cc, double v) { P9[r * 9 + cc] = v; });
    // H_x = (Wh x I)^T P9 (Wh x I), Wh(k, a) = sum_p W(k, p) c_p[a]
    double Wh[9];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int q = 0; q < 3; ++q) Wh[k * 3 + q] = tr_basis<3>(k, 0) * T.c0[q] + tr_basis<3>(k, 1) * T.c1[q];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double v = 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k)
#pragma unroll
                        for (int l = 0; l < 3; ++l) v += Wh[k * 3 + p] * Wh[l * 3 + q] * P9[(3 * k + i) * 9 + 3 * l + j];
                    dst[9 * (3 * p + q) + 3 * i + j] = v;
                }
}

// ------------------------------------------------------------------ host side
static SLPar sl_par(const Ctx& c) {
    SLPar P;
    P.k0 = c.P.slKappa[0]; P.s0 = c.P.slS[0]; P.sh0 = c.P.slSHat[0];
    P.k1 = c.P.slKappa[1]; P.s1 = c.P.slS[1]; P.sh1 = c.P.slSHat[1];
    P.h2 = shell_h2(c);
    return P;
}
bool strain_limit_on(const Ctx& c) { return c.P.slKappa[0] > 0 && c.nT > 0; }

double strain_limit_energy(Ctx& c, const double* x) {
    if (!strain_limit_on(c)) return 0.0;
    const int nb = nblocks(c.nT);
    double* part = c.red.ensure(nb);
    hipLaunchKernelGGL(k_sl_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.tri.p, c.nT, c.slIB.p, c.vol.p, sl_par(c), c.dbc.p, part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 49, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 49, c.scal.p + 49, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[49];
}
bool strain_limit_near_twice(Ctx& c, const double* x, bool first) {
    if (!strain_limit_on(c)) return false;
    uint8_t* cur = (c.slNearCur ? c.slNear1 : c.slNear0).ensure(c.nT);
    const uint8_t* prev = first ? nullptr : (c.slNearCur ? c.slNear0 : c.slNear1).p;
    const double thr = 1.0e-4 * (c.P.slS[0] - c.P.slSHat[0]);
    CIPC_HIP(hipMemsetAsync(c.counters.p + 15, 0, sizeof(uint32_t), c.s));
    hipLaunchKernelGGL(k_sl_near_limit, dim3(nblocks(c.nT)), dim3(kBlock), 0, c.s, x, c.tri.p, c.nT, c.slIB.p, c.dbc.p, c.P.slS[0], thr,
                       prev, cur, c.counters.p + 15);
    CIPC_LAUNCH_CHECK();
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 15, c.counters.p + 15, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    c.slNearCur ^= 1;
    return c.hCnt[15] != 0;
}
void strain_limit_grad_add(Ctx& c, const double* x, double* gl) {
    if (!strain_limit_on(c)) return;
    hipLaunchKernelGGL(k_sl_grad_add, dim3(nblocks(c.nT)), dim3(kBlock), 0, c.s, x, c.tri.p, c.nT, c.slIB.p, c.vol.p, sl_par(c), c.dbc.p, gl);
    CIPC_LAUNCH_CHECK();
}
void strain_limit_hess_add(Ctx& c, const double* x, double* hl) {
    if (!strain_limit_on(c)) return;
    hipLaunchKernelGGL(k_sl_hess_add, dim3(nblocks(c.nT, 64)), dim3(64), 0, c.s, x, c.tri.p, c.nT, c.slIB.p, c.vol.p, sl_par(c), c.dbc.p, hl);
    CIPC_LAUNCH_CHECK();
}

static FBPar fb_par(const Ctx& c) {
    FBPar P;
    for (int i = 0; i < 4; ++i) P.k[i] = c.P.fiberK[i];
    for (int i = 0; i < 3; ++i) P.L[i] = c.P.fiberLimit[i];
    P.h2 = shell_h2(c);
    return P;
}
bool fiber_on(const Ctx& c) { return (c.P.fiberK[0] > 0 || c.P.fiberK[1] > 0) && c.nT > 0; }
double fiber_energy(Ctx& c, const double* x) {
    const int nb = nblocks(c.nT);
    double* part = c.red.ensure(nb);
    hipLaunchKernelGGL(k_fb_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.tri.p, c.nT, c.fbIB.p, c.vol.p, fb_par(c), c.dbc.p, part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 51, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 51, c.scal.p + 51, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[51];
}
void fiber_grad_local(Ctx& c, const double* x, double* gl) {
    hipLaunchKernelGGL(k_fb_grad_local, dim3(nblocks(c.nT)), dim3(kBlock), 0, c.s, x, c.tri.p, c.nT, c.fbIB.p, c.vol.p, fb_par(c), c.dbc.p, gl);
    CIPC_LAUNCH_CHECK();
}
void fiber_hess_local(Ctx& c, const double* x, double* hl) {
    hipLaunchKernelGGL(k_fb_hess_local, dim3(nblocks(c.nT, 64)), dim3(64), 0, c.s, x, c.tri.p, c.nT, c.fbIB.p, c.vol.p, fb_par(c), c.dbc.p, hl);
    CIPC_LAUNCH_CHECK();
}

}  // namespace cipc
