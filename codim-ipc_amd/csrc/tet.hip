// Volumetric Neo-Hookean tetrahedra inside the shell step, on the device.
//
// Restates the tet terms of Compute_IncPotential / _Gradient / _Hessian
// (FEM/Shell/INC_POTENTIAL.h:98-109, 226-229, 354-364): per tet, with F = Ds Dm^-1
// (Ds = [x1 - x0, x2 - x0, x3 - x0], Dm^-1 from the rest shape, FEM/DEFORMATION_GRADIENT.h:
// 39-77, 113-141), psi = mu/2 (|F|^2 - 3) - mu ln J + lambda/2 ln^2 J
// (Physics/NEOHOOKEAN.h:136-142) weighted by h^2 vol; the stress P = mu (F - F^-T) +
// lambda ln J F^-T (NEOHOOKEAN.h:100-109); the stress derivative built in the singular
// basis of F and projected there -- the 3x3 d2psi/dsigma2 by its eigen-decomposition and
// each of the three twist/flip 2x2 blocks by clamping its two eigenvalues
// (NEOHOOKEAN.h:116-300) -- which, psi being isotropic, is the exact PSD projection of
// dP/dF.  An inverted tet (det F <= 0) makes the potential invalid (INC_POTENTIAL.h:
// 98-107): the energy is +inf.  Compute_Inversion_Free_StepSize (FEM/IPC.h:1785-1877)
// is provided as a component (the shell step's line search uses the validity halving).
//
// One thread per tet; the 3x3 eigenproblems (F^T F for the SVD, d2psi/dsigma2) are cyclic
// Jacobi in registers.  Local gradients (4 slots) and Hessian blocks (16 slots) go to the
// elastic gather plans after the rods' slots.
#include "ctx.hpp"
#include "dev_math.hpp"
#include <algorithm>

namespace cipc {

struct TetPar {
    double lam, mu, h2;
};

CIPC_DEV double det3(const double* M) {
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// F (row-major) of tet t from SoA indices [4][nTet] and Dm^-1 [9][nTet]
CIPC_DEV void tet_F(const double* __restrict__ x, const int* __restrict__ tv, const double* __restrict__ IB, int nTet, int t,
                    double* F) {
    const d3 x0 = ld3(x, tv[t]);
    const d3 e0 = ld3(x, tv[nTet + t]) - x0, e1 = ld3(x, tv[2 * nTet + t]) - x0, e2 = ld3(x, tv[3 * nTet + t]) - x0;
    const double Ds[9] = {e0.x, e1.x, e2.x, e0.y, e1.y, e2.y, e0.z, e1.z, e2.z};
    double B[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) B[k] = IB[(long)k * nTet + t];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[3 * i + j] = Ds[3 * i] * B[j] + Ds[3 * i + 1] * B[3 + j] + Ds[3 * i + 2] * B[6 + j];
}

CIPC_DEV void inv3(const double* M, double* I) {
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

// symmetric 3x3 eigen-decomposition by cyclic Jacobi: A (row-major, overwritten), V
// columns = eigenvectors, w unsorted eigenvalues
CIPC_DEV void sym_eig3(double* A, double* w, double* V) {
#pragma unroll
    for (int k = 0; k < 9; ++k) V[k] = (k % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 12; ++sweep) {
        const double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
        const double dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
        if (off <= 1e-30 * dia) break;
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            const double apq = A[3 * p + q];
            if (apq == 0) continue;
            const double th = (A[3 * q + q] - A[3 * p + p]) / (2 * apq);
            const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
            const double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {   // A <- A J (columns p, q)
                const double akp = A[3 * k + p], akq = A[3 * k + q];
                A[3 * k + p] = c * akp - s * akq;
                A[3 * k + q] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {   // A <- J^T A (rows p, q)
                const double apk = A[3 * p + k], aqk = A[3 * q + k];
                A[3 * p + k] = c * apk - s * aqk;
                A[3 * q + k] = s * apk + c * aqk;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double vkp = V[3 * k + p], vkq = V[3 * k + q];
                V[3 * k + p] = c * vkp - s * vkq;
                V[3 * k + q] = s * vkp + c * vkq;
            }
        }
    }
    w[0] = A[0]; w[1] = A[4]; w[2] = A[8];
}

// F = U diag(s) V^T with V a rotation, s[0] >= s[1] >= |s[2]| (s[2] < 0 for an inverted F)
CIPC_DEV void svd3(const double* F, double* U, double* s, double* V) {
    double C[9], w[3], Q[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    sym_eig3(C, w, Q);
    int o[3] = {0, 1, 2};   // descending order
    if (w[o[0]] < w[o[1]]) { const int t = o[0]; o[0] = o[1]; o[1] = t; }
    if (w[o[1]] < w[o[2]]) { const int t = o[1]; o[1] = o[2]; o[2] = t; }
    if (w[o[0]] < w[o[1]]) { const int t = o[0]; o[0] = o[1]; o[1] = t; }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) V[3 * r + c] = Q[3 * r + o[c]];
    if (det3(V) < 0)
#pragma unroll
        for (int r = 0; r < 3; ++r) V[3 * r + 2] = -V[3 * r + 2];
    d3 u[3];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const d3 fv = mk(F[0] * V[c] + F[1] * V[3 + c] + F[2] * V[6 + c], F[3] * V[c] + F[4] * V[3 + c] + F[5] * V[6 + c],
                         F[6] * V[c] + F[7] * V[3 + c] + F[8] * V[6 + c]);
        s[c] = sqrt(sq(fv));
        u[c] = (1.0 / s[c]) * fv;
    }
    u[2] = cross(u[0], u[1]);
    const d3 fv2 = mk(F[0] * V[2] + F[1] * V[5] + F[2] * V[8], F[3] * V[2] + F[4] * V[5] + F[5] * V[8], F[6] * V[2] + F[7] * V[5] + F[8] * V[8]);
    s[2] = dot(u[2], fv2);
#pragma unroll
    for (int c = 0; c < 3; ++c) { U[c] = u[c].x; U[3 + c] = u[c].y; U[6 + c] = u[c].z; }
}

CIPC_DEV double nh_psi(const double* F, double lam, double mu) {
    double f2 = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) f2 += F[k] * F[k];
    const double lj = log(det3(F));
    return mu / 2 * (f2 - 3) - (mu - lam / 2 * lj) * lj;
}

__global__ void k_tet_energy(const double* __restrict__ x, const int* __restrict__ tv, int nTet, const double* __restrict__ IB,
                             const double* __restrict__ vol, TetPar P, double* __restrict__ partial) {
    __shared__ double lds[4];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (t < nTet) {
        double F[9];
        tet_F(x, tv, IB, nTet, t, F);
        e = det3(F) > 0 ? P.h2 * vol[t] * nh_psi(F, P.lam, P.mu) : __builtin_inf();
    }
    const double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_tet_invalid(const double* __restrict__ x, const int* __restrict__ tv, int nTet, const double* __restrict__ IB,
                              uint32_t* __restrict__ flag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nTet) return;
    double F[9];
    tet_F(x, tv, IB, nTet, t, F);
    if (!(det3(F) > 0)) atomicOr(flag, 1u);
}

// gradient slots off + 4 t + a: g_a = h^2 vol P Dm^-T e_a (a = 1..3), g_0 = -sum
__global__ void k_tet_grad_local(const double* __restrict__ x, const int* __restrict__ tv, int nTet, const double* __restrict__ IB,
                                 const double* __restrict__ vol, TetPar P, double* __restrict__ gl, long off) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nTet) return;
    double F[9], Fi[9];
    tet_F(x, tv, IB, nTet, t, F);
    inv3(F, Fi);
    const double lj = log(det3(F)), w = P.h2 * vol[t];
    double Pk[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double fit = Fi[3 * j + i];
            Pk[3 * i + j] = w * (P.mu * (F[3 * i + j] - fit) + P.lam * lj * fit);
        }
    double g0[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double s = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) s += Pk[3 * i + j] * IB[(long)(3 * a + j) * nTet + t];
            gl[3 * (off + 4l * t + a + 1) + i] = s;
            g0[i] -= s;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) gl[3 * (off + 4l * t) + i] = g0[i];
}

// projected stress derivative (9x9, vec(F) column-major), weight w
CIPC_DEV void nh_dPdF_proj(const double* F, double lam, double mu, double w, double* D) {
    double U[9], s[3], V[9];
    svd3(F, U, s, V);
    const double ls = log(s[0] * s[1] * s[2]);
    double dE[3], A[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double inv = 1.0 / s[i];
        dE[i] = mu * (s[i] - inv) + lam * inv * ls;
        A[4 * i] = mu * (1.0 + inv * inv) - lam * inv * inv * (ls - 1.0);
    }
    A[1] = A[3] = lam / s[0] / s[1];
    A[5] = A[7] = lam / s[1] / s[2];
    A[2] = A[6] = lam / s[2] / s[0];
    // A <- A+ (eigenvalues clamped at zero)
    {
        double Aw[9], ew[3], EV[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Aw[k] = A[k];
        sym_eig3(Aw, ew, EV);
        if (ew[0] < 0 || ew[1] < 0 || ew[2] < 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double v = 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) v += EV[3 * i + k] * dmax(ew[k], 0.0) * EV[3 * j + k];
                    A[3 * i + j] = v;
                }
        }
    }
    const double mid = mu - lam * ls;
    double Bd[3], Bo[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int d = (c + 1) % 3;
        const double l = (mu + mid / s[c] / s[d]) / 2.0;
        double r = dE[c] + dE[d];
        const double sum = s[c] + s[d];
        r /= (sum < 1.0e-6 ? 2 * 1.0e-6 : 2 * sum);
        // eigenvalues 2l on (1, 1), 2r on (1, -1), each clamped at zero
        const double lp = dmax(2 * l, 0.0) / 2, rp = dmax(2 * r, 0.0) / 2;
        Bd[c] = lp + rp;
        Bo[c] = lp - rp;
    }
    // D(p, q) = w sum M(k, m) Q(p, k) Q(q, m), Q(i + 3 j, a + 3 b) = U_ia V_jb, M nonzero on
    // the (aa, cc) block (A) and the twist/flip pairs
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int i = p % 3, j = p / 3;
#pragma unroll
        for (int q = p; q < 9; ++q) {
            const int r = q % 3, sI = q / 3;
            double acc = 0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) acc += A[3 * a + c] * U[3 * i + a] * V[3 * j + a] * U[3 * r + c] * V[3 * sI + c];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int d = (c + 1) % 3;
                const double qcd_p = U[3 * i + c] * V[3 * j + d], qdc_p = U[3 * i + d] * V[3 * j + c];
                const double qcd_q = U[3 * r + c] * V[3 * sI + d], qdc_q = U[3 * r + d] * V[3 * sI + c];
                acc += Bd[c] * (qcd_p * qcd_q + qdc_p * qdc_q) + Bo[c] * (qcd_p * qdc_q + qdc_p * qcd_q);
            }
            D[p * 9 + q] = w * acc;
            D[q * 9 + p] = w * acc;
        }
    }
}

// Hessian blocks at slots off + 16 t + 4 a + b (9 doubles each): H = D^T dPdF D with
// dF_ij/dx_a,k = delta_ik c_a(j), c_a = row a-1 of Dm^-1, c_0 = -sum
__global__ void __launch_bounds__(64) k_tet_hess_local(const double* __restrict__ x, const int* __restrict__ tv, int nTet,
                                                       const double* __restrict__ IB, const double* __restrict__ vol, TetPar P,
                                                       double* __restrict__ hl, long off) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nTet) return;
    double F[9], D[81];
    tet_F(x, tv, IB, nTet, t, F);
    nh_dPdF_proj(F, P.lam, P.mu, P.h2 * vol[t], D);
    double cc[4][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        cc[1][j] = IB[(long)j * nTet + t];
        cc[2][j] = IB[(long)(3 + j) * nTet + t];
        cc[3][j] = IB[(long)(6 + j) * nTet + t];
        cc[0][j] = -(cc[1][j] + cc[2][j] + cc[3][j]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            double* dst = hl + 9 * (off + 16l * t + 4 * a + b);
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    double acc = 0;
#pragma unroll
                    for (int j = 0; j < 3; ++j)
#pragma unroll
                        for (int s = 0; s < 3; ++s) acc += cc[a][j] * cc[b][s] * D[(k + 3 * j) * 9 + (l + 3 * s)];
                    dst[3 * k + l] = acc;
                }
        }
}

// ------------------------------------------------------------------ inversion-free step size
struct cplx {
    double re, im;
};
CIPC_DEV cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
CIPC_DEV cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
CIPC_DEV cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
CIPC_DEV cplx cdiv(cplx a, cplx b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
CIPC_DEV cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
// principal square root and principal cube root (exp(log(z) / 3))
CIPC_DEV cplx csqrt_p(cplx z) {
    const double r = hypot(z.re, z.im);
    if (r == 0) return {0, 0};
    const double re = sqrt(0.5 * (r + fabs(z.re))), im = 0.5 * z.im / re;
    return z.re >= 0 ? cplx{re, im} : cplx{fabs(im), z.im >= 0 ? re : -re};
}
CIPC_DEV cplx ccbrt_p(cplx z) {
    const double r = hypot(z.re, z.im);
    if (r == 0) return {0, 0};
    const double m = cbrt(r), th = atan2(z.im, z.re) / 3;
    return {m * cos(th), m * sin(th)};
}
CIPC_DEV double quad_root_dev(double a, double b, double c, double tol) {
    double t;
    if (fabs(a) <= tol) t = -c / b;
    else {
        const double desc = b * b - 4 * a * c;
        if (desc > 0) {
            t = (-b - sqrt(desc)) / (2 * a);
            if (t < 0) t = (-b + sqrt(desc)) / (2 * a);
        } else t = -1;
    }
    return t;
}
// smallest positive real root of a t^3 + b t^2 + c t + d, negative if none (FEM/IPC.h:
// 1751-1782: Cardano's form with the three complex cube roots, imaginary parts < tol)
CIPC_DEV double cubic_root_dev(double a, double b, double c, double d, double tol) {
    if (fabs(a) <= tol) return quad_root_dev(b, c, d, tol);
    const cplx d0 = {b * b - 3 * a * c, 0}, d1 = {2 * b * b * b - 9 * a * b * c + 27 * a * a * d, 0};
    const cplx disc = csqrt_p(csub(cmul(d1, d1), cscale(cmul(d0, cmul(d0, d0)), 4.0)));
    cplx C = ccbrt_p(cscale(cadd(d1, disc), 0.5));
    if (C.re == 0 && C.im == 0) C = ccbrt_p(cscale(csub(d1, disc), 0.5));
    const cplx u2 = {-0.5, 0.5 * sqrt(3.0)}, u3 = {-0.5, -0.5 * sqrt(3.0)};
    const cplx bb = {b, 0};
    const double m3a = -3.0 * a;
    const cplx t1 = cscale(cadd(cadd(bb, C), cdiv(d0, C)), 1.0 / m3a);
    const cplx u2C = cmul(u2, C), u3C = cmul(u3, C);
    const cplx t2 = cscale(cadd(cadd(bb, u2C), cdiv(d0, u2C)), 1.0 / m3a);
    const cplx t3 = cscale(cadd(cadd(bb, u3C), cdiv(d0, u3C)), 1.0 / m3a);
    double t = -1;
    if (fabs(t1.im) < tol && t1.re > 0) t = t1.re;
    if (fabs(t2.im) < tol && t2.re > 0 && (t2.re < t || t < 0)) t = t2.re;
    if (fabs(t3.im) < tol && t3.re > 0 && (t3.re < t || t < 0)) t = t3.re;
    return t;
}

// per tet: the smallest positive t with det(Ds(x + t p)) = 0.2 det(Ds(x)) (no root: no
// bound); global minimum by atomicMin on the bit pattern of positive doubles
__global__ void k_tet_inversion_free(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ tv, int nTet,
                                     unsigned long long* __restrict__ best) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nTet) return;
    const int v0 = tv[t];
    const d3 x0 = ld3(x, v0), p0 = ld3(p, v0);
    d3 e[3], d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int vk = tv[(k + 1) * nTet + t];
        e[k] = ld3(x, vk) - x0;
        d[k] = ld3(p, vk) - p0;
    }
    const double a = dot(d[0], cross(d[1], d[2]));
    const double b = dot(e[0], cross(d[1], d[2])) + dot(d[0], cross(e[1], d[2])) + dot(d[0], cross(d[1], e[2]));
    const double c = dot(d[0], cross(e[1], e[2])) + dot(e[0], cross(d[1], e[2])) + dot(e[0], cross(e[1], d[2]));
    const double dd = (1.0 - 0.2) * dot(e[0], cross(e[1], e[2]));
    const double r = cubic_root_dev(a, b, c, dd, 1.e-6);
    if (r >= 0) atomicMin(best, (unsigned long long)__double_as_longlong(r));
}

// ------------------------------------------------------------------ host side
static TetPar tet_par(const Ctx& c) { return {c.tetLam, c.tetMu, c.P.dt * c.P.dt}; }

double tet_energy(Ctx& c, const double* x) {
    if (!c.nTet) return 0.0;
    const int nb = nblocks(c.nTet);
    double* part = c.red.ensure(nb);
    hipLaunchKernelGGL(k_tet_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.tetIdx.p, c.nTet, c.tetIB.p, c.tetVol.p, tet_par(c), part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 53, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 53, c.scal.p + 53, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[53];
}
bool tets_valid(Ctx& c, const double* x) {
    if (!c.nTet) return true;
    CIPC_HIP(hipMemsetAsync(c.counters.p + 22, 0, sizeof(uint32_t), c.s));
    hipLaunchKernelGGL(k_tet_invalid, dim3(nblocks(c.nTet)), dim3(kBlock), 0, c.s, x, c.tetIdx.p, c.nTet, c.tetIB.p, c.counters.p + 22);
    CIPC_LAUNCH_CHECK();
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 22, c.counters.p + 22, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hCnt[22] == 0;
}
void tet_grad_local(Ctx& c, const double* x, double* gl, long off) {
    if (!c.nTet) return;
    hipLaunchKernelGGL(k_tet_grad_local, dim3(nblocks(c.nTet)), dim3(kBlock), 0, c.s, x, c.tetIdx.p, c.nTet, c.tetIB.p, c.tetVol.p,
                       tet_par(c), gl, off);
    CIPC_LAUNCH_CHECK();
}
void tet_hess_local(Ctx& c, const double* x, double* hl, long off) {
    if (!c.nTet) return;
    hipLaunchKernelGGL(k_tet_hess_local, dim3(nblocks(c.nTet, 64)), dim3(64), 0, c.s, x, c.tetIdx.p, c.nTet, c.tetIB.p, c.tetVol.p,
                       tet_par(c), hl, off);
    CIPC_LAUNCH_CHECK();
}
double inversion_free_stepsize(Ctx& c, const double* x, const double* p, double alpha) {
    if (!c.nTet) return alpha;
    c.hScal[55] = 1e300;   // no bound yet (pinned staging slot)
    uint64_t* slot = reinterpret_cast<uint64_t*>(c.scal.p + 54);
    CIPC_HIP(hipMemcpyAsync(slot, c.hScal + 55, sizeof(double), hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(k_tet_inversion_free, dim3(nblocks(c.nTet)), dim3(kBlock), 0, c.s, x, p, c.tetIdx.p, c.nTet,
                       reinterpret_cast<unsigned long long*>(slot));
    CIPC_LAUNCH_CHECK();
    CIPC_HIP(hipMemcpyAsync(c.hScal + 54, c.scal.p + 54, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return std::min(alpha, c.hScal[54]);
}

}  // namespace cipc
