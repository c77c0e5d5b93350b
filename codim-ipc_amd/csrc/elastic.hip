// Shell elasticity and inertia: Neo-Hookean membrane (FEM/Shell/MEMBRANE.h:7-290),
// hinge bending (FEM/Shell/BENDING.h:52-75, 177-208, 441-496), rod springs and rod
// hinges (FEM/Shell/Rod/ROD_STRETCHING.h:10-107, ROD_BENDING.h:10-117), inertia
// (FEM/Shell/INC_POTENTIAL.h:132-144, 250-269).  All carry the h^2 of the incremental
// potential.  Triangles / hinges whose nodes are all Dirichlet contribute nothing.
//
// Assembly is deterministic and atomic-free: every stencil writes its local gradient
// / Hessian blocks to its own slots, and a precomputed plan (built once per scene,
// the topology is static) lists for each vertex / each BSR block the slots to sum in
// a fixed order.  The BSR pattern holds every (row, col) pair of a triangle or hinge
// plus all diagonal blocks (the lumped mass lands there); Dirichlet rows and
// columns become identity / zero (Project_DBC, Math/CSR_MATRIX.h:130-141).
#include "ctx.hpp"
#include "dev_math.hpp"
#include <algorithm>
#include <cmath>

namespace cipc {

__device__ __forceinline__ bool all_dbc3(const uint8_t* dbc, int a, int b, int c) { return dbc[a] && dbc[b] && dbc[c]; }

__global__ void k_memb_energy(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ Binv,
                              const double* __restrict__ vol, double lam, double mu, double h2, const uint8_t* __restrict__ dbc,
                              double* __restrict__ partial) {
    __shared__ double lds[4];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (t < nT) {
        int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
        if (!all_dbc3(dbc, a, b, c))
            e = h2 * membrane_energy(ld3(x, a), ld3(x, b), ld3(x, c), Binv[t], Binv[nT + t], Binv[2 * nT + t], vol[t], lam, mu);
    }
    double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
__global__ void k_bend_energy(const double* __restrict__ x, const int* __restrict__ hg, int nH, const double* __restrict__ hi,
                              double kb, double h2, const uint8_t* __restrict__ dbc, double* __restrict__ partial) {
    __shared__ double lds[4];
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (h < nH) {
        int v0 = hg[h], v1 = hg[nH + h], v2 = hg[2 * nH + h], v3 = hg[3 * nH + h];
        if (!(dbc[v0] && dbc[v1] && dbc[v2] && dbc[v3])) {
            double th = dihedral_angle(ld3(x, v0), ld3(x, v1), ld3(x, v2), ld3(x, v3));
            double tb = hi[h], el = hi[nH + h], hb = hi[2 * nH + h];
            e = h2 * kb * (th - tb) * (th - tb) * el / hb;
        }
    }
    double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
// rod springs (ROD_STRETCHING.h:10-28) and rod hinges (ROD_BENDING.h:10-34); par:
// SoA [3][nR] = E, rest length, h^2 pi t^2 / 4; hw: h^2 E t^4 pi / (64 lsum)
__global__ void k_rod_spring_energy(const double* __restrict__ x, const int* __restrict__ rv, int nR, const double* __restrict__ par,
                                    double* __restrict__ partial) {
    __shared__ double lds[4];
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (r < nR) e = par[2 * nR + r] * rod_spring_energy(ld3(x, rv[r]), ld3(x, rv[nR + r]), par[r], par[nR + r]);
    double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
__global__ void k_rod_bend_energy(const double* __restrict__ x, const int* __restrict__ hv3, int nRH, const double* __restrict__ hw,
                                  double* __restrict__ partial) {
    __shared__ double lds[4];
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (h < nRH) e = hw[h] * rod_kappa2(ld3(x, hv3[h]), ld3(x, hv3[nRH + h]), ld3(x, hv3[2 * nRH + h]));
    double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
__global__ void k_inertia_energy(const double* __restrict__ x, const double* __restrict__ xt, const double* __restrict__ m, int nV,
                                 double* __restrict__ partial) {
    __shared__ double lds[4];
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (v < nV) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double d = x[3 * v + k] - xt[3 * v + k];
            e += 0.5 * m[v] * d * d;
        }
    }
    double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
// static solve: the body force's potential -sum x . (m g) over nodes with mass
// (INC_POTENTIAL.h:118-130; b = M g as in Augment_Mass_Matrix_And_Body_Force)
__global__ void k_body_energy(const double* __restrict__ x, const double* __restrict__ m, int nV, double gx, double gy, double gz,
                              double* __restrict__ partial) {
    __shared__ double lds[4];
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (v < nV && m[v] > 0) e = -(x[3 * v] * (m[v] * gx) + x[3 * v + 1] * (m[v] * gy) + x[3 * v + 2] * (m[v] * gz));
    double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

static double finish_sum(Ctx& c, const double* part, int nb, int slot) {
    sum_partials(part, nb, c.scal.p + slot, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + slot, c.scal.p + slot, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[slot];
}

double elastic_energy(Ctx& c, const double* x, double* memb, double* bend) {
    const double h2 = shell_h2(c);
    double Em = 0, Eb = 0;
    if (c.nT && !fiber_on(c)) {   // the fibre model replaces the membrane (INC_POTENTIAL.h:81-92)
        int nb = nblocks(c.nT);
        double* part = c.red.ensure(nb);
        hipLaunchKernelGGL(k_memb_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.tri.p, c.nT, c.Binv.p, c.vol.p, c.lam, c.mu, h2, c.dbc.p, part);
        CIPC_LAUNCH_CHECK();
        Em = finish_sum(c, part, nb, 3);
    }
    if (c.nH && c.P.bendMult != 0) {
        int nb = nblocks(c.nH);
        double* part = c.red.ensure(nb);
        hipLaunchKernelGGL(k_bend_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.hinge.p, c.nH, c.hInfo.p, c.kBend, h2, c.dbc.p, part);
        CIPC_LAUNCH_CHECK();
        Eb = finish_sum(c, part, nb, 4);
    }
    // rods: springs join the stretching energy, rod hinges the bending energy
    if (c.nRod) {
        int nb = nblocks(c.nRod);
        double* part = c.red.ensure(nb);
        hipLaunchKernelGGL(k_rod_spring_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.rod.p, c.nRod, c.rodPar.p, part);
        CIPC_LAUNCH_CHECK();
        Em += finish_sum(c, part, nb, 3);
    }
    if (c.nRH) {
        int nb = nblocks(c.nRH);
        double* part = c.red.ensure(nb);
        hipLaunchKernelGGL(k_rod_bend_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.rodHinge.p, c.nRH, c.rodHW.p, part);
        CIPC_LAUNCH_CHECK();
        Eb += finish_sum(c, part, nb, 4);
    }
    if (c.nTet) Em += tet_energy(c, x);   // +inf when a tet is inverted
    if (memb) *memb = Em;
    if (bend) *bend = Eb;
    return Em + Eb;
}
double inertia_energy(Ctx& c, const double* x) {
    int nb = nblocks(c.nV);
    double* part = c.red.ensure(nb);
    if (c.P.staticSolve)
        hipLaunchKernelGGL(k_body_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.mass.p, c.nV, c.P.grav[0], c.P.grav[1], c.P.grav[2], part);
    else
        hipLaunchKernelGGL(k_inertia_energy, dim3(nb), dim3(kBlock), 0, c.s, x, c.xt.p, c.mass.p, c.nV, part);
    CIPC_LAUNCH_CHECK();
    return finish_sum(c, part, nb, 5);
}

// ------------------------------------------------------------------ gradient
__global__ void k_memb_grad_local(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ Binv,
                                  const double* __restrict__ vol, double lam, double mu, double h2, const uint8_t* __restrict__ dbc,
                                  double* __restrict__ gl, long ns) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    double g[9];
    if (all_dbc3(dbc, a, b, c)) {
#pragma unroll
        for (int k = 0; k < 9; ++k) g[k] = 0;
    } else {
        membrane_grad(ld3(x, a), ld3(x, b), ld3(x, c), Binv[t], Binv[nT + t], Binv[2 * nT + t], vol[t], lam, mu, g);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) gl[3 * (3l * t + q) + k] = h2 * g[3 * q + k];
}
__global__ void k_bend_grad_local(const double* __restrict__ x, const int* __restrict__ hg, int nH, const double* __restrict__ hi,
                                  double kb, double h2, const uint8_t* __restrict__ dbc, double* __restrict__ gl, long ns, long off) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nH) return;
    int v[4] = {hg[h], hg[nH + h], hg[2 * nH + h], hg[3 * nH + h]};
    double g[12];
    if (dbc[v[0]] && dbc[v[1]] && dbc[v[2]] && dbc[v[3]]) {
#pragma unroll
        for (int k = 0; k < 12; ++k) g[k] = 0;
    } else {
        d3 p0 = ld3(x, v[0]), p1 = ld3(x, v[1]), p2 = ld3(x, v[2]), p3 = ld3(x, v[3]);
        double th = dihedral_angle(p0, p1, p2, p3);
        dihedral_angle_grad(p0, p1, p2, p3, g);
        double s = h2 * kb * 2 * (th - hi[h]) * hi[nH + h] / hi[2 * nH + h];
#pragma unroll
        for (int k = 0; k < 12; ++k) g[k] *= s;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) gl[3 * (off + 4l * h + q) + k] = g[3 * q + k];
}
__global__ void k_rod_grad_local(const double* __restrict__ x, const int* __restrict__ rv, int nR, const double* __restrict__ par,
                                 double* __restrict__ gl, long ns, long off) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nR) return;
    double g[6];
    rod_spring_grad(ld3(x, rv[r]), ld3(x, rv[nR + r]), par[r], par[nR + r], g);
    const double w = par[2 * nR + r];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) gl[3 * (off + 2l * r + q) + k] = w * g[3 * q + k];
}
__global__ void k_rod_bend_grad_local(const double* __restrict__ x, const int* __restrict__ hv3, int nRH, const double* __restrict__ hw,
                                      double* __restrict__ gl, long ns, long off) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nRH) return;
    double g[9];
    rod_kappa2_grad(ld3(x, hv3[h]), ld3(x, hv3[nRH + h]), ld3(x, hv3[2 * nRH + h]), g);
    const double w = hw[h];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) gl[3 * (off + 3l * h + q) + k] = w * g[3 * q + k];
}
// g[v] = sum of local slots + m (x - xt)
__global__ void k_el_grad_gather(const double* __restrict__ gl, long ns, const int* __restrict__ off, const int* __restrict__ slot,
                                 const double* __restrict__ x, const double* __restrict__ xt, const double* __restrict__ m, int nV,
                                 int withInertia, double gx, double gy, double gz, double* __restrict__ g) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    double a = 0, b = 0, d = 0;
    for (int k = off[v]; k < off[v + 1]; ++k) {
        long q = slot[k];
        a += gl[3 * q]; b += gl[3 * q + 1]; d += gl[3 * q + 2];
    }
    if (withInertia == 2 && m[v] > 0) {   // static solve: - m g (INC_POTENTIAL.h:233-249)
        a -= m[v] * gx; b -= m[v] * gy; d -= m[v] * gz;
    } else if (withInertia == 1) {
        a += m[v] * (x[3 * v] - xt[3 * v]);
        b += m[v] * (x[3 * v + 1] - xt[3 * v + 1]);
        d += m[v] * (x[3 * v + 2] - xt[3 * v + 2]);
    }
    g[3 * v] = a; g[3 * v + 1] = b; g[3 * v + 2] = d;
}

static void elastic_grad_impl(Ctx& c, const double* x, double* g, int withInertia) {
    const double h2 = shell_h2(c);
    long ns = c.nGSlots;
    double* gl = c.gLocal.ensure(3 * std::max<long>(ns, 1));
    if (c.nT && fiber_on(c)) {
        fiber_grad_local(c, x, gl);        // INC_POTENTIAL.h:213-216
        strain_limit_grad_add(c, x, gl);
    } else if (c.nT) {
        hipLaunchKernelGGL(k_memb_grad_local, dim3(nblocks(c.nT)), dim3(kBlock), 0, c.s, x, c.tri.p, c.nT, c.Binv.p, c.vol.p, c.lam, c.mu,
                           h2, c.dbc.p, gl, ns);
        CIPC_LAUNCH_CHECK();
        strain_limit_grad_add(c, x, gl);   // INC_POTENTIAL.h:220-222
    }
    if (c.nH) {
        double kb = c.P.bendMult != 0 ? c.kBend : 0.0;
        hipLaunchKernelGGL(k_bend_grad_local, dim3(nblocks(c.nH)), dim3(kBlock), 0, c.s, x, c.hinge.p, c.nH, c.hInfo.p, kb, h2, c.dbc.p,
                           gl, ns, 3l * c.nT);
        CIPC_LAUNCH_CHECK();
    }
    if (c.nRod) {
        hipLaunchKernelGGL(k_rod_grad_local, dim3(nblocks(c.nRod)), dim3(kBlock), 0, c.s, x, c.rod.p, c.nRod, c.rodPar.p, gl, ns,
                           3l * c.nT + 4l * c.nH);
        CIPC_LAUNCH_CHECK();
    }
    if (c.nRH) {
        hipLaunchKernelGGL(k_rod_bend_grad_local, dim3(nblocks(c.nRH)), dim3(kBlock), 0, c.s, x, c.rodHinge.p, c.nRH, c.rodHW.p, gl, ns,
                           3l * c.nT + 4l * c.nH + 2l * c.nRod);
        CIPC_LAUNCH_CHECK();
    }
    tet_grad_local(c, x, gl, 3l * c.nT + 4l * c.nH + 2l * c.nRod + 3l * c.nRH);
    hipLaunchKernelGGL(k_el_grad_gather, dim3(nblocks(c.nV)), dim3(kBlock), 0, c.s, gl, ns, c.gplanOff.p, c.gplanSlot.p, x, c.xt.p,
                       c.mass.p, c.nV, withInertia && c.P.staticSolve ? 2 : withInertia, c.P.grav[0], c.P.grav[1], c.P.grav[2], g);
    CIPC_LAUNCH_CHECK();
}
void elastic_gradient(Ctx& c, const double* x, double* g) { elastic_grad_impl(c, x, g, 1); }
void elastic_gradient_only(Ctx& c, const double* x, double* g) { elastic_grad_impl(c, x, g, 0); }

// ------------------------------------------------------------------ Hessian
// local blocks, one 3x3 block (9 contiguous doubles, row-major) per slot, so that the
// assembly gathers each slot with one or two cache lines; triangle t block (a, b) at slot
// 9 t + 3 a + b, hinge h block (a, b) at 9 nT + 16 h + 4 a + b
__global__ void __launch_bounds__(64) k_memb_hess_local(const double* __restrict__ x, const int* __restrict__ tri, int nT,
                                                        const double* __restrict__ Binv, const double* __restrict__ vol, double lam,
                                                        double mu, double h2, const uint8_t* __restrict__ dbc, double* __restrict__ hl,
                                                        long ns) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    double* dst = hl + 81l * t;   // block (p, q) at 9 (3 p + q), entry (k, l) at 3 k + l
    if (all_dbc3(dbc, a, b, c)) {
#pragma unroll
        for (int k = 0; k < 81; ++k) dst[k] = 0;
        return;
    }
    double H[81];
    membrane_hess(ld3(x, a), ld3(x, b), ld3(x, c), Binv[t], Binv[nT + t], Binv[2 * nT + t], vol[t], lam, mu, H);
#pragma unroll
    for (int k = 0; k < 81; ++k) H[k] *= h2;
    // the membrane energy depends on edge vectors only: translation-reduced projection
    make_pd_translation<3>(H, 9, [&](int r, int cc, double v) { dst[9 * (3 * (r / 3) + cc / 3) + 3 * (r % 3) + cc % 3] = v; });
}
// Hinge Hessians in two kernels, so that the register-hungry row evaluation and the
// eigensolver each run at their own occupancy (Math/UTILS.h:10-27 for the projection).
// Stage 1: the symmetrised, scaled hinge Hessian, row r by lane r, into H[144 h + 12 r + c]
// (zero for inactive hinges).
__global__ void __launch_bounds__(256) k_bend_hess_rows(const double* __restrict__ x, const int* __restrict__ hg, int nH,
                                                        const double* __restrict__ hi, double kb, double h2,
                                                        const uint8_t* __restrict__ dbc, double* __restrict__ H) {
    __shared__ double sh[16][144];
    const int grp = threadIdx.x >> 4, r = threadIdx.x & 15;
    const int h = blockIdx.x * 16 + grp;
    const bool live = h < nH;
    double* M = sh[grp];
    int v[4] = {0, 0, 0, 0};
    if (live) { v[0] = hg[h]; v[1] = hg[nH + h]; v[2] = hg[2 * nH + h]; v[3] = hg[3 * nH + h]; }
    const bool active = live && kb != 0.0 && !(dbc[v[0]] && dbc[v[1]] && dbc[v[2]] && dbc[v[3]]);
    double A[12], g[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) { A[c] = 0; g[c] = 0; }
    double s1 = 0, s2 = 0;
    if (active) {
        d3 p0 = ld3(x, v[0]), p1 = ld3(x, v[1]), p2 = ld3(x, v[2]), p3 = ld3(x, v[3]);
        const double th = dihedral_angle(p0, p1, p2, p3);
        dihedral_angle_grad(p0, p1, p2, p3, g);
        s1 = h2 * kb * 2 * (th - hi[h]) * hi[nH + h] / hi[2 * nH + h];
        s2 = h2 * kb * 2 * hi[nH + h] / hi[2 * nH + h];
        if (r < 12) dihedral_hess_row(p0, p1, p2, p3, r, A);
    }
    if (r < 12)
#pragma unroll
        for (int c = 0; c < 12; ++c) M[r * 12 + c] = A[c];
    __syncthreads();
    double gr = 0;
#pragma unroll
    for (int c = 0; c < 12; ++c) gr = (c == r) ? g[c] : gr;
    if (live && r < 12)
#pragma unroll
        for (int c = 0; c < 12; ++c) H[144l * h + 12 * r + c] = s1 * (0.5 * (M[r * 12 + c] + M[c * 12 + r])) + s2 * gr * g[c];
}
// Stage 2, one hinge per thread: the dihedral energy is invariant under translations, so
// H (x + t) = 0 for every rigid translation t = (u, u, u, u), and with the orthonormal
// basis Q = Wq (x) I3 of the translation complement (Wq: the three +-1/2 patterns
// orthogonal to (1, 1, 1, 1)) the projection is H+ = Q proj(Q^T H Q) Q^T.  The 9x9
// problem fits in registers (45 upper entries + 81 eigenvector entries): cyclic Jacobi,
// fully unrolled over the 36 pairs of a sweep, stopping at off^2 <= 1e-24 ||H9||_F^2.
__device__ __forceinline__ constexpr int u9(int i, int j) { return i <= j ? i * 9 - i * (i - 1) / 2 + (j - i) : j * 9 - j * (j - 1) / 2 + (i - j); }
__device__ __forceinline__ constexpr int wsign(int k, int a) {
    // Wq[k][a] = wsign / 2: columns (+,-,+,-), (+,+,-,-), (+,-,-,+)
    return a == 0 ? ((k & 1) ? -1 : 1) : a == 1 ? ((k >> 1) ? -1 : 1) : (((k & 1) ^ (k >> 1)) ? -1 : 1);
}
__global__ void __launch_bounds__(64) k_bend_hess_proj9(const double* __restrict__ H, int nH, double* __restrict__ hl, long off) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nH) return;
    const double* Hh = H + 144l * h;
    // H9 = Q^T H Q, upper triangle: block (a, b), entry (i, j) = 1/4 sum_kl s_ka s_lb H[3k+i][3l+j]
    double a9[45];
#pragma unroll
    for (int e = 0; e < 45; ++e) a9[e] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double v = 0.25 * Hh[(3 * k + i) * 12 + 3 * l + j];
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) {
                            const int R = 3 * a + i, C = 3 * b + j;
                            if (R <= C) a9[u9(R, C)] += (wsign(k, a) * wsign(l, b) > 0) ? v : -v;
                        }
                }
    double V[81];
#pragma unroll
    for (int e = 0; e < 81; ++e) V[e] = (e % 10 == 0) ? 1.0 : 0.0;
    double fro = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = i; j < 9; ++j) fro += (i == j ? 1.0 : 2.0) * a9[u9(i, j)] * a9[u9(i, j)];
    const double tol = 1e-24 * fro;
    for (int sweep = 0; sweep < 16; ++sweep) {
        double offd = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = i + 1; j < 9; ++j) offd += a9[u9(i, j)] * a9[u9(i, j)];
        if (offd <= 0.5 * tol) break;   // off^2 of the full matrix is twice the upper sum
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int q = p + 1; q < 9; ++q) {
                const double apq = a9[u9(p, q)];
                if (apq == 0.0) continue;
                const double app = a9[u9(p, p)], aqq = a9[u9(q, q)];
                const double theta = (aqq - app) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                a9[u9(p, p)] = app - t * apq;
                a9[u9(q, q)] = aqq + t * apq;
                a9[u9(p, q)] = 0.0;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    if (k == p || k == q) continue;
                    const double akp = a9[u9(k, p)], akq = a9[u9(k, q)];
                    a9[u9(k, p)] = c * akp - s * akq;
                    a9[u9(k, q)] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const double vkp = V[9 * k + p], vkq = V[9 * k + q];
                    V[9 * k + p] = c * vkp - s * vkq;
                    V[9 * k + q] = s * vkp + c * vkq;
                }
            }
    }
    // H9+ = V diag(max(lambda, 0)) V^T (upper), reusing a9
    double lam[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) lam[i] = a9[u9(i, i)] > 0 ? a9[u9(i, i)] : 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = i; j < 9; ++j) {
            double sum = 0;
#pragma unroll
            for (int m = 0; m < 9; ++m) sum += V[9 * i + m] * lam[m] * V[9 * j + m];
            a9[u9(i, j)] = sum;
        }
    // H+ = Q H9+ Q^T: block (k, l), entry (i, j) = 1/4 sum_ab s_ka s_lb H9+[3a+i][3b+j],
    // written as the hinge's 16 local 3x3 blocks
    double* dst = hl + 9 * (off + 16l * h);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double sum = 0;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) {
                            const double v = a9[u9(3 * a + i, 3 * b + j)];
                            sum += (wsign(k, a) * wsign(l, b) > 0) ? v : -v;
                        }
                    dst[9 * (4 * k + l) + 3 * i + j] = 0.25 * sum;
                }
}

// rod segment s block (a, b) at off + 4 s + 2 a + b; rod hinge h block (a, b) at
// off2 + 9 h + 3 a + b
__global__ void k_rod_hess_local(const double* __restrict__ x, const int* __restrict__ rv, int nR, const double* __restrict__ par,
                                 double* __restrict__ hl, long ns, long off) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nR) return;
    double K[9];
    rod_spring_hess_pd(ld3(x, rv[r]), ld3(x, rv[nR + r]), par[r], par[nR + r], K);
    const double w = par[2 * nR + r];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const double sg = (p == q) ? w : -w;
#pragma unroll
            for (int e = 0; e < 9; ++e) hl[9 * (off + 4l * r + 2 * p + q) + e] = sg * K[e];
        }
}
__global__ void __launch_bounds__(64) k_rod_bend_hess_local(const double* __restrict__ x, const int* __restrict__ hv3, int nRH,
                                                            const double* __restrict__ hw, double* __restrict__ hl, long ns, long off) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nRH) return;
    double H[81];
    rod_kappa2_hess(ld3(x, hv3[h]), ld3(x, hv3[nRH + h]), ld3(x, hv3[2 * nRH + h]), H);
    // |kappa|^2 depends on the two edge vectors only: translation-reduced projection (the
    // positive weight commutes with it)
    const double w = hw[h];
    double* dst = hl + 9 * (off + 9l * h);
    make_pd_translation<3>(H, 9, [&](int r, int c, double v) { dst[9 * (3 * (r / 3) + c / 3) + 3 * (r % 3) + c % 3] = w * v; });
}
__global__ void k_el_assemble(const double* __restrict__ hl, long ns, const int* __restrict__ off, const int* __restrict__ slot,
                              const int* __restrict__ rowOf, const int* __restrict__ col, int nnzb, const double* __restrict__ m,
                              const uint8_t* __restrict__ dbc, double* __restrict__ val) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nnzb) return;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = off[b]; k < off[b + 1]; ++k) {
        long q = slot[k];
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] += hl[9 * q + e];
    }
    int r = rowOf[b], cl = col[b];
    if (m && r == cl) { acc[0] += m[r]; acc[4] += m[r]; acc[8] += m[r]; }   // no mass in a static solve
    if (dbc[r] || dbc[cl]) {
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] = (r == cl && (e == 0 || e == 4 || e == 8)) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) val[(long)e * nnzb + b] = acc[e];
}

void elastic_hessian(Ctx& c, const double* x) {
    const double h2 = shell_h2(c);
    long ns = c.nElSlots;
    double* hl = c.elLocal.ensure(9 * std::max<long>(ns, 1));
    if (c.nT && fiber_on(c)) {
        fiber_hess_local(c, x, hl);        // INC_POTENTIAL.h:340-343
        strain_limit_hess_add(c, x, hl);
    } else if (c.nT) {
        hipLaunchKernelGGL(k_memb_hess_local, dim3(nblocks(c.nT, 64)), dim3(64), 0, c.s, x, c.tri.p, c.nT, c.Binv.p, c.vol.p, c.lam, c.mu,
                           h2, c.dbc.p, hl, ns);
        CIPC_LAUNCH_CHECK();
        strain_limit_hess_add(c, x, hl);   // INC_POTENTIAL.h:347-349
    }
    if (c.nH) {
        double kb = c.P.bendMult != 0 ? c.kBend : 0.0;
        double* Hh = c.hingeH.ensure(144 * (size_t)c.nH);
        hipLaunchKernelGGL(k_bend_hess_rows, dim3((c.nH + 15) / 16), dim3(256), 0, c.s, x, c.hinge.p, c.nH, c.hInfo.p, kb, h2, c.dbc.p, Hh);
        hipLaunchKernelGGL(k_bend_hess_proj9, dim3(nblocks(c.nH, 64)), dim3(64), 0, c.s, (const double*)Hh, c.nH, hl, 9l * c.nT);
        CIPC_LAUNCH_CHECK();
    }
    if (c.nRod) {
        hipLaunchKernelGGL(k_rod_hess_local, dim3(nblocks(c.nRod)), dim3(kBlock), 0, c.s, x, c.rod.p, c.nRod, c.rodPar.p, hl, ns,
                           9l * c.nT + 16l * c.nH);
        CIPC_LAUNCH_CHECK();
    }
    if (c.nRH) {
        hipLaunchKernelGGL(k_rod_bend_hess_local, dim3(nblocks(c.nRH, 64)), dim3(64), 0, c.s, x, c.rodHinge.p, c.nRH, c.rodHW.p, hl, ns,
                           9l * c.nT + 16l * c.nH + 4l * c.nRod);
        CIPC_LAUNCH_CHECK();
    }
    tet_hess_local(c, x, hl, 9l * c.nT + 16l * c.nH + 4l * c.nRod + 9l * c.nRH);
    BSR& A = c.Ael;
    hipLaunchKernelGGL(k_el_assemble, dim3(nblocks(A.nnzb)), dim3(kBlock), 0, c.s, hl, ns, c.planOff.p, c.planSlot.p, A.rowOf.p,
                       A.col.p, A.nnzb, c.P.staticSolve ? nullptr : c.mass.p, proj_mask(c), A.val.p);
    CIPC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ static plans (host, once per scene)
void build_elastic_plans(Ctx& c) {
    const int nT = c.nT, nH = c.nH, nV = c.nV;
    // gradient plan: slot -> vertex
    {
        const int nR = c.nRod, nRH = c.nRH, nTet = c.nTet;
        long ns = 3l * nT + 4l * nH + 2l * nR + 3l * nRH + 4l * nTet;
        c.nGSlots = (int)ns;
        std::vector<std::pair<int, int>> vs;
        vs.reserve(ns);
        for (int t = 0; t < nT; ++t)
            for (int a = 0; a < 3; ++a) vs.push_back({c.hTri[3 * t + a], (int)(3l * t + a)});
        for (int h = 0; h < nH; ++h)
            for (int a = 0; a < 4; ++a) vs.push_back({c.hHinge[4 * h + a], (int)(3l * nT + 4l * h + a)});
        const long oR = 3l * nT + 4l * nH, oRH = oR + 2l * nR;
        for (int r = 0; r < nR; ++r)
            for (int a = 0; a < 2; ++a) vs.push_back({c.hRod[2 * r + a], (int)(oR + 2l * r + a)});
        for (int h = 0; h < nRH; ++h)
            for (int a = 0; a < 3; ++a) vs.push_back({c.hRodHinge[3 * h + a], (int)(oRH + 3l * h + a)});
        const long oTet = oRH + 3l * nRH;
        for (int t = 0; t < nTet; ++t)
            for (int a = 0; a < 4; ++a) vs.push_back({c.hTet[4 * t + a], (int)(oTet + 4l * t + a)});
        std::sort(vs.begin(), vs.end());
        std::vector<int> off(nV + 1, 0), slot(ns);
        for (auto& p : vs) off[p.first + 1]++;
        for (int v = 0; v < nV; ++v) off[v + 1] += off[v];
        for (long i = 0; i < ns; ++i) slot[i] = vs[i].second;
        upload(c.gplanOff, off, c.s);
        upload(c.gplanSlot, slot, c.s);
    }
    // Hessian pattern and plan
    {
        const int nR = c.nRod, nRH = c.nRH, nTet = c.nTet;
        long ns = 9l * nT + 16l * nH + 4l * nR + 9l * nRH + 16l * nTet;
        c.nElSlots = (int)ns;
        std::vector<std::pair<uint64_t, int>> ks;
        ks.reserve(ns + nV);
        const long oR = 9l * nT + 16l * nH, oRH = oR + 4l * nR;
        for (int r = 0; r < nR; ++r)
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    ks.push_back({(uint64_t)c.hRod[2 * r + a] * nV + c.hRod[2 * r + b], (int)(oR + 4l * r + 2 * a + b)});
        for (int h = 0; h < nRH; ++h)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    ks.push_back({(uint64_t)c.hRodHinge[3 * h + a] * nV + c.hRodHinge[3 * h + b], (int)(oRH + 9l * h + 3 * a + b)});
        for (int t = 0; t < nT; ++t)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    ks.push_back({(uint64_t)c.hTri[3 * t + a] * nV + c.hTri[3 * t + b], (int)(9l * t + 3 * a + b)});
        for (int h = 0; h < nH; ++h)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    ks.push_back({(uint64_t)c.hHinge[4 * h + a] * nV + c.hHinge[4 * h + b], (int)(9l * nT + 16l * h + 4 * a + b)});
        const long oTet = oRH + 9l * nRH;
        for (int t = 0; t < nTet; ++t)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    ks.push_back({(uint64_t)c.hTet[4 * t + a] * nV + c.hTet[4 * t + b], (int)(oTet + 16l * t + 4 * a + b)});
        for (int v = 0; v < nV; ++v) ks.push_back({(uint64_t)v * nV + v, -1});
        std::sort(ks.begin(), ks.end());
        std::vector<int> rowOf, col, off, slot;
        for (size_t i = 0; i < ks.size();) {
            size_t j = i;
            off.push_back((int)slot.size());
            while (j < ks.size() && ks[j].first == ks[i].first) {
                if (ks[j].second >= 0) slot.push_back(ks[j].second);
                ++j;
            }
            rowOf.push_back((int)(ks[i].first / nV));
            col.push_back((int)(ks[i].first % nV));
            i = j;
        }
        off.push_back((int)slot.size());
        int nnzb = (int)rowOf.size();
        std::vector<int> rowptr(nV + 1, 0);
        for (int b = 0; b < nnzb; ++b) rowptr[rowOf[b] + 1]++;
        for (int v = 0; v < nV; ++v) rowptr[v + 1] += rowptr[v];
        if (slot.empty()) slot.push_back(0);
        BSR& A = c.Ael;
        A.nrow = nV;
        A.nnzb = nnzb;
        upload(A.rowptr, rowptr, c.s);
        upload(A.rowOf, rowOf, c.s);
        upload(A.col, col, c.s);
        A.val.ensure_exact(9l * nnzb);
        upload(c.planOff, off, c.s);
        upload(c.planSlot, slot, c.s);
    }
    CIPC_HIP(hipStreamSynchronize(c.s));
}

}  // namespace cipc
