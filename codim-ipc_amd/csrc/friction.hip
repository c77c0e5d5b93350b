// Lagged, C1-smoothed friction of the IPC contact model on the device.
//
// friction_basis restates Compute_Friction_Basis (FEM/FRICTION.h:17-112): the friction
// stencils are the constraints without a mollifier (FRICTION.h:37-41), in constraint
// order (flag + exclusive scan + scatter); per stencil the closest-point coordinates
// (FEM/FRICTION_UTILS.h:55-66, 123-144, 204-212), a 3x2 tangent basis (:43-53,
// :111-121, :192-202, :242-259) and the normal force -b'(d^2 - xi^2) 2 d
// (FRICTION.h:104-107).  friction_energy / friction_gradient / friction_hess_local
// restate Compute_Friction_Potential / _Gradient / _Hessian (FRICTION.h:173-661) with
// the smoothing of FRICTION_UTILS.h:9-37.
//
// One thread per stencil.  The relative tangential displacement is
// u = B^T sum_a c_a (x_a - xn_a) with per-vertex weights c_a; the gradient of vertex a
// is c_a B w (w = mu lambda f1(|u|)/|u| u) and the Hessian block (a, b) is
// c_a c_b B M B^T with the 2x2 matrix M of the reference, so each stencil costs one 3x3
// product instead of a 12x12 one.  Gradients go through the contact gradient's
// deterministic vertex gather; Hessian blocks join the contact BSR assembly.
//
// Layout: stencils SoA [4][nF] (reference int4 encoding), data SoA [9][nF]:
// cp0, cp1, basis column 0 (xyz), basis column 1 (xyz), normal force.
#include "ctx.hpp"
#include "dev_math.hpp"

namespace cipc {

#pragma clang fp contract(off)

struct FricSt {
    int nv, vid[4];
    double coef[4], mult;
};

CIPC_DEV void fric_stencil(const int* c, double cp0, double cp1, FricSt& s) {
    s.mult = 1;
    if (c[0] >= 0) {   // edge-edge
        s.nv = 4;
        for (int i = 0; i < 4; ++i) s.vid[i] = c[i];
        s.coef[0] = 1.0 - cp0; s.coef[1] = cp0; s.coef[2] = cp1 - 1.0; s.coef[3] = -cp1;
        return;
    }
    s.vid[0] = -c[0] - 1;
    s.vid[1] = c[1];
    s.vid[2] = c[2];
    s.vid[3] = c[3];
    if (c[2] < 0) {    // point-point
        s.nv = 2; s.mult = -c[3];
        s.coef[0] = 1; s.coef[1] = -1; s.coef[2] = 0; s.coef[3] = 0;
    } else if (c[3] < 0) {   // point-edge
        s.nv = 3; s.mult = -c[3];
        s.coef[0] = 1; s.coef[1] = cp0 - 1.0; s.coef[2] = -cp0; s.coef[3] = 0;
    } else {           // point-triangle
        s.nv = 4;
        s.coef[0] = 1; s.coef[1] = -1.0 + cp0 + cp1; s.coef[2] = -cp0; s.coef[3] = -cp1;
    }
}

CIPC_DEV d3 nrmz(d3 a) { return (1.0 / sqrt(sq(a))) * a; }

__global__ void k_fric_flag(const int* __restrict__ C, int nC, uint32_t* __restrict__ head) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC) return;
    const int c0 = C[i], c2 = C[2l * nC + i], c3 = C[3l * nC + i];
    head[i] = (c0 >= 0 && (c2 < 0 || c3 < 0)) ? 0u : 1u;
}

// getCompI (FRICTION.h:114-124): first component whose range end exceeds the vertex
CIPC_DEV int fric_comp(const int* __restrict__ range, int nComp, int v) {
    for (int i = 0; i < nComp; ++i)
        if (v < range[i]) return i;
    return nComp - 1;
}

__global__ void k_fric_basis(const double* __restrict__ x, const int* __restrict__ C, int nC, const uint32_t* __restrict__ head,
                             const uint32_t* __restrict__ pos, int nF, double kappa, double xi2, double shat, int* __restrict__ fC,
                             double* __restrict__ fI, const int* __restrict__ compRange, int nComp, const double* __restrict__ muComp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC || !head[i]) return;
    const int o = (int)pos[i];
    int c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[k] = C[(long)k * nC + i];
        fC[(long)k * nF + o] = c[k];
    }
    double cp0 = 0, cp1 = 0, d2;
    d3 b0, b1;
    if (c[0] >= 0) {
        d3 a0 = ld3(x, c[0]), a1 = ld3(x, c[1]), q0 = ld3(x, c[2]), q1 = ld3(x, c[3]);
        d3 ea = a1 - a0, eb = q1 - q0, w = a0 - q0;
        double m00 = sq(ea), m01 = -dot(eb, ea), m11 = sq(eb);
        double r0 = -dot(w, ea), r1 = dot(w, eb);
        double det = m00 * m11 - m01 * m01;
        cp0 = (m11 * r0 - m01 * r1) / det;
        cp1 = (m00 * r1 - m01 * r0) / det;
        b0 = nrmz(ea);
        b1 = nrmz(cross(cross(ea, eb), ea));
        d2 = dist2_ee(a0, a1, q0, q1);
    } else {
        d3 p = ld3(x, -c[0] - 1);
        if (c[2] < 0) {
            d3 p1 = ld3(x, c[1]);
            d3 d = p1 - p;
            d3 xc = cross(mk(1, 0, 0), d), yc = cross(mk(0, 1, 0), d);
            d3 t = sq(xc) > sq(yc) ? xc : yc;
            b0 = nrmz(t);
            b1 = nrmz(cross(d, t));
            d2 = dist2_pp(p, p1);
        } else if (c[3] < 0) {
            d3 e0 = ld3(x, c[1]), e1 = ld3(x, c[2]);
            d3 e = e1 - e0;
            cp0 = dot(p - e0, e) / sq(e);
            b0 = nrmz(e);
            b1 = nrmz(cross(e, p - e0));
            d2 = dist2_pe(p, e0, e1);
        } else {
            d3 t0 = ld3(x, c[1]), t1 = ld3(x, c[2]), t2 = ld3(x, c[3]);
            d3 e1 = t1 - t0, e2 = t2 - t0, w = p - t0;
            double m00 = sq(e1), m01 = dot(e1, e2), m11 = sq(e2);
            double r0 = dot(e1, w), r1 = dot(e2, w);
            double det = m00 * m11 - m01 * m01;
            cp0 = (m11 * r0 - m01 * r1) / det;
            cp1 = (m00 * r1 - m01 * r0) / det;
            b0 = nrmz(e1);
            b1 = nrmz(cross(cross(e1, e2), e1));
            d2 = dist2_pt(p, t0, t1, t2);
        }
    }
    double lam = -barrier_g(d2 - xi2, shat, kappa) * 2 * sqrt(d2);
    if (nComp) {   // Compute_Friction_Coef (FRICTION.h:127-171): coefficient of the two primitives' components
        const int pa = c[0] >= 0 ? c[0] : -c[0] - 1, pb = c[0] >= 0 ? c[2] : c[1];
        lam *= muComp[fric_comp(compRange, nComp, pa) + fric_comp(compRange, nComp, pb) * nComp];
    }
    const double v[9] = {cp0, cp1, b0.x, b0.y, b0.z, b1.x, b1.y, b1.z, lam};
#pragma unroll
    for (int k = 0; k < 9; ++k) fI[(long)k * nF + o] = v[k];
}

#pragma clang fp contract(on)

// stencil i: weights, basis and u = B^T sum_a c_a (x_a - xn_a)
struct FricEval {
    FricSt s;
    d3 b0, b1;
    double lam, u0, u1, q;
};
CIPC_DEV void fric_eval(const double* __restrict__ x, const double* __restrict__ xn, const int* __restrict__ fC,
                        const double* __restrict__ fI, int nF, int i, FricEval& e) {
    int c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = fC[(long)k * nF + i];
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = fI[(long)k * nF + i];
    fric_stencil(c, v[0], v[1], e.s);
    d3 rel = mk(0, 0, 0);
    for (int a = 0; a < e.s.nv; ++a) rel = rel + e.s.coef[a] * (ld3(x, e.s.vid[a]) - ld3(xn, e.s.vid[a]));
    e.b0 = mk(v[2], v[3], v[4]);
    e.b1 = mk(v[5], v[6], v[7]);
    e.lam = v[8];
    e.u0 = dot(e.b0, rel);
    e.u1 = dot(e.b1, rel);
    e.q = e.u0 * e.u0 + e.u1 * e.u1;
}
CIPC_DEV double f0_smooth(double q, double ev) {
    if (q >= ev * ev) return sqrt(q);
    return q * (-sqrt(q) / 3.0 + ev) / (ev * ev) + ev / 3.0;
}
CIPC_DEV double f1_div_norm(double q, double ev) {
    if (q >= ev * ev) return 1.0 / sqrt(q);
    return (-sqrt(q) + 2.0 * ev) / (ev * ev);
}

__global__ void k_fric_energy(const double* __restrict__ x, const double* __restrict__ xn, const int* __restrict__ fC,
                              const double* __restrict__ fI, int nF, double ev, double* __restrict__ partial) {
    __shared__ double lds[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (i < nF) {
        FricEval E;
        fric_eval(x, xn, fC, fI, nF, i, E);
        e = E.s.mult * E.lam * f0_smooth(E.q, ev);
    }
    double t = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// local gradients SoA [3][4 nF] and the vertex of each slot (-1 unused)
__global__ void k_fric_grad_local(const double* __restrict__ x, const double* __restrict__ xn, const int* __restrict__ fC,
                                  const double* __restrict__ fI, int nF, double ev, double mu, double* __restrict__ gl,
                                  int* __restrict__ gv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nF) return;
    FricEval E;
    fric_eval(x, xn, fC, fI, nF, i, E);
    const double s = f1_div_norm(E.q, ev) * mu * E.s.mult * E.lam;
    const d3 w = (s * E.u0) * E.b0 + (s * E.u1) * E.b1;
    const long ns = 4l * nF;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const bool used = a < E.s.nv;
        gv[4l * i + a] = used ? E.s.vid[a] : -1;
        const double ca = used ? E.s.coef[a] : 0.0;
        gl[4l * i + a] = ca * w.x;
        gl[ns + 4l * i + a] = ca * w.y;
        gl[2 * ns + 4l * i + a] = ca * w.z;
    }
}

// Hessian blocks of stencil i at slots off + 16 i + 4 a + b of the SoA [9][ns] array
__global__ void k_fric_hess_local(const double* __restrict__ x, const double* __restrict__ xn, const int* __restrict__ fC,
                                  const double* __restrict__ fI, int nF, double ev, double mu, double* __restrict__ hl, long ns,
                                  long off) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nF) return;
    FricEval E;
    fric_eval(x, xn, fC, fI, nF, i, E);
    const double sc = mu * E.s.mult * E.lam;
    const double f1d = f1_div_norm(E.q, ev);
    double M0, M1, M3;
    if (E.q >= ev * ev) {
        // sliding: f1/|u| / |u|^2 ubar ubar^T, ubar = (-u1, u0)
        const double c = sc * f1d / E.q;
        M0 = c * E.u1 * E.u1; M1 = -c * E.u1 * E.u0; M3 = c * E.u0 * E.u0;
    } else if (E.q == 0) {
        M0 = M3 = sc * f1d; M1 = 0;
    } else {
        const double un = sqrt(E.q), f2 = -1.0 / (ev * ev);
        double a = f2 / un * E.u0 * E.u0 + f1d, b = f2 / un * E.u0 * E.u1, d = f2 / un * E.u1 * E.u1 + f1d;
        // 2x2 PSD clamp (Math/UTILS.h:10-27, n = 2)
        const double hh = 0.5 * (a - d), r = sqrt(hh * hh + b * b);
        const double l0 = 0.5 * (a + d) - r, l1 = 0.5 * (a + d) + r;
        if (l0 < 0) {
            double vx, vy;
            if (r == 0) { vx = 1; vy = 0; }
            else if (hh >= 0) { vx = hh + r; vy = b; }
            else { vx = b; vy = r - hh; }
            const double n = sqrt(vx * vx + vy * vy);
            vx /= n; vy /= n;
            const double l = l1 > 0 ? l1 : 0.0;
            a = l * vx * vx; b = l * vx * vy; d = l * vy * vy;
        }
        M0 = sc * a; M1 = sc * b; M3 = sc * d;
    }
    const double B[3][2] = {{E.b0.x, E.b1.x}, {E.b0.y, E.b1.y}, {E.b0.z, E.b1.z}};
    double K[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q)
            K[r * 3 + q] = B[r][0] * (M0 * B[q][0] + M1 * B[q][1]) + B[r][1] * (M1 * B[q][0] + M3 * B[q][1]);
    const long base = off + 16l * i;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double cab = (a < E.s.nv && b < E.s.nv) ? E.s.coef[a] * E.s.coef[b] : 0.0;
#pragma unroll
            for (int e = 0; e < 9; ++e) hl[(long)e * ns + base + 4 * a + b] = cab * K[e];
        }
}

// BSR keys of the friction blocks (row * nV + col; unused or Dirichlet: nV * nV)
__global__ void k_fric_hess_keys(const int* __restrict__ fC, int nF, int nV, const uint8_t* __restrict__ dbc,
                                 uint64_t* __restrict__ key) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 16l * nF) return;
    const int i = (int)(q / 16), a = (int)((q / 4) % 4), b = (int)(q % 4);
    int c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = fC[(long)k * nF + i];
    FricSt s;
    fric_stencil(c, 0.0, 0.0, s);
    const bool keep = a < s.nv && b < s.nv && !dbc[s.vid[a]] && !dbc[s.vid[b]];
    key[q] = keep ? (uint64_t)s.vid[a] * (uint64_t)nV + (uint64_t)s.vid[b] : (uint64_t)nV * (uint64_t)nV;
}

// ------------------------------------------------------------------ host side
static double fric_ev(const Ctx& c) { return std::sqrt(c.P.epsv2 * c.P.dt * c.P.dt); }
// the coefficient multiplying every stencil: mu, or 1 when it comes per component
static double fric_mu(const Ctx& c) { return c.nFricComp ? 1.0 : c.P.fricMu; }

void friction_basis(Ctx& c, const double* x, const int* C, int nC, double kappa) {
    c.nF = 0;
    if (nC == 0) return;
    uint32_t* head = c.u32a.ensure(nC);
    uint32_t* pos = c.u32b.ensure(nC);
    hipLaunchKernelGGL(k_fric_flag, dim3(nblocks(nC)), dim3(kBlock), 0, c.s, C, nC, head);
    CIPC_LAUNCH_CHECK();
    scan_exclusive(head, pos, nC, c.counters.p + 40, c.pw, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 40, c.counters.p + 40, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    const int nF = (int)c.hCnt[40];
    c.nF = nF;
    if (nF == 0) return;
    c.fC.ensure(4l * nF);
    c.fI.ensure(9l * nF);
    const double xi2 = c.P.xi * c.P.xi, shat = c.dHat2 + 2 * std::sqrt(c.dHat2) * c.P.xi;
    hipLaunchKernelGGL(k_fric_basis, dim3(nblocks(nC)), dim3(kBlock), 0, c.s, x, C, nC, (const uint32_t*)head,
                       (const uint32_t*)pos, nF, kappa, xi2, shat, c.fC.p, c.fI.p, c.fricCompRange.p, c.nFricComp, c.fricMuComp.p);
    CIPC_LAUNCH_CHECK();
}

double friction_energy(Ctx& c, const double* x, const double* xn) {
    if (c.nF == 0) return 0.0;
    const int nb = nblocks(c.nF);
    double* part = c.red.ensure(nb);
    hipLaunchKernelGGL(k_fric_energy, dim3(nb), dim3(kBlock), 0, c.s, x, xn, c.fC.p, c.fI.p, c.nF, fric_ev(c), part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 48, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 48, c.scal.p + 48, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return fric_mu(c) * c.hScal[48];
}

void friction_gradient(Ctx& c, const double* x, const double* xn, double* g, const uint8_t* skip) {
    if (c.nF == 0) return;
    const long ns = 4l * c.nF;
    double* gl = c.fGLocal.ensure(3 * ns);
    int* gv = c.fGVert.ensure(ns);
    hipLaunchKernelGGL(k_fric_grad_local, dim3(nblocks(c.nF)), dim3(kBlock), 0, c.s, x, xn, c.fC.p, c.fI.p, c.nF, fric_ev(c),
                       fric_mu(c), gl, gv);
    CIPC_LAUNCH_CHECK();
    scatter_local_gradient(c, gl, gv, ns, g, skip);
}

void friction_hess_local(Ctx& c, const double* x, const double* xn, double* hl, long ns, long off) {
    if (c.nF == 0) return;
    hipLaunchKernelGGL(k_fric_hess_local, dim3(nblocks(c.nF)), dim3(kBlock), 0, c.s, x, xn, c.fC.p, c.fI.p, c.nF, fric_ev(c),
                       fric_mu(c), hl, ns, off);
    CIPC_LAUNCH_CHECK();
}

void friction_hess_keys(Ctx& c, uint64_t* key) {
    if (c.nF == 0) return;
    hipLaunchKernelGGL(k_fric_hess_keys, dim3(nblocks(16l * c.nF)), dim3(kBlock), 0, c.s, c.fC.p, c.nF, c.nV, proj_mask(c), key);
    CIPC_LAUNCH_CHECK();
}

}  // namespace cipc
