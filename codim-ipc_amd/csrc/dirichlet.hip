// Moving Dirichlet nodes on the device.
//
// step_dirichlet restates Step_Dirichlet (FEM/BOUNDARY_CONDITION.h:107-155): every
// motion group rotates its targets about its centre and axis by angVelDeg h (Rodrigues'
// matrix of the axis as given, as Eigen::AngleAxis builds it) and then translates them
// by v h.  The targets stay resident in device memory; the step reads them there.
//
// When the CCD does not allow the whole Dirichlet displacement, the step keeps the
// nodes where they are and pulls them with the mass-weighted penalty
// 1/2 k sum m |x - target|^2 (Shell/DIRICHLET.h:26-93, IMPLICIT_EULER.h:384-401),
// doubling k up to 1e8 until 99% of the distance is covered (:712-768).  Meanwhile
// only the nodes whose target equals their position are projected out (DBCb_fixed,
// :307-326); dbc_fixed_mask builds that mask.
#include "ctx.hpp"

namespace cipc {

__global__ void k_dbc_motion(int begin, int end, double* __restrict__ tgt, double r00, double r01, double r02, double r10,
                             double r11, double r12, double r20, double r21, double r22, double cx, double cy, double cz,
                             int rotate, double dx, double dy, double dz) {
    const int i = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    double* t = tgt + 3l * i;
    double x = t[0], y = t[1], z = t[2];
    if (rotate) {
        const double a = x - cx, b = y - cy, c = z - cz;
        x = r00 * a + r01 * b + r02 * c + cx;
        y = r10 * a + r11 * b + r12 * c + cy;
        z = r20 * a + r21 * b + r22 * c + cz;
    }
    t[0] = x + dx;
    t[1] = y + dy;
    t[2] = z + dz;
}

// per Dirichlet entry: m |x - target|^2 (withMass) or |x - target|^2, block partials
__global__ void k_dbc_dist2(int n, const int* __restrict__ idx, const double* __restrict__ tgt, const double* __restrict__ x,
                            const double* __restrict__ mass, int withMass, double* __restrict__ partial) {
    __shared__ double lds[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (i < n) {
        const int v = idx[i];
        double d2 = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double d = tgt[3 * i + k] - x[3 * v + k];
            d2 += d * d;
        }
        e = withMass ? d2 * mass[v] : d2;
    }
    const double s = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_dbc_pen_grad(int n, const int* __restrict__ idx, const double* __restrict__ tgt, const double* __restrict__ x,
                               const double* __restrict__ mass, double k, double* __restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    const double km = k * mass[v];
#pragma unroll
    for (int c = 0; c < 3; ++c) g[3 * v + c] += km * (x[3 * v + c] - tgt[3 * i + c]);
}

// k m added to the diagonal block of each Dirichlet row of the elastic BSR (values SoA [9][nnzb])
__global__ void k_dbc_pen_hess(int n, const int* __restrict__ idx, const double* __restrict__ mass, double k,
                               const int* __restrict__ rowptr, const int* __restrict__ col, int nnzb, double* __restrict__ val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    for (int b = rowptr[v]; b < rowptr[v + 1]; ++b)
        if (col[b] == v) {
            const double km = k * mass[v];
            val[0l * nnzb + b] += km;
            val[4l * nnzb + b] += km;
            val[8l * nnzb + b] += km;
            return;
        }
}

__global__ void k_dbc_fixed(int n, const int* __restrict__ idx, const double* __restrict__ tgt, const double* __restrict__ x,
                            uint8_t* __restrict__ fixed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    fixed[v] = (tgt[3 * i] == x[3 * v] && tgt[3 * i + 1] == x[3 * v + 1] && tgt[3 * i + 2] == x[3 * v + 2]) ? 1 : 0;
}

// ------------------------------------------------------------------ host side
void dirichlet_device_init(Ctx& c) {
    if (c.dbcDevReady || c.nDBC == 0) return;
    c.dbcIdxD.ensure_exact(c.nDBC);
    c.dbcTgt.ensure_exact(3l * c.nDBC);
    CIPC_HIP(hipMemcpyAsync(c.dbcIdxD.p, c.hDbcIdx.data(), c.nDBC * sizeof(int), hipMemcpyHostToDevice, c.s));
    CIPC_HIP(hipMemcpyAsync(c.dbcTgt.p, c.hDbcTarget.data(), 3l * c.nDBC * sizeof(double), hipMemcpyHostToDevice, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    c.dbcDevReady = true;
}

void step_dirichlet(Ctx& c) {
    dirichlet_device_init(c);
    const double h = c.P.dt;
    for (const DBCMotion& M : c.dbcMotion) {
        if (M.end <= M.begin) continue;
        const double th = M.angVelDeg / 180 * M_PI * h;
        const double co = std::cos(th), sn = std::sin(th);
        const double* a = M.axis;
        hipLaunchKernelGGL(k_dbc_motion, dim3(nblocks(M.end - M.begin)), dim3(kBlock), 0, c.s, M.begin, M.end, c.dbcTgt.p,
                           co + (1 - co) * a[0] * a[0], (1 - co) * a[0] * a[1] - sn * a[2], (1 - co) * a[0] * a[2] + sn * a[1],
                           (1 - co) * a[1] * a[0] + sn * a[2], co + (1 - co) * a[1] * a[1], (1 - co) * a[1] * a[2] - sn * a[0],
                           (1 - co) * a[2] * a[0] - sn * a[1], (1 - co) * a[2] * a[1] + sn * a[0], co + (1 - co) * a[2] * a[2],
                           M.center[0], M.center[1], M.center[2], M.angVelDeg != 0 ? 1 : 0, M.v[0] * h, M.v[1] * h, M.v[2] * h);
        CIPC_LAUNCH_CHECK();
    }
    CIPC_HIP(hipStreamSynchronize(c.s));
}

static double dbc_sum(Ctx& c, const double* x, int withMass) {
    const int nb = nblocks(c.nDBC);
    double* part = c.red.ensure(nb);
    hipLaunchKernelGGL(k_dbc_dist2, dim3(nb), dim3(kBlock), 0, c.s, c.nDBC, c.dbcIdxD.p, c.dbcTgt.p, x, c.mass.p, withMass, part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 50, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 50, c.scal.p + 50, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[50];
}
double dbc_dist2(Ctx& c, const double* x) { return c.nDBC ? dbc_sum(c, x, 0) : 0.0; }
double dbc_penalty_energy(Ctx& c, const double* x) { return (c.nDBC && c.dbcStiff) ? 0.5 * c.dbcStiff * dbc_sum(c, x, 1) : 0.0; }

void dbc_penalty_grad(Ctx& c, const double* x, double* g) {
    if (!c.nDBC || !c.dbcStiff) return;
    hipLaunchKernelGGL(k_dbc_pen_grad, dim3(nblocks(c.nDBC)), dim3(kBlock), 0, c.s, c.nDBC, c.dbcIdxD.p, c.dbcTgt.p, x, c.mass.p,
                       c.dbcStiff, g);
    CIPC_LAUNCH_CHECK();
}
void dbc_penalty_hess(Ctx& c) {
    if (!c.nDBC || !c.dbcStiff) return;
    const BSR& A = c.Ael;
    hipLaunchKernelGGL(k_dbc_pen_hess, dim3(nblocks(c.nDBC)), dim3(kBlock), 0, c.s, c.nDBC, c.dbcIdxD.p, c.mass.p, c.dbcStiff,
                       A.rowptr.p, A.col.p, A.nnzb, A.val.p);
    CIPC_LAUNCH_CHECK();
}
void dbc_fixed_mask(Ctx& c, const double* x) {
    c.dbcFixed.ensure(c.nV);
    CIPC_HIP(hipMemsetAsync(c.dbcFixed.p, 0, c.nV, c.s));
    if (!c.nDBC) return;
    hipLaunchKernelGGL(k_dbc_fixed, dim3(nblocks(c.nDBC)), dim3(kBlock), 0, c.s, c.nDBC, c.dbcIdxD.p, c.dbcTgt.p, x, c.dbcFixed.p);
    CIPC_LAUNCH_CHECK();
}

}  // namespace cipc
