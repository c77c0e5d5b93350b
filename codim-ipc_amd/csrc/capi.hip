// C ABI (include/cipc.h).  Every entry point converts HIP failures and invalid
// arguments into negative status codes and records a message for cipc_last_error().
#include "ctx.hpp"
#include <stdexcept>
#include <string>
#include <cstdio>
#include <cmath>
#include "learn.hpp"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

using namespace cipc;

struct cipc_ctx {
    Ctx c;
    DevBuf<double> qx, qp, qb, qsol, qout;
    DevBuf<int> qC;
    unsigned char* scratch = nullptr;   // device address of this device's `scratch` array
};

// ---- static scratch of the allocation-free entry point (cipc_barrier_energy_noalloc)
// One array per GPU, alive until the process ends.  Its host-side name is only a stand-in:
// device code reaches it through the address hipGetSymbolAddress returns for the current
// device, resolved once per device when the first context is created there.
constexpr size_t kScratchBytes = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
unsigned char* g_scratchPtr[kMaxDevices] = {};

unsigned char* resolve_scratch() {
    int dev = 0;
    CIPC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* ptr = nullptr;
        CIPC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratchPtr[dev] = static_cast<unsigned char*>(ptr);
    }
    return g_scratchPtr[dev];
}
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

namespace cipc {
void ctx_setup(Ctx& c, const std::vector<double>& X, const std::vector<int>& tri);
void elastic_gradient_only(Ctx& c, const double* x, double* g);
}

// kappa[0] > 0 turns the upper limit on (needs s[0] > shat[0] >= 0); kappa[1] > 0 the
// lower one (needs s[1] < shat[1]); kappa[1] alone is ignored, as in INEXT.h:157-175
static bool strain_limit_args_ok(const double* k, const double* s, const double* sh) {
    if (k[0] < 0 || k[1] < 0) return false;
    if (k[0] > 0 && !(s[0] > sh[0] && sh[0] >= 0)) return false;
    if (k[0] > 0 && k[1] > 0 && !(s[1] < sh[1] && s[1] >= 0 && sh[1] < s[0])) return false;
    return true;
}

// fibre model: stiffness multipliers >= 0; when on, every limit > 0
static bool fiber_args_ok(const double* k, const double* lim) {
    for (int i = 0; i < 4; ++i) if (k[i] < 0) return false;
    if (k[0] > 0 || k[1] > 0)
        for (int i = 0; i < 3; ++i) if (!(lim[i] > 0)) return false;
    return true;
}

template <class F>
static int guarded(cipc_ctx* ctx, F f) {
    try {
        return f();
    } catch (const HipError& e) {
        if (ctx) ctx->c.lastError = e.what();
        return CIPC_E_HIP;
    } catch (const std::exception& e) {
        if (ctx) ctx->c.lastError = e.what();
        return CIPC_E_INVALID;
    }
}

static void upload_x(cipc_ctx* ctx, DevBuf<double>& b, const double* h, size_t n) {
    b.ensure(n);
    CIPC_HIP(hipMemcpyAsync(b.p, h, n * sizeof(double), hipMemcpyHostToDevice, ctx->c.s));
}
// constraints arrive as int4 rows; the device keeps them SoA [4][n]
// Every entry of the reference encoding is a vertex v, a -v-1, a -multiplicity or -1
// (IPC.h:144-661), so all lie in [-nV, nV); anything else would be read out of bounds on
// the device and is rejected here (CIPC_E_INVALID through guarded()).
static void upload_cs(cipc_ctx* ctx, const int32_t* cs, int32_t n) {
    const int nV = ctx->c.nV;
    for (size_t i = 0; i < 4 * (size_t)n; ++i)
        if (cs[i] < -nV || cs[i] >= nV)
            throw std::invalid_argument("constraint " + std::to_string(i / 4) + " has an entry outside [-n_vertices, n_vertices)");
    std::vector<int> soa(4 * (size_t)std::max(n, 1));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) soa[(size_t)k * n + i] = cs[4 * (size_t)i + k];
    ctx->qC.ensure(soa.size());
    CIPC_HIP(hipMemcpyAsync(ctx->qC.p, soa.data(), soa.size() * sizeof(int), hipMemcpyHostToDevice, ctx->c.s));
    CIPC_HIP(hipStreamSynchronize(ctx->c.s));
}

extern "C" {

int cipc_create(const cipc_scene_desc* d, cipc_ctx** out) {
    if (!d || !out || d->n_vertices <= 0 || !d->x0 || d->n_triangles < 0 || (d->n_triangles && !d->triangles) ||
        d->n_dbc < 0 || (d->n_dbc && (!d->dbc_index || !d->dbc_target)))
        return CIPC_E_INVALID;
    for (int i = 0; i < 3 * d->n_triangles; ++i)
        if (d->triangles[i] < 0 || d->triangles[i] >= d->n_vertices) return CIPC_E_INVALID;
    for (int i = 0; i < d->n_dbc; ++i)
        if (d->dbc_index[i] < 0 || d->dbc_index[i] >= d->n_vertices) return CIPC_E_INVALID;
    if (d->n_rod_segments < 0 || (d->n_rod_segments && !d->rod_segments) || d->n_particles < 0 ||
        (d->n_particles && !d->particles))
        return CIPC_E_INVALID;
    for (int i = 0; i < 2 * d->n_rod_segments; ++i)
        if (d->rod_segments[i] < 0 || d->rod_segments[i] >= d->n_vertices) return CIPC_E_INVALID;
    for (int i = 0; i < d->n_rod_segments; ++i)
        if (d->rod_segments[2 * i] == d->rod_segments[2 * i + 1]) return CIPC_E_INVALID;
    if (d->n_tets < 0 || (d->n_tets && !d->tets)) return CIPC_E_INVALID;
    for (int i = 0; i < 4 * d->n_tets; ++i)
        if (d->tets[i] < 0 || d->tets[i] >= d->n_vertices) return CIPC_E_INVALID;
    if (d->n_tets && !(d->tet_density >= 0 && d->tet_youngs > 0 && d->tet_poisson > -1 && d->tet_poisson < 0.5)) return CIPC_E_INVALID;
    for (int i = 0; i < d->n_particles; ++i)
        if (d->particles[i] < 0 || d->particles[i] >= d->n_vertices) return CIPC_E_INVALID;
    if (d->n_vertices >= (1 << 21) - 1) return CIPC_E_INVALID;   // contact-key packing limit
    if (d->friction_mu < 0 || (d->friction_mu > 0 && !(d->friction_epsv2 > 0))) return CIPC_E_INVALID;
    if (!strain_limit_args_ok(d->strain_limit_kappa, d->strain_limit_s, d->strain_limit_shat)) return CIPC_E_INVALID;
    if (!fiber_args_ok(d->fiber_stiff_mult, d->fiber_limit)) return CIPC_E_INVALID;
    cipc_ctx* ctx = new cipc_ctx();
    int rc = guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        Params& P = c.P;
        P.rho = d->density; P.E = d->youngs; P.nu = d->poisson; P.thick = d->thickness; P.bendMult = d->bending_stiff_mult;
        P.dt = d->dt; P.dHat = d->dhat; P.xi = d->offset; P.PNTol = d->newton_tol;
        for (int k = 0; k < 3; ++k) P.grav[k] = d->gravity[k];
        P.kappaStiffMult = d->kappa_stiff_mult;
        P.pcgTol = d->pcg_tol > 0 ? d->pcg_tol : 1e-5;
        P.pcgMaxIter = d->pcg_max_iter > 0 ? d->pcg_max_iter : 100000;
        P.collision = d->with_collision;
        P.rodRho = d->rod_density; P.rodE = d->rod_youngs; P.rodBendMult = d->rod_bend_stiff_mult; P.rodThick = d->rod_thickness;
        P.partRho = d->particle_density; P.partThick = d->particle_thickness;
        P.fricMu = d->friction_mu;
        P.epsv2 = d->friction_mu > 0 ? d->friction_epsv2 : 1e-6;
        P.fricIterAmt = d->friction_iter_amt;
        for (int i = 0; i < 4; ++i) P.fiberK[i] = d->fiber_stiff_mult[i];
        for (int i = 0; i < 3; ++i) P.fiberLimit[i] = d->fiber_limit[i];
        if (d->strain_limit_kappa[0] > 0)
            for (int i = 0; i < 2; ++i) {
                P.slKappa[i] = d->strain_limit_kappa[i]; P.slS[i] = d->strain_limit_s[i]; P.slSHat[i] = d->strain_limit_shat[i];
            }
        c.hRod.assign(d->rod_segments, d->rod_segments + 2 * (size_t)d->n_rod_segments);
        c.hTet.assign(d->tets, d->tets + 4 * (size_t)d->n_tets);
        P.tetRho = d->tet_density; P.tetE = d->tet_youngs; P.tetNu = d->tet_poisson;
        c.hPart.assign(d->particles, d->particles + (size_t)d->n_particles);
        CIPC_HIP(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));
        ctx->scratch = resolve_scratch();
        CIPC_HIP(hipHostMalloc((void**)&c.hScal, 64 * sizeof(double), hipHostMallocDefault));
        CIPC_HIP(hipHostMalloc((void**)&c.hCnt, 64 * sizeof(uint32_t), hipHostMallocDefault));
        c.nV = d->n_vertices;
        c.nDBC = d->n_dbc;
        c.hDbcIdx.assign(d->dbc_index, d->dbc_index + d->n_dbc);
        c.hDbcTarget.assign(d->dbc_target, d->dbc_target + 3 * (size_t)d->n_dbc);
        std::vector<double> X(d->x0, d->x0 + 3 * (size_t)c.nV);
        std::vector<int> T(d->triangles, d->triangles + 3 * (size_t)d->n_triangles);
        ctx_setup(c, X, T);
        CIPC_HIP(hipMemcpyAsync(c.x.p, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
        if (d->v0) CIPC_HIP(hipMemcpyAsync(c.v.p, d->v0, X.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
        else CIPC_HIP(hipMemsetAsync(c.v.p, 0, X.size() * sizeof(double), c.s));
        CIPC_HIP(hipMemcpyAsync(c.xt.p, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
    if (rc != CIPC_OK) {
        cipc_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return CIPC_OK;
}

void cipc_destroy(cipc_ctx* ctx) {
    if (!ctx) return;
    Ctx& c = ctx->c;
    if (c.s) (void)hipStreamSynchronize(c.s);
    for (hipEvent_t e : c.evPool) (void)hipEventDestroy(e);
    if (c.hScal) (void)hipHostFree(c.hScal);
    if (c.hCnt) (void)hipHostFree(c.hCnt);
    hipStream_t s = c.s;
    delete ctx;
    if (s) (void)hipStreamDestroy(s);
}

const char* cipc_last_error(const cipc_ctx* ctx) { return ctx ? ctx->c.lastError.c_str() : "null context"; }

int cipc_get_sizes(const cipc_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    const Ctx& c = ctx->c;
    out[0] = c.nV; out[1] = c.nT; out[2] = c.nBN; out[3] = c.nBE; out[4] = c.nH; out[5] = c.nC;
    return CIPC_OK;
}
int cipc_get_scalars(const cipc_ctx* ctx, double* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    const Ctx& c = ctx->c;
    out[0] = c.kappa; out[1] = c.kappaMax; out[2] = c.dHat2; out[3] = c.lam; out[4] = c.mu; out[5] = c.kBend;
    return CIPC_OK;
}
int cipc_get_setup(const cipc_ctx* ctx, double* mass, int32_t* hinges, double* hinge_info, int32_t* bedges) {
    if (!ctx) return CIPC_E_INVALID;
    const Ctx& c = ctx->c;
    if (mass) std::memcpy(mass, c.hMass.data(), c.hMass.size() * sizeof(double));
    if (hinges) std::memcpy(hinges, c.hHinge.data(), c.hHinge.size() * sizeof(int));
    if (hinge_info) std::memcpy(hinge_info, c.hHingeInfo.data(), c.hHingeInfo.size() * sizeof(double));
    if (bedges) std::memcpy(bedges, c.hBEdge.data(), c.hBEdge.size() * sizeof(int));
    return CIPC_OK;
}
int cipc_get_codim_sizes(const cipc_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    const Ctx& c = ctx->c;
    out[0] = c.nRod; out[1] = c.nRH; out[2] = c.nPart; out[3] = c.codimBN0; out[4] = c.codimBN1;
    return CIPC_OK;
}
int cipc_get_codim_setup(const cipc_ctx* ctx, double* rod_info, int32_t* rod_hinges, double* rod_hinge_info, int32_t* bnodes) {
    if (!ctx) return CIPC_E_INVALID;
    const Ctx& c = ctx->c;
    if (rod_info) std::memcpy(rod_info, c.hRodInfo.data(), c.hRodInfo.size() * sizeof(double));
    if (rod_hinges) std::memcpy(rod_hinges, c.hRodHinge.data(), c.hRodHinge.size() * sizeof(int));
    if (rod_hinge_info) std::memcpy(rod_hinge_info, c.hRodHingeInfo.data(), c.hRodHingeInfo.size() * sizeof(double));
    if (bnodes) std::memcpy(bnodes, c.hBNode.data(), c.hBNode.size() * sizeof(int));
    return CIPC_OK;
}
int cipc_set_state(cipc_ctx* ctx, const double* x, const double* v) {
    if (!ctx) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        if (x) CIPC_HIP(hipMemcpyAsync(c.x.p, x, n * sizeof(double), hipMemcpyHostToDevice, c.s));
        if (v) CIPC_HIP(hipMemcpyAsync(c.v.p, v, n * sizeof(double), hipMemcpyHostToDevice, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
}
int cipc_get_state(const cipc_ctx* cctx, double* x, double* v) {
    cipc_ctx* ctx = const_cast<cipc_ctx*>(cctx);
    if (!ctx) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        if (x) CIPC_HIP(hipMemcpyAsync(x, c.x.p, n * sizeof(double), hipMemcpyDeviceToHost, c.s));
        if (v) CIPC_HIP(hipMemcpyAsync(v, c.v.p, n * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
}
int cipc_set_kappa(cipc_ctx* ctx, double kappa) {
    if (!ctx || !(kappa > 0)) return CIPC_E_INVALID;
    ctx->c.kappa = kappa;
    return CIPC_OK;
}

int cipc_step(cipc_ctx* ctx, cipc_step_stats* stats) {
    if (!ctx) return CIPC_E_INVALID;
    int rc = guarded(ctx, [&]() { return advance_one_step(ctx->c); });
    ctx->c.stats.status = rc;
    if (stats) *stats = ctx->c.stats;
    return rc;
}

int cipc_batch_step(cipc_ctx* const* ctxs, int32_t n, int32_t threads, cipc_step_stats* stats, double* step_ms) {
    if (!ctxs || n < 0) return CIPC_E_INVALID;
    for (int i = 0; i < n; ++i)
        if (!ctxs[i]) return CIPC_E_INVALID;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (ctxs[i] == ctxs[j]) return CIPC_E_INVALID;   // one scene, one stepping thread
    if (n == 0) return CIPC_OK;
    int T = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    T = std::min(T, (int)n);
    std::atomic<int> next(0), firstErr(0);
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n) return;
            const auto t0 = std::chrono::steady_clock::now();
            cipc_step_stats st{};
            const int rc = cipc_step(ctxs[i], &st);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (stats) stats[i] = st;
            if (step_ms) step_ms[i] = ms;
            if (rc != CIPC_OK) {
                int z = 0;
                firstErr.compare_exchange_strong(z, rc);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return firstErr.load();
}

int cipc_constraint_set(cipc_ctx* ctx, const double* x, int32_t* out, int32_t cap, int32_t* n) {
    if (!ctx || !x || !n) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        upload_x(ctx, ctx->qx, x, 3 * (size_t)c.nV);
        int nC = 0;
        DevBuf<int>& outD = ctx->qC;
        constraint_set(c, ctx->qx.p, outD, nC);
        *n = nC;
        if (nC > cap || !out) return nC > cap ? CIPC_E_CAPACITY : CIPC_OK;
        std::vector<int> soa(4 * (size_t)std::max(nC, 1));
        CIPC_HIP(hipMemcpyAsync(soa.data(), outD.p, 4 * (size_t)nC * sizeof(int), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        for (int i = 0; i < nC; ++i)
            for (int k = 0; k < 4; ++k) out[4 * (size_t)i + k] = soa[(size_t)k * nC + i];
        return CIPC_OK;
    });
}

int cipc_barrier_energy(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, double* e) {
    if (!ctx || !x || !e || n_cs < 0 || (n_cs && !cs)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        upload_x(ctx, ctx->qx, x, 3 * (size_t)c.nV);
        upload_cs(ctx, cs, n_cs);
        int bad = 0;
        *e = barrier_energy(c, ctx->qx.p, ctx->qC.p, n_cs, kappa, &bad);
        if (bad) { c.lastError = "non-positive constraint distance"; return CIPC_E_PENETRATION; }
        return CIPC_OK;
    });
}

int cipc_barrier_energy_noalloc(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, double* e) {
    if (!ctx || !x || !e || n_cs < 0 || (n_cs && !cs)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        unsigned char* base = ctx->scratch;
        if (!base) throw std::runtime_error("the scratch array of this context's device is not resolved");
        // layout: x [3 nV doubles] | constraint rows [4 n_cs ints] | block partials | sum | flags
        const size_t nx = 3 * (size_t)c.nV, nb = (size_t)nblocks(n_cs);
        const size_t offC = align256(nx * sizeof(double));
        const size_t offP = offC + align256(4 * (size_t)n_cs * sizeof(int32_t));
        const size_t offO = offP + align256(std::max<size_t>(nb, 1) * sizeof(double));
        const size_t offB = offO + 256, need = offB + 256;
        if (need > kScratchBytes) {
            c.lastError = "cipc_barrier_energy_noalloc: " + std::to_string(c.nV) + " vertices and " + std::to_string(n_cs) +
                          " constraints need " + std::to_string(need) + " bytes of scratch, more than its " +
                          std::to_string(kScratchBytes) + "; use cipc_barrier_energy";
            return CIPC_E_CAPACITY;
        }
        const int nV = c.nV;
        for (size_t i = 0; i < 4 * (size_t)n_cs; ++i)   // same rule as every other constraint input
            if (cs[i] < -nV || cs[i] >= nV) {
                c.lastError = "constraint " + std::to_string(i / 4) + " has an entry outside [-n_vertices, n_vertices)";
                return CIPC_E_INVALID;
            }
        double* xd = reinterpret_cast<double*>(base);
        int* Cd = reinterpret_cast<int*>(base + offC);
        double* part = reinterpret_cast<double*>(base + offP);
        double* out = reinterpret_cast<double*>(base + offO);
        uint32_t* bad = reinterpret_cast<uint32_t*>(base + offB);
        CIPC_HIP(hipMemcpyAsync(xd, x, nx * sizeof(double), hipMemcpyHostToDevice, c.s));
        if (n_cs) CIPC_HIP(hipMemcpyAsync(Cd, cs, 4 * (size_t)n_cs * sizeof(int32_t), hipMemcpyHostToDevice, c.s));
        barrier_energy_rows(c, xd, Cd, n_cs, kappa, part, out, bad);
        CIPC_HIP(hipMemcpyAsync(c.hScal + 62, out, sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipMemcpyAsync(c.hCnt + 62, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        *e = c.hScal[62];
        if (c.hCnt[62] & 2u) { c.lastError = "constraint entry outside the vertex range"; return CIPC_E_INVALID; }
        if (c.hCnt[62]) { c.lastError = "non-positive constraint distance"; return CIPC_E_PENETRATION; }
        return CIPC_OK;
    });
}

int cipc_barrier_gradient(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, double* g) {
    if (!ctx || !x || !g || n_cs < 0 || (n_cs && !cs)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        upload_cs(ctx, cs, n_cs);
        double* gd = ctx->qout.ensure(n);
        CIPC_HIP(hipMemsetAsync(gd, 0, n * sizeof(double), c.s));
        barrier_gradient(c, ctx->qx.p, ctx->qC.p, n_cs, kappa, gd);
        CIPC_HIP(hipMemcpyAsync(g, gd, n * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
}

int cipc_elastic_energy(cipc_ctx* ctx, const double* x, double* out) {
    if (!ctx || !x || !out) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        upload_x(ctx, ctx->qx, x, 3 * (size_t)c.nV);
        elastic_energy(c, ctx->qx.p, &out[0], &out[1]);
        return CIPC_OK;
    });
}

int cipc_elastic_gradient(cipc_ctx* ctx, const double* x, double* g) {
    if (!ctx || !x || !g) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        double* gd = ctx->qout.ensure(n);
        elastic_gradient_only(c, ctx->qx.p, gd);
        CIPC_HIP(hipMemcpyAsync(g, gd, n * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
}

static void download_bsr(Ctx& c, const BSR& A, int32_t* brow, int32_t* bcol, double* bval, int off) {
    if (A.nnzb == 0) return;
    std::vector<int> r(A.nnzb), cl(A.nnzb);
    std::vector<double> v(9 * (size_t)A.nnzb);
    CIPC_HIP(hipMemcpyAsync(r.data(), A.rowOf.p, A.nnzb * sizeof(int), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipMemcpyAsync(cl.data(), A.col.p, A.nnzb * sizeof(int), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipMemcpyAsync(v.data(), A.val.p, v.size() * sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    for (int b = 0; b < A.nnzb; ++b) {
        brow[off + b] = r[b];
        bcol[off + b] = cl[b];
        for (int k = 0; k < 9; ++k) bval[9 * (size_t)(off + b) + k] = v[(size_t)k * A.nnzb + b];
    }
}

int cipc_hessian(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t which, int32_t* brow,
                 int32_t* bcol, double* bval, int32_t cap, int32_t* nblocks_out) {
    if (!ctx || !x || !nblocks_out || n_cs < 0 || (n_cs && !cs)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        upload_x(ctx, ctx->qx, x, 3 * (size_t)c.nV);
        int ne = 0, nk = 0;
        if (which & 1) { elastic_hessian(c, ctx->qx.p); ne = c.Ael.nnzb; }
        if (which & 2) {
            upload_cs(ctx, cs, n_cs);
            barrier_hessian(c, ctx->qx.p, ctx->qC.p, n_cs, kappa);
            nk = c.Acon.nnzb;
        }
        CIPC_HIP(hipStreamSynchronize(c.s));
        *nblocks_out = ne + nk;
        if (ne + nk > cap || !brow || !bcol || !bval) return (ne + nk > cap) ? CIPC_E_CAPACITY : CIPC_OK;
        if (which & 1) download_bsr(c, c.Ael, brow, bcol, bval, 0);
        if (which & 2) download_bsr(c, c.Acon, brow, bcol, bval, ne);
        return CIPC_OK;
    });
}

int cipc_min_dist2(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double* d2, double* min_val) {
    if (!ctx || !x || n_cs < 0 || (n_cs && !cs)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        upload_x(ctx, ctx->qx, x, 3 * (size_t)c.nV);
        upload_cs(ctx, cs, n_cs);
        double* dd = ctx->qout.ensure(std::max(n_cs, 1));
        double m = min_dist2(c, ctx->qx.p, ctx->qC.p, n_cs, dd);
        if (min_val) *min_val = m;
        if (d2 && n_cs) {
            CIPC_HIP(hipMemcpyAsync(d2, dd, n_cs * sizeof(double), hipMemcpyDeviceToHost, c.s));
            CIPC_HIP(hipStreamSynchronize(c.s));
        }
        return CIPC_OK;
    });
}

// ---- frame output
int cipc_write_obj(cipc_ctx* ctx, const char* path) {
    if (!ctx || !path) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        std::vector<double> x(3 * (size_t)c.nV);
        CIPC_HIP(hipMemcpyAsync(x.data(), c.x.p, x.size() * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        FILE* f = std::fopen(path, "w");
        if (!f) { c.lastError = std::string("cannot open ") + path; return CIPC_E_INVALID; }
        for (int i = 0; i < c.nV; ++i) std::fprintf(f, "v %le %le %le\n", x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        for (int t = 0; t < c.nT; ++t) std::fprintf(f, "f %d %d %d\n", c.hTri[3 * t] + 1, c.hTri[3 * t + 1] + 1, c.hTri[3 * t + 2] + 1);
        const bool ok = std::fclose(f) == 0;
        if (!ok) { c.lastError = std::string("write failed: ") + path; return CIPC_E_INVALID; }
        return CIPC_OK;
    });
}

// ---- moving Dirichlet nodes
int cipc_dirichlet_add_motion(cipc_ctx* ctx, int32_t begin, int32_t end, const double* v, const double* rot_center,
                              const double* rot_axis, double ang_vel_deg) {
    if (!ctx || !v || !rot_center || !rot_axis || begin < 0 || end < begin || end > ctx->c.nDBC) return CIPC_E_INVALID;
    DBCMotion M;
    M.begin = begin; M.end = end; M.angVelDeg = ang_vel_deg;
    for (int k = 0; k < 3; ++k) { M.v[k] = v[k]; M.center[k] = rot_center[k]; M.axis[k] = rot_axis[k]; }
    ctx->c.dbcMotion.push_back(M);
    return CIPC_OK;
}

int cipc_step_dirichlet(cipc_ctx* ctx) {
    if (!ctx) return CIPC_E_INVALID;
    return guarded(ctx, [&]() { step_dirichlet(ctx->c); return CIPC_OK; });
}

int cipc_turn_dirichlet(cipc_ctx* ctx) {
    if (!ctx) return CIPC_E_INVALID;
    for (DBCMotion& M : ctx->c.dbcMotion)
        for (int k = 0; k < 3; ++k) M.v[k] = -M.v[k];
    return CIPC_OK;
}

int cipc_get_dirichlet_targets(cipc_ctx* ctx, double* targets) {
    if (!ctx || (!targets && ctx->c.nDBC)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        if (!c.nDBC) return CIPC_OK;
        dirichlet_device_init(c);
        CIPC_HIP(hipMemcpyAsync(targets, c.dbcTgt.p, 3l * c.nDBC * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
}

int cipc_set_dirichlet_targets(cipc_ctx* ctx, const double* targets) {
    if (!ctx || (!targets && ctx->c.nDBC)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        if (!c.nDBC) return CIPC_OK;
        dirichlet_device_init(c);
        CIPC_HIP(hipMemcpyAsync(c.dbcTgt.p, targets, 3l * c.nDBC * sizeof(double), hipMemcpyHostToDevice, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
}

// ---- strain limiting
int cipc_set_strain_limit(cipc_ctx* ctx, const double* kappa, const double* s, const double* shat) {
    if (!ctx || !kappa || !s || !shat || !strain_limit_args_ok(kappa, s, shat)) return CIPC_E_INVALID;
    Params& P = ctx->c.P;
    for (int i = 0; i < 2; ++i) { P.slKappa[i] = kappa[i]; P.slS[i] = s[i]; P.slSHat[i] = shat[i]; }
    if (!(kappa[0] > 0)) P.slKappa[0] = P.slKappa[1] = 0;
    return CIPC_OK;
}

int cipc_get_tet_sizes(const cipc_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    out[0] = ctx->c.nTet;
    out[1] = ctx->c.nBT;
    return CIPC_OK;
}

int cipc_inversion_free_stepsize(cipc_ctx* ctx, const double* x, const double* p, double alpha, double* out) {
    if (!ctx || !x || !p || !out || !(alpha > 0)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        const size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        upload_x(ctx, ctx->qp, p, n);
        *out = inversion_free_stepsize(c, ctx->qx.p, ctx->qp.p, alpha);
        return CIPC_OK;
    });
}

int cipc_get_strain_limit(const cipc_ctx* ctx, double* kappa, double* s, double* shat) {
    if (!ctx) return CIPC_E_INVALID;
    const Params& P = ctx->c.P;
    for (int i = 0; i < 2; ++i) {
        if (kappa) kappa[i] = P.slKappa[i];
        if (s) s[i] = P.slS[i];
        if (shat) shat[i] = P.slSHat[i];
    }
    return CIPC_OK;
}

int cipc_strain_limit_energy(cipc_ctx* ctx, const double* x, double* e, int32_t* valid) {
    if (!ctx || !x || !e || !valid) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        upload_x(ctx, ctx->qx, x, 3 * (size_t)c.nV);
        const double E = strain_limit_energy(c, ctx->qx.p);
        *valid = std::isinf(E) ? 0 : 1;
        *e = *valid ? E : 0.0;
        return CIPC_OK;
    });
}

int cipc_set_fiber(cipc_ctx* ctx, const double* stiff_mult, const double* limit) {
    if (!ctx || !stiff_mult || !limit || !fiber_args_ok(stiff_mult, limit)) return CIPC_E_INVALID;
    Params& P = ctx->c.P;
    for (int i = 0; i < 4; ++i) P.fiberK[i] = stiff_mult[i];
    for (int i = 0; i < 3; ++i) P.fiberLimit[i] = limit[i];
    return CIPC_OK;
}

// ---- lagged friction
int cipc_set_friction(cipc_ctx* ctx, double mu, double epsv2, int32_t iter_amt) {
    if (!ctx || mu < 0 || (mu > 0 && !(epsv2 > 0))) return CIPC_E_INVALID;
    Params& P = ctx->c.P;
    P.fricMu = mu;
    if (mu > 0) P.epsv2 = epsv2;
    P.fricIterAmt = iter_amt;
    return CIPC_OK;
}

int cipc_set_friction_components(cipc_ctx* ctx, int32_t n_comp, const int32_t* comp_node_range, const double* mu_comp) {
    if (!ctx || n_comp < 0 || (n_comp && (!comp_node_range || !mu_comp))) return CIPC_E_INVALID;
    for (int i = 0; i < n_comp; ++i)
        if (comp_node_range[i] <= (i ? comp_node_range[i - 1] : 0)) return CIPC_E_INVALID;
    if (n_comp && comp_node_range[n_comp - 1] < ctx->c.nV) return CIPC_E_INVALID;   // getCompI exits on an uncovered vertex
    for (int i = 0; i < n_comp * n_comp; ++i)
        if (mu_comp[i] < 0) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        c.nFricComp = n_comp;
        if (n_comp) {
            c.fricCompRange.ensure_exact(n_comp);
            c.fricMuComp.ensure_exact((size_t)n_comp * n_comp);
            CIPC_HIP(hipMemcpyAsync(c.fricCompRange.p, comp_node_range, n_comp * sizeof(int), hipMemcpyHostToDevice, c.s));
            CIPC_HIP(hipMemcpyAsync(c.fricMuComp.p, mu_comp, (size_t)n_comp * n_comp * sizeof(double), hipMemcpyHostToDevice, c.s));
            CIPC_HIP(hipStreamSynchronize(c.s));
        }
        return CIPC_OK;
    });
}

int cipc_friction_basis(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t* out_cs,
                        double* out_info, int32_t cap, int32_t* n) {
    if (!ctx || !x || !n || n_cs < 0 || (n_cs && !cs) || cap < 0) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        upload_x(ctx, ctx->qx, x, 3 * (size_t)c.nV);
        upload_cs(ctx, cs, n_cs);
        friction_basis(c, ctx->qx.p, ctx->qC.p, n_cs, kappa);
        CIPC_HIP(hipStreamSynchronize(c.s));
        const int nF = c.nF;
        *n = nF;
        if (nF > cap) return CIPC_E_CAPACITY;
        if (nF && out_cs && out_info) {
            std::vector<int> fc(4 * (size_t)nF);
            std::vector<double> fi(9 * (size_t)nF);
            CIPC_HIP(hipMemcpyAsync(fc.data(), c.fC.p, fc.size() * sizeof(int), hipMemcpyDeviceToHost, c.s));
            CIPC_HIP(hipMemcpyAsync(fi.data(), c.fI.p, fi.size() * sizeof(double), hipMemcpyDeviceToHost, c.s));
            CIPC_HIP(hipStreamSynchronize(c.s));
            for (int i = 0; i < nF; ++i) {
                for (int k = 0; k < 4; ++k) out_cs[4 * (size_t)i + k] = fc[(size_t)k * nF + i];
                for (int k = 0; k < 9; ++k) out_info[9 * (size_t)i + k] = fi[(size_t)k * nF + i];
            }
        }
        return CIPC_OK;
    });
}

int cipc_friction_energy(cipc_ctx* ctx, const double* x, const double* xn, double* e) {
    if (!ctx || !x || !xn || !e) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        upload_x(ctx, ctx->qp, xn, n);
        *e = friction_energy(c, ctx->qx.p, ctx->qp.p);
        return CIPC_OK;
    });
}

int cipc_friction_gradient(cipc_ctx* ctx, const double* x, const double* xn, double* g) {
    if (!ctx || !x || !xn || !g) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        upload_x(ctx, ctx->qp, xn, n);
        double* gd = ctx->qout.ensure(n);
        CIPC_HIP(hipMemsetAsync(gd, 0, n * sizeof(double), c.s));
        friction_gradient(c, ctx->qx.p, ctx->qp.p, gd);
        CIPC_HIP(hipMemcpyAsync(g, gd, n * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        return CIPC_OK;
    });
}

int cipc_friction_hessian(cipc_ctx* ctx, const double* x, const double* xn, int32_t* brow, int32_t* bcol, double* bval,
                          int32_t cap, int32_t* nblocks_out) {
    if (!ctx || !x || !xn || !nblocks_out) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        upload_x(ctx, ctx->qp, xn, n);
        barrier_hessian(c, ctx->qx.p, nullptr, 0, 0.0, ctx->qp.p);
        CIPC_HIP(hipStreamSynchronize(c.s));
        *nblocks_out = c.Acon.nnzb;
        if (c.Acon.nnzb > cap || !brow || !bcol || !bval) return (c.Acon.nnzb > cap) ? CIPC_E_CAPACITY : CIPC_OK;
        download_bsr(c, c.Acon, brow, bcol, bval, 0);
        return CIPC_OK;
    });
}

int cipc_ccd_stepsize(cipc_ctx* ctx, const double* x, const double* p, double alpha_in, double* alpha_out, double* alpha_clamped) {
    if (!ctx || !x || !p || !alpha_out || !(alpha_in > 0)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        upload_x(ctx, ctx->qp, p, n);
        int z = 0;
        double cl = 0;
        *alpha_out = ccd_stepsize(c, ctx->qx.p, ctx->qp.p, alpha_in, &cl, &z);
        if (alpha_clamped) *alpha_clamped = cl;
        if (z) { c.lastError = "zero time of impact"; return CIPC_E_ZERO_TOI; }
        return CIPC_OK;
    });
}

int cipc_linear_solve(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, const double* b, double* sol,
                      double tol, int32_t* iters, double* relres) {
    if (!ctx || !x || !b || !sol || n_cs < 0 || (n_cs && !cs)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        Ctx& c = ctx->c;
        size_t n = 3 * (size_t)c.nV;
        upload_x(ctx, ctx->qx, x, n);
        upload_x(ctx, ctx->qb, b, n);
        upload_cs(ctx, cs, n_cs);
        elastic_hessian(c, ctx->qx.p);
        barrier_hessian(c, ctx->qx.p, ctx->qC.p, n_cs, kappa);
        double* sd = ctx->qsol.ensure(n);
        int it = 0;
        double rr = 0;
        c.lpStepStart = true;   // a standalone solve evaluates the learned network afresh
        c.cg.stepStart = true;
        c.lpOffThisStep = false;
        c.fp64ThisStep = false;
        c.solveX = ctx->qx.p;
        int rc = pcg_solve(c, ctx->qb.p, sd, tol > 0 ? tol : c.P.pcgTol, c.P.pcgMaxIter, &it, &rr);
        c.solveX = nullptr;
        CIPC_HIP(hipMemcpyAsync(sol, sd, n * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        if (iters) *iters = it;
        if (relres) *relres = rr;
        if (rc == 1) { c.lastError = "matrix not positive definite"; return CIPC_E_NO_CONVERGENCE; }
        if (rc == 2) { c.lastError = "PCG iteration limit"; return CIPC_E_NO_CONVERGENCE; }
        if (rc == 3) { c.lastError = "system too large for the direct solve"; return CIPC_E_UNSUPPORTED; }
        return CIPC_OK;
    });
}

int cipc_set_precision(cipc_ctx* ctx, int32_t mode) {
    if (!ctx) return CIPC_E_INVALID;
    if (mode < 0 || mode > 2) {
        ctx->c.lastError = "precision mode must be 0 (fp64), 1 (mixed) or 2 (fp32)";
        return CIPC_E_INVALID;
    }
    ctx->c.P.precision = mode;
    return CIPC_OK;
}
int cipc_get_solver_info(const cipc_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    out[0] = ctx->c.lastRefinements;
    out[1] = (int32_t)ctx->c.pcgGraphCaptures;
    out[2] = (int32_t)ctx->c.pcgGraphLaunches;
    out[3] = (int32_t)ctx->c.directSolves;
    return CIPC_OK;
}

int cipc_set_coarse(cipc_ctx* ctx, int32_t every, int32_t max_unknowns, int32_t flags) {
    if (!ctx || every < 0 || max_unknowns < 0 || (flags & ~1)) return CIPC_E_INVALID;
    Ctx& c = ctx->c;
    c.coarseOn = every > 0;
    c.cg.every = every > 0 ? every : 1;
    c.cg.maxDof = max_unknowns;
    c.cg.bf16 = (flags & 1) != 0;
    c.cg.ready = false;
    return CIPC_OK;
}
int cipc_get_coarse_info(const cipc_ctx* ctx, double* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    const CoarseSpace& g = ctx->c.cg;
    out[0] = g.built ? g.nAgg : -1;
    out[1] = g.nc;
    out[2] = (double)g.setups;
    out[3] = g.lastSetupMs;
    out[4] = g.ready ? 1 : 0;
    out[5] = (double)g.failures;
    out[6] = (double)ctx->c.fp64Retries;
    return CIPC_OK;
}

int cipc_set_linear_solver(cipc_ctx* ctx, int32_t mode) {
    if (!ctx) return CIPC_E_INVALID;
    Ctx& c = ctx->c;
    if (mode < 0 || mode > 1) {
        c.lastError = "linear solver must be 0 (PCG, direct fallback) or 1 (direct)";
        return CIPC_E_INVALID;
    }
    if (mode == 1 && 3l * c.nV > kDirectMaxDof) {
        c.lastError = "direct solve limited to " + std::to_string(kDirectMaxDof) + " unknowns";
        return CIPC_E_UNSUPPORTED;
    }
    c.P.linSolver = mode;
    return CIPC_OK;
}

int cipc_set_output_folder(cipc_ctx* ctx, const char* folder) {
    if (!ctx) return CIPC_E_INVALID;
    ctx->c.outputFolder = folder ? folder : "";
    return CIPC_OK;
}

int cipc_set_static_solve(cipc_ctx* ctx, int32_t enable) {
    if (!ctx) return CIPC_E_INVALID;
    ctx->c.P.staticSolve = enable ? 1 : 0;
    return CIPC_OK;
}

int cipc_set_graphs(cipc_ctx* ctx, int32_t enable) {
    if (!ctx) return CIPC_E_INVALID;
    ctx->c.useGraphs = enable != 0;
    if (!enable) ctx->c.pcgGraph.reset();
    return CIPC_OK;
}

int32_t cipc_precond_num_params(void) { return lp_num_params(); }
int cipc_precond_init(cipc_ctx* ctx, uint64_t seed) {
    if (!ctx) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        lp_init_weights(ctx->c, seed);
        return CIPC_OK;
    });
}
int cipc_precond_set_weights(cipc_ctx* ctx, const float* w, int32_t n) {
    if (!ctx || !w || n != lp_num_params()) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        set_weights(ctx->c, w, n);
        return CIPC_OK;
    });
}
int cipc_precond_get_weights(cipc_ctx* ctx, float* w, int32_t n) {
    if (!ctx || !w || n != lp_num_params()) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        get_weights(ctx->c, w, n);
        return CIPC_OK;
    });
}
int cipc_set_preconditioner(cipc_ctx* ctx, int32_t mode, int32_t budget) {
    if (!ctx || mode < 0 || mode > 1 || budget < 0) return CIPC_E_INVALID;
    if (mode == 1 && (!ctx->c.lp || !ctx->c.lp->weightsReady)) {
        ctx->c.lastError = "learned preconditioner has no weights (cipc_precond_init / cipc_precond_set_weights)";
        return CIPC_E_INVALID;
    }
    ctx->c.precond = mode;
    if (budget > 0) ctx->c.lpBudget = budget;
    return CIPC_OK;
}
int cipc_autotune(cipc_ctx* ctx, int32_t what) {
    if (!ctx || what < 0 || what > 3) return CIPC_E_INVALID;
    autotune_start(ctx->c, what);
    return CIPC_OK;
}
int cipc_get_tuning(const cipc_ctx* ctx, cipc_tuning* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    const Ctx& c = ctx->c;
    *out = cipc_tuning{};
    out->spmv_grid = c.spmvGrid; out->spmv_lanes = c.spmvLanes; out->spmv_unroll = c.spmvUnroll;
    out->factor_grid = c.lpGrid; out->factor_lanes = c.lpLanes;
    out->tuning_pcg = c.tune.pcg ? 1 : 0; out->tuning_cell = c.tune.cell ? 1 : 0;
    out->cell_mult = c.cellMult;
    return CIPC_OK;
}
int cipc_set_tuning(cipc_ctx* ctx, const cipc_tuning* t) {
    if (!ctx || !t) return CIPC_E_INVALID;
    auto okGrid = [](int g) { return g >= 8 && g % 8 == 0 && g <= 65536; };
    const int fl = t->factor_lanes, sl = t->spmv_lanes;
    if (!okGrid(t->spmv_grid) || !okGrid(t->factor_grid) || !(sl == 4 || sl == 8 || sl == 16) ||
        !(t->spmv_unroll == 1 || t->spmv_unroll == 2) || !(fl == 2 || fl == 4 || fl == 8 || fl == 16) ||
        !(t->cell_mult >= 0.125 && t->cell_mult <= 16))
        return CIPC_E_INVALID;
    Ctx& c = ctx->c;
    c.spmvGrid = t->spmv_grid; c.spmvLanes = sl; c.spmvUnroll = t->spmv_unroll;
    c.lpGrid = t->factor_grid; c.lpLanes = fl; c.cellMult = t->cell_mult;
    c.tune.pcg = c.tune.cell = c.tune.enabled = false;
    return CIPC_OK;
}
int cipc_set_precond_reuse(cipc_ctx* ctx, int32_t every) {
    if (!ctx || every < 1) return CIPC_E_INVALID;
    ctx->c.lpEvery = every;
    return CIPC_OK;
}
int cipc_set_warm_start(cipc_ctx* ctx, int32_t enable) {
    if (!ctx) return CIPC_E_INVALID;
    ctx->c.warmStart = enable != 0;
    return CIPC_OK;
}
static int precond_common(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa) {
    Ctx& c = ctx->c;
    if (!c.lp || !c.lp->weightsReady) {
        c.lastError = "learned preconditioner has no weights";
        return CIPC_E_INVALID;
    }
    size_t n = 3 * (size_t)c.nV;
    upload_x(ctx, ctx->qx, x, n);
    upload_cs(ctx, cs, n_cs);
    elastic_hessian(c, ctx->qx.p);
    barrier_hessian(c, ctx->qx.p, ctx->qC.p, n_cs, kappa);
    return CIPC_OK;
}
int cipc_precond_loss_grad(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t n_probes, uint64_t seed,
                           double* loss, float* grad) {
    if (!ctx || !x || !loss || n_cs < 0 || (n_cs && !cs) || n_probes < 1) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        int rc = precond_common(ctx, x, cs, n_cs, kappa);
        if (rc != CIPC_OK) return rc;
        *loss = lp_loss_grad(ctx->c, n_probes, seed, grad);
        return CIPC_OK;
    });
}
int cipc_precond_train_step(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t n_probes, uint64_t seed,
                            double lr, double* loss) {
    if (!ctx || !x || n_cs < 0 || (n_cs && !cs) || n_probes < 1 || !(lr > 0)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        int rc = precond_common(ctx, x, cs, n_cs, kappa);
        if (rc != CIPC_OK) return rc;
        const double l = lp_train_step(ctx->c, n_probes, seed, lr);
        if (loss) *loss = l;
        return CIPC_OK;
    });
}
int cipc_precond_factor(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t use_mfma, float* g,
                        int32_t* nnzb) {
    if (!ctx || !x || !nnzb || n_cs < 0 || (n_cs && !cs)) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        *nnzb = ctx->c.Ael.nnzb;
        if (!g) return CIPC_OK;
        int rc = precond_common(ctx, x, cs, n_cs, kappa);
        if (rc != CIPC_OK) return rc;
        lp_factor(ctx->c, use_mfma != 0, g, 9 * ctx->c.Ael.nnzb);
        return CIPC_OK;
    });
}
int cipc_precond_get_adam(cipc_ctx* ctx, float* m, float* v, int32_t* t, int32_t n) {
    if (!ctx || !m || !v || !t || n != lp_num_params()) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        get_adam(ctx->c, m, v, t, n);
        return CIPC_OK;
    });
}
int cipc_precond_set_adam(cipc_ctx* ctx, const float* m, const float* v, int32_t t, int32_t n) {
    if (!ctx || !m || !v || t < 0 || n != lp_num_params()) return CIPC_E_INVALID;
    return guarded(ctx, [&]() {
        set_adam(ctx->c, m, v, t, n);
        return CIPC_OK;
    });
}
int cipc_precond_info(const cipc_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    out[0] = ctx->c.lastPrecondFallback;
    out[1] = (int32_t)ctx->c.precondSolves;
    out[2] = (int32_t)ctx->c.precondFallbacks;
    out[3] = ctx->c.lastLearnedRc;
    out[4] = ctx->c.lastLearnedIters;
    out[5] = ctx->c.lastFallbackContinued;
    out[6] = ctx->c.lastJacobiIters;
    return CIPC_OK;
}

int cipc_set_profiling(cipc_ctx* ctx, int32_t enable) {
    if (!ctx) return CIPC_E_INVALID;
    ctx->c.profiling = enable ? 1 : 0;
    ctx->c.profPrecision = ctx->c.P.precision;
    ctx->c.profLaunches = 0;
    ctx->c.profMs = 0;
    ctx->c.profBytes = 0;
    ctx->c.profFlops = 0;
    return CIPC_OK;
}
int cipc_get_profile(const cipc_ctx* ctx, double* out) {
    if (!ctx || !out) return CIPC_E_INVALID;
    const Ctx& c = ctx->c;
    double nrow = c.Ael.nrow, nnzb = c.Ael.nnzb;
    out[0] = (double)c.profLaunches;
    out[1] = c.profMs;
    out[2] = nrow;
    out[3] = nnzb;
    const double L = c.profLaunches > 0 ? (double)c.profLaunches : 1.0;
    out[4] = c.profBytes / L;
    out[5] = c.profFlops / L;
    return CIPC_OK;
}

}  // extern "C"
