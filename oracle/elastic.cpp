// Oracle: shell elasticity and inertia summed over the mesh.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Rods (springs, hinges) are added by codim.cpp: stretching energy with the membrane,
// rod bending with the hinge bending.
// Restates FEM/Shell/MEMBRANE.h:7-290 (Neo-Hookean membrane, triangles whose three
// nodes are all Dirichlet are skipped), FEM/Shell/BENDING.h:52-75, 177-208, 441-496
// (hinge bending, KL = false), and the inertia term of FEM/Shell/INC_POTENTIAL.h:
// 132-144, 250-269.  All energies carry the h^2 factor of the incremental potential.
// Parity pins: per-stencil derivatives against torch autograd / finite differences
// (tests/test_oracle_geom.py); assembled sums against numpy re-evaluations.
#include "oracle.h"

namespace orc {

static inline V3 X3(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

static inline bool tri_all_dbc(const Scene& S, int t) {
    const int* v = &S.tri[3 * t];
    return S.isDBC[v[0]] && S.isDBC[v[1]] && S.isDBC[v[2]];
}
static inline bool hinge_all_dbc(const Scene& S, int h) {
    const int* v = &S.hinge[4 * h];
    return S.isDBC[v[0]] && S.isDBC[v[1]] && S.isDBC[v[2]] && S.isDBC[v[3]];
}

double elastic_energy(const Scene& S, const double* x, double* membOut, double* bendOut) {
    const double h2 = shell_h2(S);
    double Em = 0, Eb = 0;
    const bool memb = !fiber_active(S);   // the fibre model replaces the membrane (INC_POTENTIAL.h:81-92)
    for (int t = 0; t < S.nT && memb; ++t) {
        if (tri_all_dbc(S, t)) continue;
        const int* v = &S.tri[3 * t];
        Em += h2 * membrane_energy(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), &S.Binv[3 * t], S.vol[t], S.lam, S.mu);
    }
    if (S.P.bendMult != 0) {
        int nH = (int)S.hinge.size() / 4;
        for (int h = 0; h < nH; ++h) {
            if (hinge_all_dbc(S, h)) continue;
            const int* v = &S.hinge[4 * h];
            double th = dihedral_angle(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), X3(x, v[3]));
            const double* I = &S.hingeInfo[3 * h];
            Eb += h2 * S.kBend * (th - I[0]) * (th - I[0]) * I[1] / I[2];
        }
    }
    codim_energy(S, x, &Em, &Eb);
    if (!S.tet.empty()) Em += tet_energy(S, x);   // volumetric Neo-Hookean with the stretching terms
    if (membOut) *membOut = Em;
    if (bendOut) *bendOut = Eb;
    return Em + Eb;
}

void elastic_grad(const Scene& S, const double* x, double* g) {
    const double h2 = shell_h2(S);
    const bool memb = !fiber_active(S);
    for (int t = 0; t < S.nT && memb; ++t) {
        if (tri_all_dbc(S, t)) continue;
        const int* v = &S.tri[3 * t];
        double gl[9];
        membrane_grad(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), &S.Binv[3 * t], S.vol[t], S.lam, S.mu, gl);
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) g[3 * v[a] + k] += h2 * gl[3 * a + k];
    }
    if (S.P.bendMult != 0) {
        int nH = (int)S.hinge.size() / 4;
        for (int h = 0; h < nH; ++h) {
            if (hinge_all_dbc(S, h)) continue;
            const int* v = &S.hinge[4 * h];
            V3 p[4] = {X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), X3(x, v[3])};
            double th = dihedral_angle(p[0], p[1], p[2], p[3]);
            double gl[12];
            dihedral_angle_grad(p[0], p[1], p[2], p[3], gl);
            const double* I = &S.hingeInfo[3 * h];
            double s = h2 * S.kBend * 2 * (th - I[0]) * I[1] / I[2];
            for (int a = 0; a < 4; ++a)
                for (int k = 0; k < 3; ++k) g[3 * v[a] + k] += s * gl[3 * a + k];
        }
    }
    codim_grad(S, x, g);
    if (!S.tet.empty()) tet_grad(S, x, g);
}

// Local blocks are written straight into their final slots of `out` (the reference's
// per-element triplet offsets, MEMBRANE.h:170-180, BENDING.h:443-455).
void elastic_hess(const Scene& S, const double* x, std::vector<Block>& out) {
    const double h2 = shell_h2(S);
    std::vector<long> offT(S.nT + 1, 0);
    const bool memb = !fiber_active(S);
    for (int t = 0; t < S.nT; ++t) offT[t + 1] = offT[t] + ((tri_all_dbc(S, t) || !memb) ? 0 : 9);
    const long base = (long)out.size();
    out.resize(base + offT[S.nT]);
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (int t = 0; t < S.nT; ++t) {
        if (tri_all_dbc(S, t) || !memb) continue;
        const int* v = &S.tri[3 * t];
        double H[81];
        membrane_hess(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), &S.Binv[3 * t], S.vol[t], S.lam, S.mu, H);
        for (int i = 0; i < 81; ++i) H[i] *= h2;
        if (S.P.projectPD) make_pd(9, H);
        Block* dst = &out[base + offT[t]];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                Block& B = dst[3 * a + b];
                B.r = v[a]; B.c = v[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) B.b[3 * k + l] = H[(3 * a + k) * 9 + 3 * b + l];
            }
    }
    codim_hess(S, x, out);
    if (!S.tet.empty()) tet_hess(S, x, out);
    if (S.P.bendMult == 0) return;
    int nH = (int)S.hinge.size() / 4;
    std::vector<long> offH(nH + 1, 0);
    for (int h = 0; h < nH; ++h) offH[h + 1] = offH[h] + (hinge_all_dbc(S, h) ? 0 : 16);
    const long baseH = (long)out.size();
    out.resize(baseH + offH[nH]);
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (int h = 0; h < nH; ++h) {
        if (hinge_all_dbc(S, h)) continue;
        const int* v = &S.hinge[4 * h];
        V3 p[4] = {X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), X3(x, v[3])};
        double th = dihedral_angle(p[0], p[1], p[2], p[3]);
        double gl[12], H[144];
        dihedral_angle_grad(p[0], p[1], p[2], p[3], gl);
        dihedral_angle_hess(p[0], p[1], p[2], p[3], H);
        const double* I = &S.hingeInfo[3 * h];
        double s1 = h2 * S.kBend * 2 * (th - I[0]) * I[1] / I[2];
        double s2 = h2 * S.kBend * 2 * I[1] / I[2];
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) H[i * 12 + j] = s1 * H[i * 12 + j] + s2 * gl[i] * gl[j];
        if (S.P.projectPD) make_pd(12, H);
        Block* dst = &out[baseH + offH[h]];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                Block& B = dst[4 * a + b];
                B.r = v[a]; B.c = v[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) B.b[3 * k + l] = H[(3 * a + k) * 12 + 3 * b + l];
            }
    }
}

double shell_h2(const Scene& S) { return S.P.staticSolve ? 1.0 : S.P.dt * S.P.dt; }

// inertia 1/2 (x - xt)^T M (x - xt), or in a static solve the body force's -x . (M g)
// (INC_POTENTIAL.h:118-144)
double inertia_energy(const Scene& S, const double* x, const double* xt) {
    double E = 0;
    if (S.P.staticSolve) {
        for (int i = 0; i < S.nV; ++i)
            if (S.mass[i] > 0)
                for (int k = 0; k < 3; ++k) E -= x[3 * i + k] * (S.mass[i] * S.P.grav[k]);
        return E;
    }
    for (int i = 0; i < S.nV; ++i)
        for (int k = 0; k < 3; ++k) {
            double d = x[3 * i + k] - xt[3 * i + k];
            E += 0.5 * S.mass[i] * d * d;
        }
    return E;
}

}  // namespace orc
