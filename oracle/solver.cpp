// Oracle: sparse assembly and linear solves.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// assemble() restates CSR_MATRIX::Construct_From_Triplet (Math/CSR_MATRIX.h:49-56,
// duplicate summation), the mass-matrix add (FEM/Shell/INC_POTENTIAL.h:383-386) and
// Project_DBC (Math/CSR_MATRIX.h:130-141: every stored entry in a Dirichlet row or
// column becomes 0, stored diagonal entries become 1).
//
// solve_direct() stands in for Solve_Direct (Math/DIRECT_SOLVER.h:14-66, CHOLMOD):
// reverse Cuthill-McKee ordering of the free degrees of freedom followed by an
// envelope Cholesky factorisation; a non-positive pivot reports "not SPD" exactly
// where CHOLMOD reports CHOLMOD_NOT_POSDEF.  solve_pcg() is the block-Jacobi
// preconditioned conjugate gradient the GPU path uses, with the same stopping rule.
// Parity pins: ||Ax-b||/||b|| <= 1e-12 and agreement with scipy.sparse.linalg.spsolve
// on committed fixtures (tests/test_oracle_solver.py).
#include "oracle.h"
#include <algorithm>
#include <queue>
#include <numeric>

namespace orc {

void assemble(const Scene& S, const std::vector<Block>& blocks, CSR& A) {
    const int n = 3 * S.nV;
    struct T { int r, c; double v; };
    std::vector<T> trip;
    trip.reserve(blocks.size() * 9 + n);
    for (const Block& B : blocks)
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) trip.push_back({3 * B.r + k, 3 * B.c + l, B.b[3 * k + l]});
    for (int i = 0; i < S.nV; ++i)
        for (int k = 0; k < 3; ++k) trip.push_back({3 * i + k, 3 * i + k, S.mass[i]});
    std::stable_sort(trip.begin(), trip.end(), [](const T& a, const T& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
    A.n = n;
    A.rp.assign(n + 1, 0);
    A.ci.clear(); A.val.clear();
    for (size_t i = 0; i < trip.size();) {
        size_t j = i;
        double s = 0;
        while (j < trip.size() && trip[j].r == trip[i].r && trip[j].c == trip[i].c) s += trip[j++].v;
        A.ci.push_back(trip[i].c); A.val.push_back(s);
        A.rp[trip[i].r + 1]++;
        i = j;
    }
    for (int r = 0; r < n; ++r) A.rp[r + 1] += A.rp[r];
    for (int r = 0; r < n; ++r)
        for (int k = A.rp[r]; k < A.rp[r + 1]; ++k) {
            int c = A.ci[k];
            if (S.isDBC[r / 3] || S.isDBC[c / 3]) A.val[k] = (r == c) ? 1.0 : 0.0;
        }
}

int solve_direct(const CSR& A, const double* b, double* x) {
    const int n = A.n;
    // identity rows (Dirichlet) are solved directly; the rest is factorised
    std::vector<char> fixed(n, 0);
    for (int r = 0; r < n; ++r) {
        bool id = true;
        for (int k = A.rp[r]; k < A.rp[r + 1]; ++k)
            if ((A.ci[k] == r && A.val[k] != 1.0) || (A.ci[k] != r && A.val[k] != 0.0)) { id = false; break; }
        fixed[r] = id;
    }
    std::vector<int> fidx;
    std::vector<int> map(n, -1);
    for (int r = 0; r < n; ++r) if (!fixed[r]) { map[r] = (int)fidx.size(); fidx.push_back(r); }
    const int m = (int)fidx.size();
    for (int r = 0; r < n; ++r) if (fixed[r]) x[r] = b[r];
    if (m == 0) return 0;
    // adjacency on free dofs
    std::vector<std::vector<int>> adj(m);
    for (int i = 0; i < m; ++i) {
        int r = fidx[i];
        for (int k = A.rp[r]; k < A.rp[r + 1]; ++k) {
            int c = map[A.ci[k]];
            if (c >= 0 && c != i && A.val[k] != 0.0) adj[i].push_back(c);
        }
    }
    // reverse Cuthill-McKee
    std::vector<int> order; order.reserve(m);
    std::vector<char> seen(m, 0);
    std::vector<int> deg(m);
    for (int i = 0; i < m; ++i) deg[i] = (int)adj[i].size();
    for (int start = 0; start < m; ++start) {
        if (seen[start]) continue;
        // pseudo-peripheral start: minimum degree node of this component
        std::vector<int> comp; std::queue<int> q; q.push(start); seen[start] = 1;
        while (!q.empty()) { int u = q.front(); q.pop(); comp.push_back(u); for (int w : adj[u]) if (!seen[w]) { seen[w] = 1; q.push(w); } }
        for (int u : comp) seen[u] = 0;
        int s0 = comp[0];
        for (int u : comp) if (deg[u] < deg[s0]) s0 = u;
        q.push(s0); seen[s0] = 1;
        while (!q.empty()) {
            int u = q.front(); q.pop(); order.push_back(u);
            std::vector<int> nb;
            for (int w : adj[u]) if (!seen[w]) { seen[w] = 1; nb.push_back(w); }
            std::sort(nb.begin(), nb.end(), [&](int a, int c) { return deg[a] < deg[c] || (deg[a] == deg[c] && a < c); });
            for (int w : nb) q.push(w);
        }
    }
    std::reverse(order.begin(), order.end());
    std::vector<int> pos(m);
    for (int i = 0; i < m; ++i) pos[order[i]] = i;
    // envelope: first column of each row (lower triangle)
    std::vector<int> first(m);
    for (int i = 0; i < m; ++i) first[i] = i;
    for (int i = 0; i < m; ++i) {
        int pi = pos[i];
        for (int w : adj[i]) { int pw = pos[w]; if (pw < pi) first[pi] = std::min(first[pi], pw); }
    }
    std::vector<long long> off(m + 1, 0);
    for (int i = 0; i < m; ++i) off[i + 1] = off[i] + (i - first[i] + 1);
    std::vector<double> L(off[m], 0.0);
    for (int i = 0; i < m; ++i) {
        int r = fidx[i]; int pi = pos[i];
        for (int k = A.rp[r]; k < A.rp[r + 1]; ++k) {
            int c = map[A.ci[k]];
            if (c < 0) continue;
            int pc = pos[c];
            if (pc <= pi) L[off[pi] + (pc - first[pi])] += A.val[k];
        }
    }
    // envelope Cholesky, row by row
    for (int i = 0; i < m; ++i) {
        double* Li = &L[off[i]];
        int fi = first[i];
        for (int j = fi; j < i; ++j) {
            double* Lj = &L[off[j]];
            int fj = first[j];
            int k0 = std::max(fi, fj);
            double s = Li[j - fi];
            for (int k = k0; k < j; ++k) s -= Li[k - fi] * Lj[k - fj];
            Li[j - fi] = s / Lj[j - fj];
        }
        double s = Li[i - fi];
        for (int k = fi; k < i; ++k) s -= Li[k - fi] * Li[k - fi];
        if (!(s > 0)) return 1;
        Li[i - fi] = std::sqrt(s);
    }
    // solve L L^T y = b (permuted)
    std::vector<double> y(m);
    for (int i = 0; i < m; ++i) y[pos[i]] = b[fidx[i]];
    for (int i = 0; i < m; ++i) {
        double s = y[i];
        const double* Li = &L[off[i]];
        for (int k = first[i]; k < i; ++k) s -= Li[k - first[i]] * y[k];
        y[i] = s / Li[i - first[i]];
    }
    for (int i = m - 1; i >= 0; --i) {
        const double* Li = &L[off[i]];
        y[i] /= Li[i - first[i]];
        double yi = y[i];
        for (int k = first[i]; k < i; ++k) y[k] -= Li[k - first[i]] * yi;
    }
    for (int i = 0; i < m; ++i) x[fidx[i]] = y[pos[i]];
    return 0;
}

static void spmv(const CSR& A, const double* x, double* y) {
    for (int r = 0; r < A.n; ++r) {
        double s = 0;
        for (int k = A.rp[r]; k < A.rp[r + 1]; ++k) s += A.val[k] * x[A.ci[k]];
        y[r] = s;
    }
}

// 3x3 symmetric block inverse via the adjugate
static void inv3(const double* a, double* r) {
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double id = 1.0 / det;
    r[0] = c00 * id; r[1] = (a[2] * a[7] - a[1] * a[8]) * id; r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r[3] = c01 * id; r[4] = (a[0] * a[8] - a[2] * a[6]) * id; r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r[6] = c02 * id; r[7] = (a[1] * a[6] - a[0] * a[7]) * id; r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

int solve_pcg(const CSR& A, const double* b, double* x, double tol, int maxIter, int* iters) {
    const int n = A.n, nb = n / 3;
    std::vector<double> Dinv(9 * nb, 0.0);
    for (int i = 0; i < nb; ++i) {
        double D[9] = {0};
        for (int k = 0; k < 3; ++k) {
            int r = 3 * i + k;
            for (int q = A.rp[r]; q < A.rp[r + 1]; ++q) {
                int c = A.ci[q];
                if (c / 3 == i) D[3 * k + c % 3] = A.val[q];
            }
        }
        inv3(D, &Dinv[9 * i]);
    }
    std::vector<double> r(b, b + n), z(n), p(n), Ap(n);
    for (int i = 0; i < n; ++i) x[i] = 0;
    double bn = 0;
    for (int i = 0; i < n; ++i) bn += b[i] * b[i];
    bn = std::sqrt(bn);
    if (bn == 0) { if (iters) *iters = 0; return 0; }
    auto prec = [&](const std::vector<double>& in, std::vector<double>& out) {
        for (int i = 0; i < nb; ++i)
            for (int k = 0; k < 3; ++k)
                out[3 * i + k] = Dinv[9 * i + 3 * k] * in[3 * i] + Dinv[9 * i + 3 * k + 1] * in[3 * i + 1] + Dinv[9 * i + 3 * k + 2] * in[3 * i + 2];
    };
    prec(r, z);
    p = z;
    double rz = 0;
    for (int i = 0; i < n; ++i) rz += r[i] * z[i];
    int it = 0;
    for (; it < maxIter; ++it) {
        double rn = 0;
        for (int i = 0; i < n; ++i) rn += r[i] * r[i];
        if (std::sqrt(rn) <= tol * bn) break;
        spmv(A, p.data(), Ap.data());
        double pAp = 0;
        for (int i = 0; i < n; ++i) pAp += p[i] * Ap[i];
        if (!(pAp > 0)) { if (iters) *iters = it; return 1; }
        double al = rz / pAp;
        for (int i = 0; i < n; ++i) { x[i] += al * p[i]; r[i] -= al * Ap[i]; }
        prec(r, z);
        double rz2 = 0;
        for (int i = 0; i < n; ++i) rz2 += r[i] * z[i];
        double be = rz2 / rz;
        rz = rz2;
        for (int i = 0; i < n; ++i) p[i] = z[i] + be * p[i];
    }
    if (iters) *iters = it;
    return it >= maxIter ? 2 : 0;
}

}  // namespace orc
