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
// Parity pins: agreement with scipy.sparse.linalg.spsolve on contact-stack Hessians
// (tests/test_oracle_contact.py).
#include "oracle.h"
#include <algorithm>
#include <parallel/algorithm>
#include <queue>
#include <numeric>

namespace orc {

// Blocks are summed per (row, col) in their input order (stable sort), the mass is
// added to the diagonal blocks, then the block matrix is expanded to scalar CSR.
void assemble(const Scene& S, const std::vector<Block>& blocks, CSR& A, const std::vector<char>* proj) {
    const std::vector<char>& pm = proj ? *proj : S.isDBC;
    const int n = 3 * S.nV;
    const long nb = (long)blocks.size();
    std::vector<std::pair<uint64_t, long>> key(nb + S.nV);
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (long i = 0; i < nb; ++i) key[i] = {(uint64_t)blocks[i].r * S.nV + blocks[i].c, i};
    for (int v = 0; v < S.nV; ++v) key[nb + v] = {(uint64_t)v * S.nV + v, -1 - (long)v};   // mass entries
    __gnu_parallel::stable_sort(key.begin(), key.end(), [](const std::pair<uint64_t, long>& a, const std::pair<uint64_t, long>& b) {
        return a.first < b.first;
    });
    // unique blocks
    std::vector<long> start;
    start.reserve(key.size() / 4 + 1);
    for (size_t i = 0; i < key.size(); ++i)
        if (i == 0 || key[i].first != key[i - 1].first) start.push_back((long)i);
    const long nu = (long)start.size();
    start.push_back((long)key.size());
    std::vector<double> bval(9 * nu);
    std::vector<int> brow(nu), bcol(nu);
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (long u = 0; u < nu; ++u) {
        double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (long j = start[u]; j < start[u + 1]; ++j) {
            long id = key[j].second;
            if (id >= 0) for (int k = 0; k < 9; ++k) s[k] += blocks[id].b[k];
            else if (!S.P.staticSolve) { double m = S.mass[-1 - id]; s[0] += m; s[4] += m; s[8] += m; }   // no mass in a static solve
        }
        int r = (int)(key[start[u]].first / S.nV), c = (int)(key[start[u]].first % S.nV);
        brow[u] = r; bcol[u] = c;
        bool z = pm[r] || pm[c];
        for (int k = 0; k < 9; ++k) bval[9 * u + k] = z ? ((r == c && k % 4 == 0) ? 1.0 : 0.0) : s[k];
    }
    // expand block rows to scalar rows
    std::vector<long> rowBlock(S.nV + 1, 0);
    for (long u = 0; u < nu; ++u) rowBlock[brow[u] + 1]++;
    for (int v = 0; v < S.nV; ++v) rowBlock[v + 1] += rowBlock[v];
    A.n = n;
    A.rp.assign(n + 1, 0);
    for (int v = 0; v < S.nV; ++v)
        for (int k = 0; k < 3; ++k) A.rp[3 * v + k + 1] = (int)(3 * (rowBlock[v + 1] - rowBlock[v]));
    for (int r = 0; r < n; ++r) A.rp[r + 1] += A.rp[r];
    A.ci.resize(A.rp[n]);
    A.val.resize(A.rp[n]);
#pragma omp parallel for schedule(static) num_threads(S.P.threads)
    for (int v = 0; v < S.nV; ++v)
        for (int k = 0; k < 3; ++k) {
            long o = A.rp[3 * v + k];
            for (long u = rowBlock[v]; u < rowBlock[v + 1]; ++u)
                for (int l = 0; l < 3; ++l) {
                    A.ci[o] = 3 * bcol[u] + l;
                    A.val[o] = bval[9 * u + 3 * k + l];
                    ++o;
                }
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

// 3x3 symmetric block inverse via the adjugate
static void inv3(const double* a, double* r) {
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double id = 1.0 / det;
    r[0] = c00 * id; r[1] = (a[2] * a[7] - a[1] * a[8]) * id; r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r[3] = c01 * id; r[4] = (a[0] * a[8] - a[2] * a[6]) * id; r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r[6] = c02 * id; r[7] = (a[1] * a[6] - a[0] * a[7]) * id; r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// Block-Jacobi PCG with the same stopping rule as the GPU solver; OpenMP parallel
// loops (the CPU baseline runs it on all host cores).
int solve_pcg(const CSR& A, const double* b, double* x, double tol, int maxIter, int* iters, int threads) {
    const int n = A.n, nb = n / 3;
    if (threads < 1) threads = 1;
    std::vector<double> Dinv(9 * (size_t)nb, 0.0);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int i = 0; i < nb; ++i) {
        double D[9] = {0};
        for (int k = 0; k < 3; ++k) {
            int r = 3 * i + k;
            for (int q = A.rp[r]; q < A.rp[r + 1]; ++q) {
                int c = A.ci[q];
                if (c / 3 == i) D[3 * k + c % 3] = A.val[q];
            }
        }
        inv3(D, &Dinv[9 * (size_t)i]);
    }
    std::vector<double> r(b, b + n), z(n), p(n), Ap(n);
    double bn = 0, rz = 0;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : bn, rz)
    for (int i = 0; i < nb; ++i) {
        for (int k = 0; k < 3; ++k) {
            x[3 * i + k] = 0;
            z[3 * i + k] = Dinv[9 * (size_t)i + 3 * k] * r[3 * i] + Dinv[9 * (size_t)i + 3 * k + 1] * r[3 * i + 1] +
                           Dinv[9 * (size_t)i + 3 * k + 2] * r[3 * i + 2];
            p[3 * i + k] = z[3 * i + k];
            bn += b[3 * i + k] * b[3 * i + k];
            rz += r[3 * i + k] * z[3 * i + k];
        }
    }
    bn = std::sqrt(bn);
    if (bn == 0) { if (iters) *iters = 0; return 0; }
    double rn = bn * bn;
    int it = 0;
    for (; it < maxIter; ++it) {
        if (std::sqrt(rn) <= tol * bn) break;
        double pAp = 0;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : pAp)
        for (int row = 0; row < n; ++row) {
            double s = 0;
            for (int k = A.rp[row]; k < A.rp[row + 1]; ++k) s += A.val[k] * p[A.ci[k]];
            Ap[row] = s;
            pAp += p[row] * s;
        }
        if (!(pAp > 0)) { if (iters) *iters = it; return 1; }
        double al = rz / pAp, rz2 = 0, rn2 = 0;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : rz2, rn2)
        for (int i = 0; i < nb; ++i) {
            double rv[3];
            for (int k = 0; k < 3; ++k) {
                x[3 * i + k] += al * p[3 * i + k];
                rv[k] = r[3 * i + k] - al * Ap[3 * i + k];
                r[3 * i + k] = rv[k];
            }
            for (int k = 0; k < 3; ++k) {
                double zk = Dinv[9 * (size_t)i + 3 * k] * rv[0] + Dinv[9 * (size_t)i + 3 * k + 1] * rv[1] + Dinv[9 * (size_t)i + 3 * k + 2] * rv[2];
                z[3 * i + k] = zk;
                rz2 += rv[k] * zk;
                rn2 += rv[k] * rv[k];
            }
        }
        double be = rz2 / rz;
        rz = rz2;
        rn = rn2;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < n; ++i) p[i] = z[i] + be * p[i];
    }
    if (iters) *iters = it;
    return it >= maxIter ? 2 : 0;
}

}  // namespace orc
