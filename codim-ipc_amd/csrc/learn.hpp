// Learned preconditioner for the Newton PCG: a message-passing network over the block
// graph of the Hessian predicts a block lower-triangular factor G on the Hessian's own
// sparsity pattern, used as a factorised approximate inverse
//     M^-1 = S G^T G S,   S = diag(H)^-1/2,
// so applying it is two sparse products (no triangular solves).  The network sees the
// symmetrically scaled Hessian  H^ = S H S.  It is trained to minimise the Kaporin
// condition number of G H^ G^T,
//     L = log( tr(G H^ G^T) / n ) - (2/n) log det G,
// with tr() estimated by Rademacher probes and log det G exact (G is triangular with
// exponentiated diagonal).  With the decoder at zero, G = I and M^-1 is point Jacobi.
// Node features: H^_ii, m_i mean(S_i^2), log(1 + contact blocks in row i), Dirichlet
// flag; edge features: H^_ij (elastic plus contact part) on the elastic block pattern.
//
// Network (4 message-passing layers, hidden 64):
//   h0 = relu(We f_i + be)            f_i: 12 node features
//   e0 = relu(Wx a_ij + bx)           a_ij: the 9 entries of H^_ij (off-diagonal blocks)
//   layer l:  m_ij = relu(Wm [h_i, h_j, e_ij] + bm)
//             h_i += relu(Wu [h_i, sum_j m_ij] + bu),   e_ij += m_ij
//   G_ij (j < i) = Wo [h_i, h_j, e_ij] + bo
//   G_ii = lower-triangular, diagonal exp(r_0..2), below-diagonal r_3..5,  r = Wd h_i + bd
// Weights are kept in FP32 (master, Adam moments) and used rounded to BF16; products
// accumulate in FP32.
#pragma once
#include "ctx.hpp"

namespace cipc {

constexpr int kLpHidden = 64;
constexpr int kLpLayers = 4;
constexpr int kLpNodeIn = 12;
constexpr int kLpEdgeIn = 9;

// offsets of the parameter tensors in the flat parameter vector (row-major [K][N] + bias)
struct LpLayout {
    int we, be, wx, bx, wm[kLpLayers], bm[kLpLayers], wu[kLpLayers], bu[kLpLayers], wo, bo, wd, bd, total;
    LpLayout();
};

struct LearnedPrecond {
    LpLayout L;
    bool graphReady = false, weightsReady = false, forwardReady = false;
    bool mfma = true;                 // inference forward: message layers on the matrix cores
    int nV = 0, nnzb = 0, nE = 0, nL = 0;
    // static graph of the elastic BSR pattern (directed edges = off-diagonal blocks)
    DevBuf<int> eBlock, eSrc, eDst, erp, eT, lEdge, blkT, diagBlk, lSrcOff, lSrcIdx, lDstOff, lDstIdx, lowSrc, lowDst;
    // compact block CSRs of F = G S (lower triangle incl. diagonal) and of F^T (upper),
    // for the two products of the preconditioner; pos* map lower edges / diagonals to
    // their slots
    int nLow = 0;
    DevBuf<int> gLrp, gLcol, gUrp, gUcol, posLE, posLD, posUE, posUD;
    DevBuf<uint32_t> FL, FU;          // BF16 blocks, five packed words per block, SoA [5][nLow]
    // parameters
    DevBuf<float> W, Wq, gW, mW, vW;
    int adamT = 0;
    // features and activations (FP32)
    DevBuf<double> S;                 // [nV][3] inverse square roots of diag(H)
    DevBuf<float> nodeF, edgeF;       // [nV][12], [nE][9]
    DevBuf<float> h[kLpLayers + 1], e[kLpLayers + 1], m[kLpLayers], u[kLpLayers], agg[kLpLayers];
    DevBuf<float> Gl, raw;            // [nL][9], [nV][6]
    DevBuf<float> Gblk;               // SoA [9][nnzb]: lower blocks, transposes at upper positions, G_ii on the diagonal
    DevBuf<short> wt16;               // one layer's weights packed for the MFMA inference layers (BF16, transposed)
    // training scratch
    DevBuf<float> dh, dh2, de, de2, dtmp, dtmp2, dtmp3, dagg, dY, dGl, draw, wpartial;
    DevBuf<double> probes, pu, pv, lossPart;
    double lastLoss = 0, lastTrace = 0;
};

// graph + features from the current Hessian (Ael + Acon); forward pass fills Gblk
void lp_build_graph(Ctx& c);
void lp_init_weights(Ctx& c, uint64_t seed);
void set_weights(Ctx& c, const float* w, int n);      // host FP32 weights -> device (Adam state reset)
void get_weights(Ctx& c, float* w, int n);
void get_adam(Ctx& c, float* m, float* v, int* t, int n);
void set_adam(Ctx& c, const float* m, const float* v, int t, int n);
int lp_num_params();
// forward on the current Hessian with or without the MFMA message layers; G = SoA [9][nnzb]
void lp_factor(Ctx& c, bool useMfma, float* G, int n);
void lp_forward(Ctx& c);
// one training step on the current Hessian: loss, gradients, Adam update; returns loss
double lp_train_step(Ctx& c, int nProbes, uint64_t seed, double lr);
// loss and flat gradient without updating (gradient checks)
double lp_loss_grad(Ctx& c, int nProbes, uint64_t seed, float* gradOut);
// features of the current Hessian + forward pass (fills Gblk)
void lp_prepare(Ctx& c);
// reuse the last forward pass's factor with the diagonal scaling S of the current
// Hessian (F = G S); false when there is no forward pass to reuse for this pattern
bool lp_refresh(Ctx& c);
// z = S G^T G S r (V = double or float vectors); tmp holds 3 nV values; when part !=
// nullptr the per-workgroup partial sums of r.z are written there (returns their count).
// coarse: z also receives P zc, the two-level correction already solved into c.cg.zc
template <class V>
int lp_apply(Ctx& c, const V* r, V* z, V* tmp, double* part, bool coarse = false);
int lp_apply_partials(const Ctx& c);
// z = F^T y (the second product of lp_apply) with partial sums of r.z
template <class V>
void lp_apply_upper(Ctx& c, const V* y, V* z, const V* r, double* part, const int* done = nullptr, bool coarse = false);
// the factor products' launch arguments (part of a captured PCG chunk's key)
std::vector<uintptr_t> lp_graph_key(const Ctx& c);
// unpack the 9 BF16 entries of a block stored as five words (SoA [5][n]) at position b
template <class V>
__device__ __forceinline__ void lp_unpack9(const uint32_t* __restrict__ val, long n, long b, V* a) {
    uint32_t w[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) w[k] = val[k * n + b];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[2 * k] = (V)__uint_as_float(w[k] << 16);
        a[2 * k + 1] = (V)__uint_as_float(w[k] & 0xFFFF0000u);
    }
    a[8] = (V)__uint_as_float(w[4] << 16);
}

}  // namespace cipc
