"""CPU reference (torch, float64) of the learned preconditioner's network and loss.

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in codim-ipc_amd/csrc/learn.hip
(forward, Kaporin loss, backward).  Same model as the library:

  S = diag(H)^-1/2, H^ = S H S (H = full assembled Hessian incl. mass, contact, Dirichlet
  projection); graph = the elastic block pattern (all vertex pairs of a triangle or hinge,
  plus the diagonal).
  node features (12): H^_ii (9), m_i mean(S_i^2), log(1 + contact blocks in row i),
  Dirichlet flag;  edge features (9): H^_ij (all of H: elastic and contact) on the
  elastic block pattern.
  h0 = relu(We f + be), e0 = relu(Wx a + bx)
  4 x [ m = relu(Wm [h_i, h_j, e_ij] + bm); h += relu(Wu [h, sum_j m] + bu); e += m ]
  G_ij (j < i) = Wo [h_i, h_j, e_ij] + bo;  G_ii = tril, diag exp(r_0..2), r_3..5 below
  loss = log(mean_p(w_p^T G H^ G^T w_p) / n) - (2/n) sum_i sum_k r_ik
Weights enter rounded to BF16 (round to nearest even), as on the GPU.
"""
import numpy as np
import torch

HID = 64
LAYERS = 4


def layout():
    o, L = 0, {}

    def take(name, n):
        nonlocal o
        L[name] = (o, n)
        o += n

    take("we", 12 * HID); take("be", HID); take("wx", 9 * HID); take("bx", HID)
    for l in range(LAYERS):
        take("wm%d" % l, 3 * HID * HID); take("bm%d" % l, HID)
        take("wu%d" % l, 2 * HID * HID); take("bu%d" % l, HID)
    take("wo", 3 * HID * 9); take("bo", 9); take("wd", HID * 6); take("bd", 6)
    return L, o


def bf16(w):
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def elastic_pattern(nV, tri, hinge):
    pairs = [np.stack([tri[:, a], tri[:, b]], 1) for a in range(3) for b in range(3)]
    pairs += [np.stack([hinge[:, a], hinge[:, b]], 1) for a in range(4) for b in range(4)]
    pairs.append(np.stack([np.arange(nV), np.arange(nV)], 1))
    P = np.concatenate(pairs).astype(np.int64)
    k = np.unique(P[:, 0] * nV + P[:, 1])
    return (k // nV).astype(np.int64), (k % nV).astype(np.int64)


def probes(seed, nprobes, n3):
    """Rademacher vectors of the library (splitmix64 of seed * FNV prime + p * constant)."""
    M = (1 << 64) - 1
    out = []
    for p in range(nprobes):
        sp = (seed * 0x100000001B3 + p * 0x632BE59BD9B4E019) & M
        with np.errstate(over="ignore"):
            i = np.arange(1, n3 + 1, dtype=np.uint64)
            z = np.uint64(sp) + i * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        out.append(np.where((z >> np.uint64(63)) == 1, 1.0, -1.0))
    return np.array(out)


def stencil(c):
    """vertices of one constraint in the library's int4 encoding (IPC.h:144-661)"""
    c0, c1, c2, c3 = (int(v) for v in c)
    if c0 >= 0:
        if c2 >= 0 and c3 >= 0:
            return [c0, c1, c2, c3]
        if c3 >= 0:
            return [c0, c1, -c2 - 1, c3]
        if c2 >= 0:
            return [c0, -c3 - 1, c1, c2]
        return [c0, -c2 - 1, c1, -c3 - 1]
    p = -c0 - 1
    if c3 >= 0:
        return [p, c1, c2, c3]
    if c2 >= 0:
        return [p, c1, c2]
    return [p, c1]


def contact_row_counts(nV, C, dbc=None):
    """blocks per row of the contact Hessian: distinct stencil partners (incl. itself);
    blocks in a Dirichlet row or column are not stored (they are zero after projection)"""
    rows = [set() for _ in range(nV)]
    for c in C:
        v = stencil(c)
        for a in v:
            if dbc is not None and dbc[a]:
                continue
            rows[a].update(b for b in v if dbc is None or not dbc[b])
    return np.array([len(r) for r in rows], dtype=float)


def loss_and_grad(H, mass, dbc, tri, hinge, weights, nprobes, seed, contact_rows=None):
    """H: scipy CSR (3n x 3n) of the assembled Hessian; contact_rows: per vertex number of
    contact blocks in the row (None -> zeros).  Returns (loss, flat gradient float64)."""
    nV = H.shape[0] // 3
    B = H.tobsr(blocksize=(3, 3))
    B.sort_indices()
    row, col = elastic_pattern(nV, tri, hinge)
    # values of H on the elastic pattern (0 where H has no block there)
    lut = {}
    brow = np.repeat(np.arange(nV), np.diff(B.indptr))
    for k, (i, j) in enumerate(zip(brow, B.indices)):
        lut[(int(i), int(j))] = k
    d = H.diagonal()
    s = np.where(d > 0, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 1.0).reshape(nV, 3)
    Hhat_full = torch.tensor(np.array([B.data[k] * s[i][:, None] * s[j][None, :] for k, (i, j) in enumerate(zip(brow, B.indices))]))
    Hrow, Hcol = torch.tensor(brow), torch.tensor(B.indices.astype(np.int64))
    diagk = np.array([lut[(i, i)] for i in range(nV)])
    node = np.zeros((nV, 12))
    node[:, :9] = Hhat_full[diagk].reshape(nV, 9).numpy()
    node[:, 9] = mass * (s ** 2).mean(1)
    node[:, 10] = np.log1p(np.zeros(nV) if contact_rows is None else contact_rows)
    node[:, 11] = dbc.astype(float)
    off = row != col
    er, ec = row[off], col[off]
    ef = np.zeros((len(er), 9))
    for t, (i, j) in enumerate(zip(er, ec)):
        k = lut.get((int(i), int(j)))
        if k is not None:
            ef[t] = (B.data[k] * s[i][:, None] * s[j][None, :]).ravel()

    L, total = layout()
    wq = torch.tensor(bf16(weights).astype(np.float64), requires_grad=True)

    def W(name, k, n):
        o, m = L[name]
        return wq[o:o + m].reshape(k, n)

    def B_(name):
        o, m = L[name]
        return wq[o:o + m]

    er_t, ec_t = torch.tensor(er), torch.tensor(ec)
    f = torch.tensor(node)
    a = torch.tensor(ef)
    h = torch.relu(f @ W("we", 12, HID) + B_("be"))
    e = torch.relu(a @ W("wx", 9, HID) + B_("bx"))
    for l in range(LAYERS):
        m = torch.relu(torch.cat([h[er_t], h[ec_t], e], 1) @ W("wm%d" % l, 3 * HID, HID) + B_("bm%d" % l))
        agg = torch.zeros_like(h).index_add(0, er_t, m)
        h = h + torch.relu(torch.cat([h, agg], 1) @ W("wu%d" % l, 2 * HID, HID) + B_("bu%d" % l))
        e = e + m
    low = torch.tensor(ec < er)
    lr_, lc_ = er_t[low], ec_t[low]
    Gl = (torch.cat([h[lr_], h[lc_], e[low]], 1) @ W("wo", 3 * HID, 9) + B_("bo")).reshape(-1, 3, 3)
    raw = h @ W("wd", HID, 6) + B_("bd")
    Gd = torch.zeros(nV, 3, 3, dtype=torch.float64)
    Gd = Gd.index_put((torch.arange(nV).repeat_interleave(3), torch.tensor([0, 1, 2] * nV), torch.tensor([0, 1, 2] * nV)),
                      torch.exp(raw[:, :3]).reshape(-1))
    Gd = Gd.index_put((torch.arange(nV).repeat_interleave(3), torch.tensor([1, 2, 2] * nV), torch.tensor([0, 0, 1] * nV)),
                      raw[:, 3:].reshape(-1))
    n3 = 3 * nV
    w = torch.tensor(probes(seed, nprobes, n3)).reshape(nprobes, nV, 3)
    tr = 0
    S = torch.tensor(s)
    Hb = torch.tensor(B.data)
    for p in range(nprobes):
        wp = w[p]
        u = torch.einsum("nji,nj->ni", Gd, wp).index_add(0, lc_, torch.einsum("eji,ej->ei", Gl, wp[lr_]))   # G^T w
        x = S * u
        v = S * torch.zeros_like(u).index_add(0, Hrow, torch.einsum("bij,bj->bi", Hb, x[Hcol]))            # S H S u
        tr = tr + (u * v).sum()
    tr = tr / nprobes
    loss = torch.log(tr / n3) - 2.0 / n3 * raw[:, :3].sum()
    loss.backward()
    return float(loss.detach()), wq.grad.numpy().copy()
