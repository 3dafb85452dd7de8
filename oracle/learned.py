"""CPU restatement of the learned sparse preconditioner (test infrastructure).

Only tests, smoke() and the benchmark's CPU leg use this module, as the checker
of the HIP kernels.  It restates, in numpy:

* the top-K row graph, the message weights, the N values and the 8 row
  features of a CSR matrix;
* the network (embedding, 3 message-passing layers, hidden 64, tanh readout
  of g and h) with BF16 rounding at the points where the device stores BF16
  (``bf16=True``) or in plain float64 (``bf16=False``, for finite differences);
* the preconditioner M^-1 = S D^-1 S - Q (N + N^T) Q and the pair loss
  L = sum_p |M^-1 r_p - e_p|^2 / (P |e_p|^2), with its analytic gradient.

The parameter layout matches the device's single FP32 array.  This path has
no counterpart in the reference benchmark (cg.cu has no preconditioner);
its pins are mathematical: finite differences of the float64 form, the dense
matrix form of M^-1, brute-force top-K selection and the Neumann identity.
"""
from __future__ import annotations

import numpy as np

K = 4            # neighbours per row
F = 8            # features
K0 = 32          # embedding input width (features zero-padded)
H = 64           # hidden width
XW = 2 * H       # [H | aggregated H]
L = 3            # message-passing layers
P_MAX = 8        # pairs per system
GMAX = 1.0       # |g|, |h| bound

OFF_W0 = 0
OFF_B0 = OFF_W0 + K0 * H
OFF_W1 = OFF_B0 + H
LAYER = XW * H + H
OFF_R = OFF_W1 + L * LAYER
READOUT = 2 * (H + 1)
N_PARAMS = OFF_R + READOUT


def layer_w(params, l):
    o = OFF_W1 + l * LAYER
    return params[o:o + XW * H].reshape(XW, H)


def layer_b(params, l):
    o = OFF_W1 + l * LAYER + XW * H
    return params[o:o + H]


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even); float32 out."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


# ---------------------------------------------------------------------------
# Graph and features
# ---------------------------------------------------------------------------

class Graph:
    def __init__(self, n, adiag, nbr, wgt, nval, feat):
        self.n = n
        self.adiag = adiag      # float64 [n]
        self.nbr = nbr          # int32 [n, K]
        self.wgt = wgt          # float64 [n, K] (normalised message weights)
        self.nval = nval        # float64 [n, K]
        self.feat = feat        # float64 [n, F]


def graph(m) -> Graph:
    """Top-K off-diagonals per row by (|a| descending, column ascending), padded
    with zero-weight self edges; features as the device computes them."""
    n = m.n
    rs, ci, a = m.rowstr, m.colidx, m.a
    adiag = np.zeros(n)
    for j in range(n):
        cols = ci[rs[j]:rs[j + 1]]
        hit = np.nonzero(cols == j)[0]
        if len(hit):
            adiag[j] = a[rs[j] + hit[-1]]
    ad = np.where(adiag != 0.0, np.abs(adiag), 1.0)
    nbr = np.zeros((n, K), np.int32)
    wgt = np.zeros((n, K))
    nval = np.zeros((n, K))
    feat = np.zeros((n, F))
    for j in range(n):
        cols = ci[rs[j]:rs[j + 1]]
        vals = a[rs[j]:rs[j + 1]]
        off = cols != j
        c, v = cols[off], vals[off]
        av = np.abs(v)
        dk = ad[c]
        feat[j, 0] = av.sum() / ad[j]
        feat[j, 1] = (v * v).sum() / ad[j] ** 2
        feat[j, 2] = (av.max() if len(av) else 0.0) / ad[j]
        feat[j, 3] = v.sum() / ad[j]
        feat[j, 4] = (v * v / (ad[j] * dk)).sum()
        feat[j, 5] = (av / np.sqrt(ad[j] * dk)).sum()
        feat[j, 6] = np.log2(len(cols)) / 8.0
        feat[j, 7] = np.log10(ad[j]) / 3.0
        order = np.lexsort((c, -av))[:K]
        w = np.zeros(K)
        for t in range(K):
            if t < len(order):
                k = c[order[t]]
                nbr[j, t] = k
                w[t] = av[order[t]] / np.sqrt(ad[j] * ad[k])
                if adiag[j] != 0.0 and adiag[k] != 0.0:
                    nval[j, t] = v[order[t]] / (adiag[j] * adiag[k])
            else:
                nbr[j, t] = j
        s = w.sum()
        wgt[j] = w / s if s > 0 else 0.0
    return Graph(n, adiag, nbr, wgt, nval, feat)


def sym_matrix(g: Graph):
    """C = N + N^T as a dense matrix (small n only)."""
    C = np.zeros((g.n, g.n))
    for j in range(g.n):
        for t in range(K):
            C[j, g.nbr[j, t]] += g.nval[j, t]
    return C + C.T


def sym_apply(g: Graph, v):
    """C v for v of shape [n] or [P, n] without forming C."""
    v = np.atleast_2d(v)
    out = np.zeros_like(v, dtype=np.float64)
    rows = np.repeat(np.arange(g.n), K)
    cols = g.nbr.reshape(-1)
    val = g.nval.reshape(-1)
    for p in range(v.shape[0]):
        np.add.at(out[p], rows, val * v[p, cols])   # N v
        np.add.at(out[p], cols, val * v[p, rows])   # N^T v
    return out


def precond_dense(g: Graph, gh):
    """M^-1 = S D^-1 S - Q C Q as a dense matrix (small n only)."""
    gg, hh = gh[: g.n], gh[g.n:]
    s2 = np.exp(2 * gg) / g.adiag
    q = np.exp(gg + hh)
    return np.diag(s2) - q[:, None] * sym_matrix(g) * q[None, :]


# ---------------------------------------------------------------------------
# Pair loss and its gradient with respect to (g, h)
# ---------------------------------------------------------------------------

def pair_loss(g: Graph, gh, r, e):
    """Loss of pairs r, e ([P, n]) and dL/d(g, h) ([2n])."""
    n = g.n
    gg, hh = np.asarray(gh[:n], np.float64), np.asarray(gh[n:], np.float64)
    P = r.shape[0]
    omega = 1.0 / (P * np.einsum("pj,pj->p", e, e))
    s2 = np.exp(2 * gg) / g.adiag
    q = np.exp(gg + hh)
    w = sym_apply(g, q[None, :] * r)
    y = s2[None, :] * r - q[None, :] * w
    diff = y - e
    loss = float((omega[:, None] * diff * diff).sum())
    d = 2 * omega[:, None] * diff
    phi = sym_apply(g, q[None, :] * d)
    dq = (-d * w - r * phi).sum(axis=0)
    ds = (d * r).sum(axis=0)
    return loss, np.concatenate([2 * s2 * ds + q * dq, q * dq])


# ---------------------------------------------------------------------------
# Network
# ---------------------------------------------------------------------------

def _round(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, np.float64)


def forward(params, norm, g: Graph, bf16_on=True):
    """Activations and (g, h) of the network; bf16_on rounds where the device
    stores BF16 (inputs, weights, activations)."""
    n = g.n
    p = np.asarray(params, np.float64)
    x0 = np.zeros((n, K0))
    x0[:, :F] = (g.feat.astype(np.float32) - norm[:F]) / norm[F:]
    x0 = _round(x0, bf16_on)
    w0 = _round(p[OFF_W0:OFF_W0 + K0 * H].reshape(K0, H), bf16_on)
    hcur = _round(np.maximum(x0 @ w0 + p[OFF_B0:OFF_B0 + H], 0.0), bf16_on)
    xs = []
    wg = g.wgt.astype(np.float32).astype(np.float64) if bf16_on else g.wgt
    for l in range(L):
        agg = _round(np.einsum("jt,jtf->jf", wg, hcur[g.nbr]), bf16_on)
        x = np.concatenate([hcur, agg], axis=1)
        xs.append(x)
        wl = _round(layer_w(p, l), bf16_on)
        hcur = _round(np.maximum(x @ wl + layer_b(p, l), 0.0), bf16_on)
    h3 = hcur
    ro = p[OFF_R:]
    og = h3 @ ro[:H] + ro[H]
    oh = h3 @ ro[H + 1:2 * H + 1] + ro[2 * H + 1]
    gh = np.concatenate([GMAX * np.tanh(og), GMAX * np.tanh(oh)])
    return dict(x0=x0, xs=xs, h3=h3, og=og, oh=oh, gh=gh)


def backward(params, g: Graph, act, dgh, bf16_on=True):
    """Gradient of the parameters from dL/d(g, h)."""
    n = g.n
    p = np.asarray(params, np.float64)
    grad = np.zeros(N_PARAMS)
    ro = p[OFF_R:]
    dtg = dgh[:n] * GMAX * (1 - np.tanh(act["og"]) ** 2)
    dth = dgh[n:] * GMAX * (1 - np.tanh(act["oh"]) ** 2)
    h3 = act["h3"]
    grad[OFF_R:OFF_R + H] = h3.T @ dtg
    grad[OFF_R + H] = dtg.sum()
    grad[OFF_R + H + 1:OFF_R + 2 * H + 1] = h3.T @ dth
    grad[OFF_R + 2 * H + 1] = dth.sum()
    dz = (dtg[:, None] * ro[None, :H] + dth[:, None] * ro[None, H + 1:2 * H + 1]) * (h3 > 0)
    dz = _round(dz, bf16_on)
    wg = g.wgt.astype(np.float32).astype(np.float64) if bf16_on else g.wgt
    for l in range(L - 1, -1, -1):
        x = act["xs"][l]
        o = OFF_W1 + l * LAYER
        grad[o:o + XW * H] = (x.T @ dz).reshape(-1)
        grad[o + XW * H:o + LAYER] = dz.sum(axis=0)
        wl = _round(layer_w(p, l), bf16_on)
        dx = dz @ wl.T
        dh = dx[:, :H].copy()
        np.add.at(dh, g.nbr.reshape(-1), (wg.reshape(-1)[:, None] *
                                           np.repeat(dx[:, H:], K, axis=0)))
        dz = _round(dh * (x[:, :H] > 0), bf16_on)
    x0 = act["x0"]
    grad[OFF_W0:OFF_W0 + K0 * H] = (x0.T @ dz).reshape(-1)
    grad[OFF_B0:OFF_B0 + H] = dz.sum(axis=0)
    return grad


def loss_and_grad(params, norm, g: Graph, r, e, bf16_on=True):
    act = forward(params, norm, g, bf16_on)
    loss, dgh = pair_loss(g, act["gh"], r, e)
    return loss, backward(params, g, act, dgh, bf16_on), act["gh"]


def feature_norm(g: Graph):
    f = g.feat.astype(np.float32).astype(np.float64)
    m = f.mean(axis=0)
    s = np.sqrt(np.maximum((f * f).mean(axis=0) - m * m, 0.0))
    return np.concatenate([m, np.maximum(s, 1e-3)]).astype(np.float32)


def cg_pairs(m, b, xstar, pairs, spmv):
    """The first `pairs` residuals r_k of plain CG from x = 0 and the
    corrections e_k = x* - x_k (float64, [P, n])."""
    x = np.zeros(m.n)
    r = b.copy()
    p = r.copy()
    rho = r @ r
    R, E = [], []
    for k in range(pairs):
        R.append(r.copy())
        E.append(xstar - x)
        if k + 1 == pairs:
            break
        q = spmv(p)
        alpha = rho / (p @ q)
        x = x + alpha * p
        r = r - alpha * q
        rho0, rho = rho, r @ r
        p = r + (rho / rho0) * p
    return np.array(R), np.array(E)
