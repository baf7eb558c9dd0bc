"""CPU oracle of the learned preconditioner network (test infrastructure).

The network has no counterpart in the reference; this numpy restatement is
the checker of the HIP forward/backward kernels.  It is pinned by
finite-difference gradient checks in float64 (tests/test_mpnn_oracle.py).

Network (per node i of the periodic node graph, hidden width H = 64):

    x_i   = [Dn_i / sD (25), v_i * s (5), 0, 0],  s = sqrt(n_act) / ||v||  (32)
            Dn_i = diag(d)^-1 D_i diag(d)^-1,  d_r = sqrt|D_i,rr| (1 where 0)
    h0    = bf(silu(bf(x) W0 + b0))
    for l = 1..3:
        mean_i = (1/deg_i) sum_j h_j
        grad_i = (1/deg_i) sum_j (h_j - h_i) (x) e_ij       (64 x 3, c-major)
        A_i    = bf([h_i, mean_i, grad_i])                   (320)
        h      = bf(h + silu(A W_l + b_l))
    y_i   = h W_out + b_out                                  (5)
    z     = v + y / s

bf() rounds to bfloat16 (round to nearest even); products accumulate in
float32.  e_ij = (x_j - x_i) / l_i with l_i the RMS length of node i's
edges.  grad's 192 columns are ordered [c * 64 + k] (direction c, channel k).
Loss of one sample (v, t): mean over active nodes of ||y_i - y*_i||^2 with
y* = (t - v) s; n_act counts the nodes with edges (periodic masters).
"""
import numpy as np

H = 64
NIN = 32
NAGG = 5 * H  # [h, mean, grad(3H)]
NOUT = 5


def bf(a):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = a[nan]
    return out


def silu(u):
    # exp(-u) overflows float32 for u < -88.7; the quotient then takes its
    # exact limit (-0), so the overflow is not reported
    with np.errstate(over="ignore"):
        return u / (1.0 + np.exp(-u))


def dsilu(u):
    with np.errstate(over="ignore"):
        s = 1.0 / (1.0 + np.exp(-u))
    return s * (1.0 + u * (1.0 - s))


def init_params(seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    p = {"W0": rng.standard_normal((NIN, H)) * np.sqrt(2.0 / NIN), "b0": np.zeros(H)}
    for l in (1, 2, 3):
        p[f"W{l}"] = rng.standard_normal((NAGG, H)) * (0.5 / np.sqrt(NAGG))
        p[f"b{l}"] = np.zeros(H)
    p["Wout"] = rng.standard_normal((H, NOUT)) * (0.1 / np.sqrt(H))
    p["bout"] = np.zeros(NOUT)
    return {k: v.astype(dtype) for k, v in p.items()}


class Graph:
    """Periodic node graph: CSR rows over all nodes (slaves have no edges)."""

    def __init__(self, ptr, col, evec):
        self.ptr = np.asarray(ptr, dtype=np.int64)
        self.col = np.asarray(col, dtype=np.int64)
        self.evec = np.asarray(evec, dtype=np.float32).reshape(-1, 3)
        self.n = len(self.ptr) - 1
        deg = np.diff(self.ptr)
        self.deg = deg
        self.inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0).astype(np.float32)
        self.row = np.repeat(np.arange(self.n), deg)
        self.active = deg > 0


def periodic_graph(case):
    """Edges (I, J) = (iper(i), iper(j)) of the element adjacency, i != j,
    one per ordered pair, with the geometric edge vector of the original
    pair normalised by the RMS edge length of row I."""
    n = case.nshg
    colm = case.colm.astype(np.int64) - 1
    rowp = case.rowp.astype(np.int64) - 1
    per = case.iper.astype(np.int64) - 1
    deg = np.diff(colm)
    i = np.repeat(np.arange(n), deg)
    j = rowp
    I, J = per[i], per[j]
    keep = I != J
    i, j, I, J = i[keep], j[keep], I[keep], J[keep]
    vec = case.x[j] - case.x[i]
    key = I * n + J
    order = np.lexsort((J, I))
    key, I, J, vec = key[order], I[order], J[order], vec[order]
    first = np.ones(len(key), dtype=bool)
    first[1:] = key[1:] != key[:-1]
    I, J, vec = I[first], J[first], vec[first]
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, I + 1, 1)
    ptr = np.cumsum(ptr)
    l2 = np.zeros(n)
    np.add.at(l2, I, np.sum(vec * vec, axis=1))
    cnt = np.diff(ptr)
    lrms = np.sqrt(l2 / np.maximum(cnt, 1))
    lrms[cnt == 0] = 1.0
    evec = vec / lrms[I][:, None]
    return Graph(ptr, J, evec)


def features(bdiag, v, dscale, active):
    """Node inputs x (n x 32): BDiag with its diagonal scaled to +-1 (25,
    column-major block) over dscale, and the vector scaled to unit RMS over
    the active nodes; zero rows for inactive nodes."""
    n = v.shape[0]
    vn = np.linalg.norm(v)
    s = np.sqrt(float(np.count_nonzero(active))) / vn if vn > 0 else 0.0
    x = np.zeros((n, NIN), dtype=np.float32)
    d = np.sqrt(np.abs(np.einsum("irr->ir", bdiag)))
    d[d == 0] = 1.0
    dn = bdiag / (d[:, :, None] * d[:, None, :])
    x[:, :25] = dn.reshape(n, 25, order="F") / dscale[None, :]
    x[:, 25:30] = v * s
    x[~active] = 0.0
    return x, s


def aggregate(g, h):
    """A = [h, mean_j h_j, (1/deg) sum_j (h_j - h_i) e_ij] in float32."""
    n, f = g.n, h.dtype
    ev, inv = g.evec.astype(f), g.inv.astype(f)
    hj = h[g.col]
    mean = np.zeros((n, H), dtype=f)
    np.add.at(mean, g.row, hj)
    mean *= inv[:, None]
    d = hj - h[g.row]
    grad = np.zeros((n, 3, H), dtype=f)
    for c in range(3):
        np.add.at(grad[:, c, :], g.row, d * ev[:, c:c + 1])
    grad *= inv[:, None, None]
    return np.concatenate([h, mean, grad.reshape(n, 3 * H)], axis=1)


def aggregate_T(g, dA):
    """Adjoint of aggregate: dA (n x 320) -> dh (n x 64)."""
    n, f = g.n, dA.dtype
    ev, inv = g.evec.astype(f), g.inv.astype(f)
    dh = dA[:, :H].copy()
    dmean = dA[:, H:2 * H] * inv[:, None]
    dgrad = dA[:, 2 * H:].reshape(n, 3, H) * inv[:, None, None]
    # mean: dh_j += dmean_i for each edge (i -> j)
    np.add.at(dh, g.col, dmean[g.row])
    # grad: d/dh_j += sum_c dgrad_i[c] e_ij[c];  d/dh_i -= same
    w = np.einsum("ec,ech->eh", ev, dgrad[g.row])
    np.add.at(dh, g.col, w)
    np.add.at(dh, g.row, -w)
    return dh


def forward(p, g, x, rounding=True):
    """Returns (y, cache).  rounding=False gives the float64 model used by
    the finite-difference checks."""
    r = bf if rounding else (lambda a: a)
    f = np.float32 if rounding else np.float64
    cache = {}
    X = r(x.astype(f))
    U0 = X @ r(p["W0"]).astype(f) + p["b0"].astype(f)
    h = r(silu(U0))
    h[~g.active] = 0.0
    cache.update(X=X, U0=U0, h0=h)
    for l in (1, 2, 3):
        A = r(aggregate(g, h).astype(f))
        U = A @ r(p[f"W{l}"]).astype(f) + p[f"b{l}"].astype(f)
        hn = r(h + silu(U))
        hn[~g.active] = 0.0
        cache[f"A{l}"], cache[f"U{l}"], cache[f"h{l}"] = A, U, hn
        h = hn
    y = h @ r(p["Wout"]).astype(f) + p["bout"].astype(f)
    y[~g.active] = 0.0
    return y, cache


def loss_and_grad(p, g, x, ystar, rounding=True):
    """Mean squared error over active nodes and its gradient w.r.t. every
    parameter (straight-through for the bf16 roundings)."""
    r = bf if rounding else (lambda a: a)
    f = np.float32 if rounding else np.float64
    y, c = forward(p, g, x, rounding)
    na = max(1, int(g.active.sum()))
    diff = (y - ystar.astype(f))
    diff[~g.active] = 0.0
    loss = float(np.sum(diff.astype(np.float64) ** 2) / na)
    dy = (2.0 / na) * diff
    gr = {}
    h3 = c["h3"]
    gr["Wout"] = h3.T @ dy
    gr["bout"] = dy.sum(axis=0)
    dh = dy @ r(p["Wout"]).astype(f).T
    dh[~g.active] = 0.0
    for l in (3, 2, 1):
        dU = dh * dsilu(c[f"U{l}"])
        dU[~g.active] = 0.0
        gr[f"W{l}"] = c[f"A{l}"].T @ dU
        gr[f"b{l}"] = dU.sum(axis=0)
        dA = dU @ r(p[f"W{l}"]).astype(f).T
        dh = dh + aggregate_T(g, dA)
        dh[~g.active] = 0.0
    dU0 = dh * dsilu(c["U0"])
    dU0[~g.active] = 0.0
    gr["W0"] = c["X"].T @ dU0
    gr["b0"] = dU0.sum(axis=0)
    return loss, gr, y
