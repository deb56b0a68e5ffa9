"""Independent pure-Python scalar restatement -- TEST INFRASTRUCTURE ONLY.

Written separately from nmf_oracle.c, directly from the reference loops in
rust/src/lib.rs, for tiny shapes only.  Python floats are IEEE binary64 and
CPython never contracts a*b+c into an FMA, so on matching summation orders
this module and the C oracle must agree bit for bit.  It writes the golden
vectors under tests/golden/ and cross-checks the C oracle in tests/.

Matrices are lists of rows (row-major).
"""
from __future__ import annotations

import math

DBL_EPSILON = 2.220446049250313e-16


def _zeros(a, b):
    return [[0.0] * b for _ in range(a)]


# ----------------------------------------------------------------- lee ----
def lee_update_h(V, W, H, eps):
    """euclidean_update_h_impl (lib.rs:247-276) via _inner (lib.rs:210-245)."""
    n, r, p = len(W), len(W[0]), len(H[0])
    G = _zeros(r, r)
    for i in range(r):
        for j in range(r):
            s = 0.0
            for u in reversed(range(n)):
                s += W[u][i] * W[u][j]
            G[i][j] = s
    out = _zeros(r, p)
    for j in range(p):
        for i in range(r):
            num = 0.0
            for u in reversed(range(n)):
                num += W[u][i] * V[u][j]
            den = 0.0
            for l in reversed(range(r)):
                den += G[i][l] * H[l][j]
            t = H[i][j] * num
            out[i][j] = (t if t > eps else eps) / (den + eps)
    return out


def lee_update_w(V, W, H, eps):
    """euclidean_update_w_impl (lib.rs:278-322)."""
    n, r, p = len(W), len(W[0]), len(H[0])
    K = _zeros(r, r)
    for j in range(r):
        for i in range(r):
            s = 0.0
            for u in reversed(range(p)):
                s += H[j][u] * H[i][u]
            K[i][j] = s
    out = _zeros(n, r)
    for i in range(n):
        for j in range(r):
            num = 0.0
            for u in reversed(range(p)):
                num += V[i][u] * H[j][u]
            den = 0.0
            for l in reversed(range(r)):
                den += W[i][l] * K[l][j]
            t = W[i][j] * num
            out[i][j] = (eps if t < eps else t) / (den + eps)
    return out


def col_rescale(W):
    """lib.rs:1460-1470."""
    n, r = len(W), len(W[0])
    for j in range(r):
        s = 0.0
        for i in range(n):
            s += W[i][j]
        if s > 0.0:
            for i in range(n):
                W[i][j] = W[i][j] / s


# -------------------------------------------------------------- brunet ----
def brunet_update_h(V, W, H):
    """divergence_update_h_impl (lib.rs:74-144)."""
    n, r, p = len(W), len(W[0]), len(H[0])
    sw = []
    for k in range(r):
        s = 0.0
        for u in range(n):
            s += W[u][k]
        sw.append(s)
    out = _zeros(r, p)
    for j in range(p):
        hc = [H[k][j] for k in range(r)]
        q = []
        for u in range(n):
            wh = 0.0
            for k in range(r):
                wh += W[u][k] * hc[k]
            q.append(V[u][j] / wh if wh != 0.0 else 0.0)
        for i in range(r):
            tmp = 0.0
            for u in range(n):
                tmp += W[u][i] * q[u]
            out[i][j] = hc[i] * tmp / sw[i]
    return out


def brunet_update_w(V, W, H):
    """divergence_update_w_impl (lib.rs:146-204)."""
    n, r, p = len(W), len(W[0]), len(H[0])
    sh = []
    for k in range(r):
        s = 0.0
        for u in range(p):
            s += H[k][u]
        sh.append(s)
    out = _zeros(n, r)
    for i in range(n):
        q = []
        for u in range(p):
            wh = 0.0
            for k in range(r):
                wh += W[i][k] * H[k][u]
            q.append(V[i][u] / wh if wh != 0.0 else 0.0)
        for j in range(r):
            tmp = 0.0
            for u in range(p):
                tmp += H[j][u] * q[u]
            out[i][j] = W[i][j] * tmp / sh[j]
    return out


def _pmax(X, eps):
    for row in X:
        for k in range(len(row)):
            if row[k] < eps:
                row[k] = eps


# ---------------------------------------------------------------- hals ----
def _wtv(V, W):
    # matrixmultiply in the reference; the oracle's documented forward order.
    n, r, p = len(W), len(W[0]), len(V[0])
    A = _zeros(r, p)
    for u in range(p):
        for k in range(r):
            s = 0.0
            for i in range(n):
                s += W[i][k] * V[i][u]
            A[k][u] = s
    return A


def _vht(V, H):
    n, r, p = len(V), len(H), len(V[0])
    B = _zeros(n, r)
    for i in range(n):
        for k in range(r):
            s = 0.0
            for u in range(p):
                s += V[i][u] * H[k][u]
            B[i][k] = s
    return B


def _gram_cols(W):
    """compute_wtw_from_wt (lib.rs:542-560), forward."""
    n, r = len(W), len(W[0])
    G = _zeros(r, r)
    for k in range(r):
        for j in range(r):
            s = 0.0
            for i in range(n):
                s += W[i][k] * W[i][j]
            G[k][j] = s
    return G


def _gram_rows(H):
    """compute_hht (lib.rs:562-579), forward."""
    r, p = len(H), len(H[0])
    K = _zeros(r, r)
    for k in range(r):
        for j in range(r):
            s = 0.0
            for u in range(p):
                s += H[k][u] * H[j][u]
            K[k][j] = s
    return K


def _rmax(a, b):
    """Rust f64::max: NaN-ignoring."""
    if a != a:
        return b
    if b != b:
        return a
    return a if a > b else b


def _hals_sweep_h(H, A, G, eps, lam=0.0, gam=0.0, reg=False):
    r, p = len(H), len(H[0])
    for u in range(p):
        for k in range(r):
            s = A[k][u] - lam if reg else A[k][u]
            for j in range(r):
                if j != k:
                    s -= G[k][j] * H[j][u]
            b = _rmax(G[k][k] + gam, eps) if reg else _rmax(G[k][k], eps)
            val = s / b
            H[k][u] = val if val > eps else eps


def _hals_sweep_w(W, B, K, eps, lam=0.0, gam=0.0, reg=False):
    n, r = len(W), len(W[0])
    for i in range(n):
        for k in range(r):
            s = B[i][k] - lam if reg else B[i][k]
            for j in range(r):
                if j != k:
                    s -= K[k][j] * W[i][j]
            d = _rmax(K[k][k] + gam, eps) if reg else _rmax(K[k][k], eps)
            val = s / d
            W[i][k] = val if val > eps else eps


def hals_iteration(V, W, H, eps, reg=False, lh=0.0, lw=0.0, gh=0.0, gw=0.0):
    """lib.rs:666-677 (plain) and 866-872 (dnmf's regularised sweeps)."""
    A = _wtv(V, W)
    G = _gram_cols(W)
    _hals_sweep_h(H, A, G, eps, lh, gh, reg)
    B = _vht(V, H)
    K = _gram_rows(H)
    _hals_sweep_w(W, B, K, eps, lw, gw, reg)


def extract_diagonal(W, H):
    """lib.rs:808-824."""
    n, r, p = len(W), len(W[0]), len(H[0])
    d = []
    for k in range(r):
        sw = 0.0
        for i in range(n):
            sw += W[i][k] * W[i][k]
        sh = 0.0
        for j in range(p):
            sh += H[k][j] * H[k][j]
        nw, nh = math.sqrt(sw), math.sqrt(sh)
        if nw > 0.0:
            for i in range(n):
                W[i][k] /= nw
        if nh > 0.0:
            for j in range(p):
                H[k][j] /= nh
        d.append(nw * nh)
    return d


# -------------------------------------------------------------- offset ----
def offset_iteration(V, W, H, off, eps):
    """One offset iteration (lib.rs:1711-1741): offset_update_h_inner
    (lib.rs:330-362) with den_addon = W^T off, offset_update_w_impl
    (lib.rs:393-450), then the offset update.  Returns (W, H, off)."""
    n, r, p = len(W), len(W[0]), len(H[0])
    G = _zeros(r, r)
    for i in range(r):
        for j in range(r):
            s = 0.0
            for u in reversed(range(n)):
                s += W[u][i] * W[u][j]
            G[i][j] = s
    add = []
    for i in range(r):
        s = 0.0
        for u in reversed(range(n)):
            s += W[u][i] * off[u]
        add.append(s)
    Hn = _zeros(r, p)
    for j in range(p):
        for i in range(r):
            num = 0.0
            for u in reversed(range(n)):
                num += W[u][i] * V[u][j]
            den = 0.0
            for l in reversed(range(r)):
                den += G[i][l] * H[l][j]
            den += add[i]
            t = H[i][j] * num
            Hn[i][j] = (t if t > eps else eps) / (den + eps)
    K = _zeros(r, r)
    for j in range(r):
        for i in range(r):
            s = 0.0
            for u in reversed(range(p)):
                s += Hn[j][u] * Hn[i][u]
            K[i][j] = s
    rsh = []
    for i in range(r):
        s = 0.0
        for u in reversed(range(p)):
            s += Hn[i][u]
        rsh.append(s)
    Wn = _zeros(n, r)
    for i in range(n):
        for j in range(r):
            num = 0.0
            for u in reversed(range(p)):
                num += V[i][u] * Hn[j][u]
            den = 0.0
            for l in reversed(range(r)):
                den += W[i][l] * K[l][j]
            den += off[i] * rsh[j]
            t = W[i][j] * num
            Wn[i][j] = (eps if t < eps else t) / (den + eps)
    offn = list(off)
    for i in range(n):
        rv = 0.0
        rw = 0.0
        for j in range(p):
            rv += V[i][j]
            rw += _wh(Wn, Hn, i, j) + off[i]
        num = rv if rv > eps else eps
        offn[i] = off[i] * num / (rw + eps)
    return Wn, Hn, offn


def row_means(V):
    """Default initial offset (lib.rs:1669-1683)."""
    out = []
    for row in V:
        s = 0.0
        for x in row:
            s += x
        out.append(s / len(row))
    return out


# --------------------------------------------------------------- nsNMF ----
def smoothing(r, theta):
    """lib.rs:457-466."""
    o = theta / r
    return [[(1.0 - theta) + o if a == b else o for b in range(r)] for a in range(r)]


def _mm(A, B):
    """A B with a forward inner sum (the oracle's order for ndarray dot)."""
    out = _zeros(len(A), len(B[0]))
    for a in range(len(A)):
        for b in range(len(B[0])):
            s = 0.0
            for k in range(len(B)):
                s += A[a][k] * B[k][b]
            out[a][b] = s
    return out


def nsnmf_iteration(V, W, H, theta):
    """One nsNMF iteration (lib.rs:1799-1820): H by the brunet H update with
    W S, W by the brunet W update with S H, then col_rescale.  Returns (W, H)."""
    S = smoothing(len(W[0]), theta)
    Hn = brunet_update_h(V, _mm(W, S), H)
    Wn = brunet_update_w(V, W, _mm(S, Hn))
    col_rescale(Wn)
    return Wn, Hn


# --------------------------------------------------------------- lsNMF ----
def lsnmf_iteration(V, S, W, H, eps):
    """One lsNMF iteration (lib.rs:981-1024): est = W H, H by the weighted
    MU with W^T (V.S) / W^T (est.S); est = W H', W by (V.S) H'^T /
    (est.S) H'^T.  Products with a forward inner sum.  Returns (W, H)."""
    n, r, p = len(W), len(W[0]), len(H[0])
    Hn = _zeros(r, p)
    for u in range(p):
        num = [0.0] * r
        den = [0.0] * r
        for i in range(n):
            vw = V[i][u] * S[i][u]
            ew = _wh(W, H, i, u) * S[i][u]
            for k in range(r):
                num[k] += W[i][k] * vw
                den[k] += W[i][k] * ew
        for k in range(r):
            t = H[k][u] * num[k]
            Hn[k][u] = (t if t > eps else eps) / (den[k] + eps)
    Wn = _zeros(n, r)
    for i in range(n):
        num = [0.0] * r
        den = [0.0] * r
        for u in range(p):
            vw = V[i][u] * S[i][u]
            ew = _wh(W, Hn, i, u) * S[i][u]
            for k in range(r):
                num[k] += vw * Hn[k][u]
                den[k] += ew * Hn[k][u]
        for k in range(r):
            t = W[i][k] * num[k]
            Wn[i][k] = (t if t > eps else eps) / (den[k] + eps)
    return Wn, Hn


# ---------------------------------------------------------- objectives ----
def _wh(W, H, i, j):
    s = 0.0
    for k in range(len(H)):
        s += W[i][k] * H[k][j]
    return s


def euclidean_distance(V, W, H):
    """lib.rs:1432-1445 with the oracle's forward dense W H."""
    s = 0.0
    for j in range(len(V[0])):
        for i in range(len(V)):
            d = V[i][j] - _wh(W, H, i, j)
            s += d * d
    return 0.5 * s


def kl_divergence(V, W, H):
    """lib.rs:1406-1429."""
    s = 0.0
    for j in range(len(V[0])):
        for i in range(len(V)):
            vv, ww = V[i][j], _wh(W, H, i, j)
            if vv > 0.0:
                if ww > 0.0:
                    s += vv * math.log(vv / ww) - vv + ww
            else:
                s += ww
    return s


# -------------------------------------------------------------- driver ----
def run(algo, V, W0, H0, max_iter, eps=1e-9, rescale=True, sparsity=0.0, theta=0.5,
        offset=None, weight=None):
    """max_iter stop only (no deviances).  lib.rs:1528-1647 (brunet, lee),
    637-688 (hals), 710-790 (ehals), 839-906 (dnmf), 1649-1759 (offset),
    1761-1833 (nsNMF), 946-1043 (lsNMF, with `weight`).  offset returns
    (W, H, off)."""
    W = [list(map(float, row)) for row in W0]
    H = [list(map(float, row)) for row in H0]
    if algo == "offset":
        off = [float(x) for x in offset] if offset is not None else row_means(V)
        for _ in range(max_iter):
            W, H, off = offset_iteration(V, W, H, off, eps)
        return W, H, off
    if algo == "nsnmf":
        for _ in range(max_iter):
            W, H = nsnmf_iteration(V, W, H, theta)
        return W, H
    if algo == "lsnmf":
        for _ in range(max_iter):
            W, H = lsnmf_iteration(V, weight, W, H, eps)
        return W, H
    Wp = [row[:] for row in W]
    Hp = [row[:] for row in H]
    for it in range(1, max_iter + 1):
        if algo == "lee":
            H = lee_update_h(V, W, H, eps)
            W = lee_update_w(V, W, H, eps)
            if rescale:
                col_rescale(W)
        elif algo == "brunet":
            H = brunet_update_h(V, W, H)
            W = brunet_update_w(V, W, H)
            if it % 10 == 0:
                _pmax(H, DBL_EPSILON)
                _pmax(W, DBL_EPSILON)
        elif algo == "hals":
            hals_iteration(V, W, H, eps)
        elif algo == "ehals":
            Ws = [row[:] for row in W]
            Hs = [row[:] for row in H]
            for X, Xp in ((W, Wp), (H, Hp)):
                for a in range(len(X)):
                    for b in range(len(X[0])):
                        e = X[a][b] + 0.5 * (X[a][b] - Xp[a][b])
                        X[a][b] = e if e > eps else eps
            hals_iteration(V, W, H, eps)
            Wp, Hp = Ws, Hs
        elif algo == "dnmf":
            hals_iteration(V, W, H, eps, True, sparsity, sparsity, 0.0, 0.0)
            d = extract_diagonal(W, H)
            for k in range(len(H)):
                for j in range(len(H[0])):
                    H[k][j] *= d[k]
        else:
            raise ValueError(algo)
    return W, H
