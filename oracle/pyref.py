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


# ---------------------------------------------------- FCNNLS / snmf ----
def _chol(A):
    """Lower Cholesky factor of A (list of rows) or None (fcnnls.rs:40-64)."""
    r = len(A)
    L = [row[:] for row in A]
    for j in range(r):
        s = L[j][j]
        for k in range(j):
            s -= L[j][k] * L[j][k]
        if s <= 0.0:
            return None
        d = math.sqrt(s)
        L[j][j] = d
        for i in range(j + 1, r):
            t = L[i][j]
            for k in range(j):
                t -= L[i][k] * L[j][k]
            L[i][j] = t / d
    return L


def _chol_solve(L, b):
    r = len(L)
    x = list(b)
    for i in range(r):
        s = x[i]
        for k in range(i):
            s -= L[i][k] * x[k]
        x[i] = s / L[i][i]
    for i in reversed(range(r)):
        s = x[i]
        for k in range(i + 1, r):
            s -= L[k][i] * x[k]
        x[i] = s / L[i][i]
    return x


def _spd_solve(A, b):
    """Cholesky solve, else the regularised fallback (fcnnls.rs:95-131)."""
    L = _chol(A)
    if L is None:
        Areg = [row[:] for row in A]
        er = 1e-12
        while True:
            L = _chol(Areg)
            if L is not None:
                break
            for i in range(len(A)):
                Areg[i][i] += er
            er *= 10.0
            if er > 1e6:
                return [0.0] * len(b)
    return _chol_solve(L, b)


def _restricted(ctc, cta_col, passive):
    act = [i for i in range(len(ctc)) if passive[i]]
    out = [0.0] * len(ctc)
    if act:
        sol = _spd_solve([[ctc[a][b] for b in act] for a in act], [cta_col[a] for a in act])
        for a, v in zip(act, sol):
            out[a] = v
    return out


def fcnnls(X, Y, eps):
    """FCNNLS (fcnnls.rs:270-501) on lists; returns K (l x p) as lists of
    rows.  Columns are held as lists here (K_cols[j][i])."""
    nobs, l, p = len(X), len(X[0]), len(Y[0])
    ctc = [[sum_fwd(X[o][a] * X[o][b] for o in range(nobs)) for b in range(l)] for a in range(l)]
    cta = [[sum_fwd(X[o][a] * Y[o][j] for o in range(nobs)) for a in range(l)] for j in range(p)]
    full = [True] * l
    Kc = [_restricted(ctc, cta[j], full) for j in range(p)]
    P = [[Kc[j][i] > 0.0 for i in range(l)] for j in range(p)]
    for j in range(p):
        for i in range(l):
            if not P[j][i]:
                Kc[j][i] = 0.0
    D = [col[:] for col in Kc]
    fset = [j for j in range(p) if sum(P[j]) != l]
    it, max_it = 0, 3 * l
    while fset:
        for j in fset:
            Kc[j] = _restricted(ctc, cta[j], P[j])
        hset = [j for j in fset if any(x < eps for x in Kc[j])]
        while hset and it < max_it:
            it += 1
            if not any(P[j][i] and Kc[j][i] < eps for j in hset for i in range(l)):
                break
            for j in hset:
                alphas = []
                for i in range(l):
                    a = math.inf
                    if P[j][i] and Kc[j][i] < eps:
                        den = D[j][i] - Kc[j][i]
                        if den != 0.0:
                            a = D[j][i] / den
                    alphas.append(a)
                best, bi = math.inf, 0
                for i, a in enumerate(alphas):
                    if a < best:
                        best, bi = a, i
                for i in range(l):
                    D[j][i] -= best * (D[j][i] - Kc[j][i])
                D[j][bi] = 0.0
                P[j][bi] = False
            for j in hset:
                Kc[j] = _restricted(ctc, cta[j], P[j])
            hset = [j for j in range(p) if any(x < eps for x in Kc[j])]
        nf = []
        for j in fset:
            g = [cta[j][a] - sum_fwd(ctc[a][b] * Kc[j][b] for b in range(l)) for a in range(l)]
            mv, mr = -math.inf, 0
            for i in range(l):
                if not P[j][i] and g[i] > mv:
                    mv, mr = g[i], i
            if mv > eps:
                P[j][mr] = True
                D[j] = Kc[j][:]
                nf.append(j)
        fset = nf
        if it >= max_it:
            break
    return [[Kc[j][i] for j in range(p)] for i in range(l)]


def sum_fwd(terms):
    s = 0.0
    for t in terms:
        s += t
    return s


def snmf_kim_park(A, W, H, max_iter, eta, beta, eps):
    """Kim-Park driver (lib.rs:1198-1334) on A (m x n), internal W, H;
    returns (W, H, deviances)."""
    m, n, k = len(A), len(A[0]), len(W[0])
    if eta < 0.0:
        eta = 0.0
        for row in A:
            for x in row:
                eta = max(eta, x)
    sb = math.sqrt(max(beta, 1e-30))
    W = [row[:] for row in W]
    for j in range(k):
        nrm = math.sqrt(sum_fwd(W[i][j] * W[i][j] for i in range(m)))
        if nrm > 0.0:
            for i in range(m):
                W[i][j] = W[i][j] / nrm
    devs = []
    for _ in range(max_iter):
        H = fcnnls(W + [[sb] * k], A + [[0.0] * n], eps)
        Xw = [[H[q][i] for q in range(k)] for i in range(n)] + \
             [[eta if q == j else 0.0 for q in range(k)] for j in range(k)]
        Yw = [[A[j][i] for j in range(m)] for i in range(n)] + [[0.0] * m for _ in range(k)]
        Wt = fcnnls(Xw, Yw, eps)
        W = [[Wt[q][i] for q in range(k)] for i in range(m)]
        s = 0.0
        for i in range(m):
            for j in range(n):
                d = A[i][j] - _wh(W, H, i, j)
                s += d * d
        devs.append(0.5 * s)
    return W, H, devs


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
        offset=None, weight=None, smoothness=-1.0):
    """max_iter stop only (no deviances).  lib.rs:1528-1647 (brunet, lee),
    637-688 (hals), 710-790 (ehals), 839-906 (dnmf), 1649-1759 (offset),
    1761-1833 (nsNMF), 946-1043 (lsNMF, with `weight`), 1198-1334 and
    2117-2146 (snmf/r, snmf/l: beta = sparsity, eta = smoothness).  offset
    returns (W, H, off)."""
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
    if algo == "snmf/r":
        W, H, _ = snmf_kim_park(V, W, H, max_iter, smoothness, sparsity, eps)
        return W, H
    if algo == "snmf/l":
        At = [[V[i][j] for i in range(len(V))] for j in range(len(V[0]))]
        Wi = [[H[q][j] for q in range(len(H))] for j in range(len(H[0]))]
        Hi = [[W[i][q] for i in range(len(W))] for q in range(len(W[0]))]
        Wi, Hi, _ = snmf_kim_park(At, Wi, Hi, max_iter, smoothness, sparsity, eps)
        return ([[Hi[q][i] for q in range(len(Hi))] for i in range(len(Hi[0]))],
                [[Wi[j][q] for j in range(len(Wi))] for q in range(len(Wi[0]))])
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
