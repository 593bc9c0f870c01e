"""numpy interpreters of the device work lists, used by CPU tests to check plans without a GPU.

They execute the documented semantics of include/yastn_hip.h (jobs, gemm/term words) with numpy
indexing; they are independent of the HIP kernels, which the GPU tests check.
"""
import numpy as np

YH_MAXR = 8


def gather(h_jobs, src, dst):
    for w in h_jobs:
        src_off, dst_off, r = int(w[0]), int(w[1]), int(w[2])
        M = YH_MAXR
        shape = tuple(int(x) for x in w[6:6 + r])
        ss = np.array(w[6 + M:6 + M + r], dtype=np.int64)
        ds = np.array(w[6 + 2 * M:6 + 2 * M + r], dtype=np.int64)
        idx = np.indices(shape).reshape(r, -1)
        d = dst_off + ds @ idx
        if src_off < 0:
            dst[d] = 0
        else:
            dst[d] = src[src_off + ss @ idx]
    return dst


def gemm(h_gemms, h_terms, outs, ins):
    for g in h_gemms:
        M, N, c_off, c_rs, c_cs, tb, tc, fl = (int(x) for x in g)
        C = outs[(fl >> 8) & 0xff]
        acc = np.zeros((M, N), dtype=C.dtype)
        for t in h_terms[tb:tb + tc]:
            a_off, a_rs, a_cs, b_off, b_rs, b_cs, K, tf = (int(x) for x in t)
            A = ins[(tf >> 8) & 0xff]
            B = ins[(tf >> 16) & 0xff]
            ia = a_off + np.arange(M)[:, None] * a_rs + np.arange(K)[None, :] * a_cs
            ib = b_off + np.arange(K)[:, None] * b_rs + np.arange(N)[None, :] * b_cs
            a = A[ia] if K and M else np.zeros((M, K), C.dtype)
            b = B[ib] if K and N else np.zeros((K, N), C.dtype)
            if tf & 1:
                a = np.conj(a)
            if tf & 2:
                b = np.conj(b)
            acc = acc + a @ b
        ic = c_off + np.arange(M)[:, None] * c_rs + np.arange(N)[None, :] * c_cs
        if fl & 1:
            C[ic] += acc
        else:
            C[ic] = acc
    return outs


def _lor(x, eps):
    return x / (x * x + eps)


def svd_adjoint_core(descs, J, T, X, Y, S, gS, gU, gVh, eps, has_gS, has_gU, has_gV):
    """numpy transliteration of the svd adjoint core kernel (yastn_amd/csrc/misc.hip)."""
    for m, n, k, u_off, s_off, v_off, kk, tt in (tuple(int(x) for x in d) for d in descs):
        s = S[s_off:s_off + k]
        F = _lor(s[None, :] - s[:, None], eps)
        G = _lor(s[None, :] + s[:, None], eps)
        np.fill_diagonal(F, 0)
        np.fill_diagonal(G, 0)
        sinv = _lor(s, eps)
        Jb = np.zeros((k, k), dtype=J.dtype)
        if has_gU:
            x = X[kk:kk + k * k].reshape(k, k)
            Jb += 0.5 * (F + G) * (x - x.conj().T)
            if np.iscomplexobj(J):
                Jb[np.arange(k), np.arange(k)] += 1j * np.diag(x).imag * sinv
            if m > k:
                Jb -= x * sinv[None, :]
        if has_gV:
            y = Y[kk:kk + k * k].reshape(k, k)
            Jb += 0.5 * (F - G) * (y - y.conj().T)
            if n > k:
                Jb -= sinv[:, None] * y.conj().T
        if has_gS:
            Jb[np.arange(k), np.arange(k)] += gS[s_off:s_off + k]
        J[kk:kk + k * k] = Jb.ravel()
        if m >= n:
            t = np.zeros((m, k), dtype=J.dtype)
            if has_gU and m > k:
                t = gU[u_off:u_off + m * k].reshape(m, k) * sinv[None, :]
            T[tt:tt + m * k] = t.ravel()
        else:
            t = np.zeros((k, n), dtype=J.dtype)
            if has_gV:
                t = sinv[:, None] * gVh[v_off:v_off + k * n].reshape(k, n)
            T[tt:tt + k * n] = t.ravel()


QR_NB = 32


def _colmajor(buf, off, rows, cols):
    """View of a column-major rows x cols matrix stored at buf[off:]."""
    return buf[off:off + rows * cols].reshape(cols, rows).T


def qr_panel(work, qd, step):
    """numpy transliteration of the Householder panel kernel (yastn_amd/csrc/jacobi.hip)."""
    p, q, w_off, vall_off, tm_off, x_off, y_off, _ = (int(x) for x in qd)
    k0 = step * QR_NB
    nb = min(QR_NB, q - k0)
    Lr = p - k0
    W0 = _colmajor(work, w_off, p, q)
    Vall = _colmajor(work, vall_off, p, q)
    P = W0[k0:, k0:k0 + nb].copy()
    taus = []
    for j in range(nb):
        alpha = P[j, j]
        xn2 = float(np.sum(np.abs(P[j + 1:, j]) ** 2))
        if xn2 == 0.0 and np.imag(alpha) == 0.0:
            tau, beta = 0.0, np.real(alpha)
        else:
            nrm = np.sqrt(abs(alpha) ** 2 + xn2)
            beta = -nrm if np.real(alpha) >= 0 else nrm
            tau = (beta - alpha) / beta
            P[j + 1:, j] = P[j + 1:, j] / (alpha - beta)
        P[j, j] = beta
        v = np.concatenate([[1.0], P[j + 1:, j]])
        for c in range(j + 1, nb):
            w = np.vdot(v, P[j:, c])
            P[j:, c] = P[j:, c] - np.conj(tau) * v * w
        taus.append(tau)
    V = np.zeros((Lr, nb), dtype=work.dtype)
    for j in range(nb):
        V[j, j] = 1.0
        V[j + 1:, j] = P[j + 1:, j]
    T = np.zeros((nb, nb), dtype=work.dtype)
    for j in range(nb):
        T[j, j] = taus[j]
        z = V[:, :j].conj().T @ V[:, j]
        T[:j, j] = -taus[j] * (T[:j, :j] @ z)
    Tm = _colmajor(work, tm_off + step * QR_NB * QR_NB, QR_NB, QR_NB)
    Tm[:] = 0
    Tm[:nb, :nb] = -np.triu(T)
    W0[k0:, k0:k0 + nb] = np.triu(P)
    Vall[k0:, k0:k0 + nb] = V


def splitk_reduce(h_red, outs, part):
    """numpy transliteration of the split-K reduction kernel (yastn_amd/csrc/gemm.hip)."""
    for M, N, c_off, c_rs, c_cs, p_off, S, fl in (tuple(int(x) for x in r) for r in h_red):
        acc = part[p_off:p_off + M * N].reshape(M, N).copy()
        for s in range(1, S):
            acc = acc + part[p_off + s * M * N:p_off + (s + 1) * M * N].reshape(M, N)
        C = outs[(fl >> 8) & 0xff]
        ic = c_off + np.arange(M)[:, None] * c_rs + np.arange(N)[None, :] * c_cs
        if fl & 1:
            C[ic] += acc
        else:
            C[ic] = acc


class SlotOrder:
    """Mirror of the inner kernel's slot order (panel_inner_kernel's parameter lanes, jacobi.hip): slot k of
    sub-round r -> (i, j), j == nc when the column has no partner."""

    def __init__(self, intra, ci, cj, nc, k):
        self.intra, self.ci, self.cj, self.nc, self.k, self.r = intra, ci, cj, nc, k, 0
        if intra:
            ne = nc + (nc & 1)
            self.n1 = ne - 1
            m = self.n1 if self.n1 > 0 else 1
            self.ph1, self.ph2 = k % m, (m - k % m) % m
            self.kind = 0
        elif ci <= cj:
            self.kind, self.ph1 = 1, k % cj
        else:
            self.kind, self.ph1 = 2, k % ci

    def get(self):
        k, ci, cj, nc = self.k, self.ci, self.cj, self.nc
        if self.kind == 0:
            return (self.r, self.n1) if k == 0 else (min(self.ph1, self.ph2), max(self.ph1, self.ph2))
        if self.kind == 1:
            return (k, ci + self.ph1) if k < ci else (ci + self.ph1, nc)
        return (self.ph1, ci + k if k < cj else nc)

    def advance(self):
        self.r += 1
        if self.kind == 0:
            self.ph1 = 0 if self.ph1 + 1 == self.n1 else self.ph1 + 1
            self.ph2 = 0 if self.ph2 + 1 == self.n1 else self.ph2 + 1
        elif self.kind == 1:
            self.ph1 = 0 if self.ph1 + 1 == self.cj else self.ph1 + 1
        else:
            self.ph1 = 0 if self.ph1 + 1 == self.ci else self.ph1 + 1


def _rotation(a, b, g, tol, floor2):
    g2 = abs(g) ** 2
    if not (g2 > tol * tol * a * b) or a <= floor2 or b <= floor2:
        return None
    e = g / abs(g)
    zeta = 0.5 * (b - a) / abs(g)
    t = 1.0 / (abs(zeta) + np.sqrt(1 + zeta * zeta))
    if zeta < 0:
        t = -t
    c = 1 / np.sqrt(1 + t * t)
    return c, c * t, e


def inner_sweep(G, intra, ci, cj, tol, floor2, check_cover=True):
    """The kernel's inner sweep on the pair's Gram matrix G (nc x nc): per sub-round every G block
    (slot pa rows, slot pb columns) and Q piece (row, slot pb) is recomputed from the previous G, into a
    fresh buffer.  Returns (Q, rotated); asserts that each sub-round writes every entry of G."""
    nc = G.shape[0]
    ne = nc + (nc & 1)
    npr = ne // 2 if intra else max(ci, cj)
    nsub = ne - 1 if intra else npr
    Q = np.eye(nc, dtype=G.dtype)
    orders = [SlotOrder(intra, ci, cj, nc, k) for k in range(npr)]
    rotated = False
    for _ in range(nsub):
        slots = [o.get() for o in orders]
        rots = []
        for i, j in slots:
            r = None if j >= nc else _rotation(G[i, i].real, G[j, j].real, G[i, j], tol, floor2)
            rots.append(r)
            rotated |= r is not None
        Gn = np.full_like(G, np.nan)
        J = np.eye(nc, dtype=G.dtype)
        for (i, j), r in zip(slots, rots):
            if r is not None:
                c, s, e = r
                # columns (w_i, w_j) <- (c w_i - s conj(e) w_j, s e w_i + c w_j)
                J[i, i], J[j, i], J[i, j], J[j, j] = c, -s * np.conj(e), s * e, c
        blk = J.conj().T @ G @ J
        if not check_cover:
            Gn = blk
        for pa, (i, j) in enumerate(slots if check_cover else ()):
            rows = [i] + ([j] if j < nc else [])
            for pb, (i2, j2) in enumerate(slots):
                cols = [i2] + ([j2] if j2 < nc else [])
                Gn[np.ix_(rows, cols)] = blk[np.ix_(rows, cols)]
        if check_cover:
            assert not np.isnan(Gn).any(), 'a G entry was not written in a sub-round'
        G = Gn
        Q = Q @ J
        for o in orders:
            o.advance()
    # every column pair rotated exactly once in the sweep's order: each (i, j) appears once
    return Q, rotated


def panel_sweeps(plan_jobs, plan_rounds, blocks_words, work, floors, b, max_sweeps, mode_tr):
    """The panel rounds' iteration on W (p x q, column-major at work_off) and V (q x q after it):
    per sweep the rounds in order, per job the Gram matrix of its columns, the inner sweep and W <- W Q,
    V <- V Q.  Returns the sweeps used (or -1)."""
    for sweep in range(max_sweeps):
        rotated = False
        for off, cnt in plan_rounds:
            for job in plan_jobs[off:off + cnt]:
                blk, pi, pj = (int(x) for x in job)
                w = blocks_words[blk]
                m, n, w_off = int(w[1]), int(w[2]), int(w[6])
                tr = m < n
                p, q = (n, m) if tr else (m, n)
                intra = pi == pj
                c0i, c0j = pi * b, pj * b
                ci = min(b, q - c0i)
                cj = 0 if intra else min(b, q - c0j)
                cols = list(range(c0i, c0i + ci)) + ([] if intra else list(range(c0j, c0j + cj)))
                W = work[w_off:w_off + p * q].reshape(q, p).T
                V = work[w_off + p * q:w_off + p * q + q * q].reshape(q, q).T
                G = W[:, cols].conj().T @ W[:, cols]
                tol = np.finfo(float).eps * np.sqrt(max(p, 1))
                Qm, rot = inner_sweep(G, intra, ci, cj, tol, floors[blk], check_cover=sweep == 0)
                if rot:
                    W[:, cols] = W[:, cols] @ Qm
                    V[:, cols] = V[:, cols] @ Qm
                    rotated = True
        if not rotated:
            return sweep + 1
    return -1
