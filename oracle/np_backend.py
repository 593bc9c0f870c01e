"""numpy restatement of the reference backend's hot path (oracle; test infrastructure only).

Data model (reference yastn/tensor/_auxiliary.py:26-38): a tensor is one flat 1-D
array holding C-ordered blocks; ``meta`` arguments are nested tuples of ints
with ``(start, stop)`` slices into those flat arrays.

Forward functions restate yastn/backend/backend_np.py; adjoints restate the
hand-written backward passes of yastn/backend/_backend_torch_backwards.py and
yastn/backend/linalg/torch_svd_gesdd.py in numpy.
"""
import numpy as np


def _sl(x):
    return slice(x[0], x[1])


def _groups(meta_new, meta_mrg):
    """Pair each destination block of ``meta_new`` with its run of ``meta_mrg`` rows.

    meta_mrg rows of one destination are consecutive and carry its charge as the
    first entry (backend_np.py:655-656 uses itertools.groupby and asserts tn == t1).
    """
    runs, i = [], 0
    for tn, Dn, sln in meta_new:
        j = i
        while j < len(meta_mrg) and tuple(meta_mrg[j][0]) == tuple(tn):
            j += 1
        if j == i:
            raise AssertionError(f"no source rows for destination block {tn}")
        runs.append(((tn, Dn, sln), meta_mrg[i:j]))
        i = j
    if i != len(meta_mrg):
        raise AssertionError("meta_mrg rows left over")
    return runs


def transpose_and_merge(data, order, meta_new, meta_mrg, Dsize):
    """backend_np.py:653-661: zero-filled gather of permuted blocks into fused blocks."""
    out = np.zeros(Dsize, dtype=data.dtype)
    for (tn, Dn, sln), rows in _groups(meta_new, meta_mrg):
        dst = out[_sl(sln)].reshape(Dn)
        for _, slo, Do, Dslc, Drsh in rows:
            box = tuple(slice(a, b) for a, b in Dslc)
            dst[box] = np.transpose(data[_sl(slo)].reshape(Do), order).reshape(Drsh)
    return out


def transpose_and_merge_adjoint(grad, order, meta_new, meta_mrg, D_source):
    """_backend_torch_backwards.py:376-392: gather under the inverse permutation."""
    inv = np.argsort(order)
    out = np.zeros(D_source, dtype=grad.dtype)
    for (tn, Dn, sln), rows in _groups(meta_new, meta_mrg):
        src = grad[_sl(sln)].reshape(Dn)
        for _, slo, Do, Dslc, _ in rows:
            box = tuple(slice(a, b) for a, b in Dslc)
            permuted_shape = tuple(Do[k] for k in order)
            out[_sl(slo)] = np.transpose(src[box].reshape(permuted_shape), inv).ravel()
    return out


def unmerge(data, meta):
    """backend_np.py:664-669: strided sub-block copy, output fully covered."""
    out = np.empty_like(data)
    for sln, Dn, slo, Do, sub in meta:
        box = tuple(slice(a, b) for a, b in sub)
        out[_sl(sln)] = data[_sl(slo)].reshape(Do)[box].ravel()
    return out


def unmerge_adjoint(grad, meta, fwd_size):
    """_backend_torch_backwards.py:417-426: scatter back into a fully covered source."""
    out = np.empty(fwd_size, dtype=grad.dtype)
    for sln, Dn, slo, Do, sub in meta:
        box = tuple(slice(a, b) for a, b in sub)
        out[_sl(slo)].reshape(Do)[box] = grad[_sl(sln)].reshape(Dn)
    return out


def transpose(data, axes, meta):
    """backend_np.py:646-650: per-block permutation into a new block order."""
    out = np.empty_like(data)
    for sln, Dn, slo, Do in meta:
        out[_sl(sln)] = np.transpose(data[_sl(slo)].reshape(Do), axes).ravel()
    return out


def dot(A, B, meta_dot, Dsize):
    """backend_np.py:580-587: one GEMM per output block, f64 x c128 promotes."""
    dtype = np.promote_types(A.dtype, B.dtype)
    out = np.empty(Dsize, dtype=dtype)
    for slc, Dc, sla, Da, slb, Db in meta_dot:
        out[_sl(slc)] = (A[_sl(sla)].reshape(Da).astype(dtype) @ B[_sl(slb)].reshape(Db).astype(dtype)).ravel()
    return out


def dot_adjoint(gC, A, B, meta_dot):
    """_backend_torch_backwards.py:117-138: gA += gC B^H, gB += A^H gC, accumulated per block."""
    dtype = np.promote_types(A.dtype, B.dtype)
    gA = np.zeros(A.size, dtype=dtype)
    gB = np.zeros(B.size, dtype=dtype)
    for slc, Dc, sla, Da, slb, Db in meta_dot:
        c = gC[_sl(slc)].reshape(Dc)
        a = A[_sl(sla)].reshape(Da)
        b = B[_sl(slb)].reshape(Db)
        gA[_sl(sla)] += (c @ b.conj().T).ravel()
        gB[_sl(slb)] += (a.conj().T @ c).ravel()
    return gA, gB


def svd(data, meta, sizes):
    """backend_np.py:330-340 (LAPACK gesdd via numpy): thin SVD per block, S descending."""
    real = np.zeros(0, dtype=data.dtype).real.dtype
    U = np.empty(sizes[0], dtype=data.dtype)
    S = np.empty(sizes[1], dtype=real)
    V = np.empty(sizes[2], dtype=data.dtype)
    for sl, D, slU, DU, slS, slV, DV in meta:
        u, s, vh = np.linalg.svd(data[_sl(sl)].reshape(D), full_matrices=False)
        U[_sl(slU)] = u.ravel()
        S[_sl(slS)] = s
        V[_sl(slV)] = vh.ravel()
    return U, S, V


def qr(data, meta, sizes):
    """backend_np.py:511-520: economic QR of every block (scipy.linalg.qr there, numpy's LAPACK geqrf
    here), with the signs of R's rows (and Q's columns) chosen so that diag(R) is >= 0."""
    Q = np.zeros(sizes[0], dtype=data.dtype)
    R = np.zeros(sizes[1], dtype=data.dtype)
    for sl, D, slQ, DQ, slR, DR in meta:
        q, r = np.linalg.qr(data[_sl(sl)].reshape(D), mode='reduced')
        s = np.sign(np.real(np.diag(r)))
        s[s == 0] = 1
        Q[_sl(slQ)] = (q * s).ravel()
        R[_sl(slR)] = (s.reshape(-1, 1) * r).ravel()
    return Q, R


def transpose_dot_sum(Adata, Bdata, meta_dot, Areshape, Breshape, Aorder, Border, Dsize):
    """backend_np.py:590-601 (no-fusion tensordot): blocks permuted and reshaped to matrices, each output
    block the sum of its listed products, in list order."""
    dtype = np.promote_types(Adata.dtype, Bdata.dtype)
    out = np.zeros(Dsize, dtype=dtype)
    Ad = [Adata[_sl(sl)].reshape(Di).transpose(Aorder).reshape(Dl, Dr) for sl, Di, Dl, Dr in Areshape]
    Bd = [Bdata[_sl(sl)].reshape(Di).transpose(Border).reshape(Dl, Dr) for sl, Di, Dl, Dr in Breshape]
    for sl, Dslc, tab in meta_dot:
        acc = np.zeros(Dslc, dtype=dtype)
        for ta, tb in tab:
            acc = acc + Ad[ta] @ Bd[tb]
        out[_sl(sl)] = acc.ravel()
    return out


def fix_svd_signs(U, V, meta):
    """backend_np.py:447-458: make the entry of largest (2^40-quantised) modulus in each U column real positive.

    Ties in the quantised modulus go to the first row (argmax semantics).
    """
    U = U.copy()
    V = V.copy()
    q = (np.abs(U) * float(2 ** 40)).astype(np.int64)
    for _, _, slU, DU, _, slV, DV in meta:
        u = U[_sl(slU)].reshape(DU)
        v = V[_sl(slV)].reshape(DV)
        rows = np.argmax(q[_sl(slU)].reshape(DU), axis=0)
        piv = u[rows, np.arange(DU[1])]
        ph = piv / np.abs(piv)
        u *= np.conj(ph)[None, :]
        v *= ph[:, None]
    return U, V


def eigh(data, meta, sizes):
    """backend_np.py:461-474 with meta: Hermitian eigendecomposition per block, ascending."""
    real = np.zeros(0, dtype=data.dtype).real.dtype
    S = np.zeros(sizes[0], dtype=real)
    U = np.zeros(sizes[1], dtype=data.dtype)
    for sl, D, slU, DU, slS in meta:
        w, u = np.linalg.eigh(data[_sl(sl)].reshape(D))
        S[_sl(slS)] = w
        U[_sl(slU)] = u.ravel()
    return S, U


def _idx(m):
    """A mask value: an index list, or a slice (flip_charges embeds with {0: slice(None)},
    tensor/_single.py:222-225)."""
    return m if isinstance(m, slice) else np.asarray(m)


def apply_mask(A, mask, meta, Dsize, axis, ndim):
    """backend_np.py:626-633: select index lists ``mask[tm]`` along ``axis`` of every block."""
    out = np.empty(Dsize, dtype=A.dtype)
    for sln, Dn, sla, Da, tm in meta:
        idx = (slice(None),) * axis + (_idx(mask[tm]),)
        out[_sl(sln)] = A[_sl(sla)].reshape(Da)[idx].ravel()
    return out


def apply_mask_adjoint(grad, mask, meta, size_a, axis, ndim):
    """_backend_torch_backwards.py:270-279: zero-filled scatter of the selected rows."""
    out = np.zeros(size_a, dtype=grad.dtype)
    for sln, Dn, sla, Da, tm in meta:
        idx = (slice(None),) * axis + (_idx(mask[tm]),)
        out[_sl(sla)].reshape(Da)[idx] = grad[_sl(sln)].reshape(Dn)
    return out


def embed_mask(A, mask, meta, Dsize, axis, ndim):
    """backend_np.py:636-643: zero-filled scatter along ``axis``."""
    out = np.zeros(Dsize, dtype=A.dtype)
    for sln, Dn, sla, Da, tm in meta:
        idx = (slice(None),) * axis + (_idx(mask[tm]),)
        out[_sl(sln)].reshape(Dn)[idx] = A[_sl(sla)].reshape(Da)
    return out


def dot_diag(Adata, Bdata, meta, Dsize, axis, a_ndim):
    """backend_np.py:604-611: every output block is B's block scaled along ``axis`` by a diagonal block."""
    dim = [1] * a_ndim
    dim[axis] = -1
    out = np.empty(Dsize, dtype=np.promote_types(Adata.dtype, Bdata.dtype))
    for sln, slb, Db, sla in meta:
        out[_sl(sln)] = (Adata[_sl(sla)].reshape(dim) * Bdata[_sl(slb)].reshape(Db)).ravel()
    return out


def dot_diag_adjoint(g, Adata, Bdata, meta, axis, a_ndim):
    """Gradients of sum(conj(g) * dot_diag(A, B)) in torch's convention (the reference differentiates the
    torch ops of backend_torch.py:557-564): gB = g conj(A) broadcast, gA = the sum of g conj(B) over the
    other axes, accumulated over the blocks that read the same diagonal block."""
    dim = [1] * a_ndim
    dim[axis] = -1
    dt = np.promote_types(Adata.dtype, Bdata.dtype)
    gA = np.zeros(Adata.size, dtype=dt)
    gB = np.zeros(Bdata.size, dtype=dt)
    other = tuple(d for d in range(a_ndim) if d != axis)
    for sln, slb, Db, sla in meta:
        gb = g[_sl(sln)].reshape(Db)
        gB[_sl(slb)] = (gb * np.conj(Adata[_sl(sla)]).reshape(dim)).ravel()
        gA[_sl(sla)] += (gb * np.conj(Bdata[_sl(slb)].reshape(Db))).sum(axis=other)
    if not np.iscomplexobj(Adata):
        gA = gA.real
    if not np.iscomplexobj(Bdata):
        gB = gB.real
    return gA, gB


def _lorentz(x, eps=1e-12):
    """torch_svd_gesdd.py:4-6: x / (x^2 + eps) with eps fixed to 1e-12 (the passed value is ignored)."""
    return x / (x * x + eps)


def svd_block_adjoint(u, s, vh, gu, gs, gvh):
    """torch_svd_gesdd.py:296-428 (SVDGESDD.backward) for one thin block.

    Lorentzian-broadened F = 1/(s_j - s_i), G = 1/(s_j + s_i), zero diagonals;
    projector terms when the block is rectangular; the extra imaginary diagonal
    term for complex data.
    """
    m, k = u.shape
    n = vh.shape[1]
    cplx = np.iscomplexobj(u) or np.iscomplexobj(vh)
    dtype = np.complex128 if cplx else np.float64
    out = np.zeros((m, n), dtype=dtype)
    if gs is not None:
        out += (u * gs[None, :]) @ vh
    if gu is None and gvh is None:
        return out
    sinv = _lorentz(s)
    F = _lorentz(s[None, :] - s[:, None])
    np.fill_diagonal(F, 0)
    G = _lorentz(s[None, :] + s[:, None])
    np.fill_diagonal(G, 0)
    uh = u.conj().T
    v = vh.conj().T
    if gu is not None:
        t = u @ ((F + G) * (uh @ gu - gu.conj().T @ u)) * 0.5
        if m > k:
            t = t + (np.eye(m) - u @ uh) @ (gu * sinv[None, :])
        out += t @ vh
    if gvh is not None:
        gv = gvh.conj().T
        t = ((F - G) * (vh @ gv - gvh @ v)) @ vh * 0.5
        if n > k:
            t = t + sinv[:, None] * (gvh @ (np.eye(n) - v @ vh))
        out += u @ t
    if cplx and gu is not None:
        L = np.diagonal(uh @ gu).imag * sinv
        out += (u * (1j * L)[None, :]) @ vh
    return out


def svd_adjoint(gU, gS, gV, U, S, V, meta, data_size):
    """_backend_torch_backwards.py:53-66: per-block svd_block_adjoint into a zero-filled buffer."""
    out = np.zeros(data_size, dtype=U.dtype)
    for sl, D, slU, DU, slS, slV, DV in meta:
        u = U[_sl(slU)].reshape(DU)
        vh = V[_sl(slV)].reshape(DV)
        s = S[_sl(slS)]
        g = svd_block_adjoint(u, s, vh,
                              None if gU is None else gU[_sl(slU)].reshape(DU),
                              None if gS is None else gS[_sl(slS)],
                              None if gV is None else gV[_sl(slV)].reshape(DV))
        out[_sl(sl)] = g.ravel()
    return out


def eigh_block_adjoint(w, u, gw, gu):
    """Adjoint of an ascending Hermitian eigendecomposition, the formula of torch.linalg.eigh's backward
    (used by the reference, backend_torch.py:378; no broadening):
    gA = U X U^H with X = skew(U^H gU) / E, E_ij = w_j - w_i (E_ii = 1), and diag(X) = gw when gw is given.
    """
    n = u.shape[0]
    dtype = np.result_type(u.dtype, np.float64)
    X = np.zeros((n, n), dtype=dtype)
    if gu is not None:
        uhgu = u.conj().T @ gu
        X = (uhgu - uhgu.conj().T) * 0.5
        E = w[None, :] - w[:, None]
        np.fill_diagonal(E, 1.0)
        X = X / E
    if gw is not None:
        X = X.astype(dtype)
        X[np.arange(n), np.arange(n)] = gw
    return u @ X @ u.conj().T
