"""Plain per-block GPU path: the comparison baseline for the batched HIP backend.

Same module surface as ``backend_hip``; the hot-path functions are restated the way the reference
torch backend runs them on a GPU (yastn/backend/_backend_torch_backwards.py): a Python loop over the
blocks of every call with one or more torch launches per block (slice / permute / copy, one ``@``
per sector, ``torch.linalg.svd`` per block), and per-block adjoints in backward.  It is a timing
baseline only; nothing in the HIP path uses it.
"""
import numpy as np
import torch

from . import backend_hip as _hip

for _name in _hip.__all__:
    globals()[_name] = getattr(_hip, _name)

BACKEND_ID = "torch_plain"
__all__ = list(_hip.__all__)


def _sl(x):
    return slice(x[0], x[1])


def _groups(meta_new, meta_mrg):
    i = 0
    for tn, Dn, sln in meta_new:
        j = i
        while j < len(meta_mrg) and tuple(meta_mrg[j][0]) == tuple(tn):
            j += 1
        yield (tn, Dn, sln), meta_mrg[i:j]
        i = j


def qr(data, meta, sizes):
    """The reference torch backend's qr (backend_torch.py:475-484): torch.linalg.qr per block, signs so
    that diag(R) >= 0, differentiated by torch."""
    Q = torch.zeros(sizes[0], dtype=data.dtype, device=data.device)
    R = torch.zeros(sizes[1], dtype=data.dtype, device=data.device)
    for sl, D, slQ, DQ, slR, DR in meta:
        q, r = torch.linalg.qr(data[_sl(sl)].view(D))
        sg = torch.sign(torch.real(torch.diagonal(r)))
        sg = torch.where(sg == 0, torch.ones_like(sg), sg)
        Q[_sl(slQ)].view(DQ)[:] = q * sg
        R[_sl(slR)].view(DR)[:] = sg.reshape(-1, 1) * r
    return Q, R


class _TransposeAndMerge(torch.autograd.Function):
    @staticmethod
    def forward(data, order, meta_new, meta_mrg, Dsize):
        out = torch.zeros((Dsize,), dtype=data.dtype, device=data.device)
        for (tn, Dn, sln), rows in _groups(meta_new, meta_mrg):
            blk = out[_sl(sln)].reshape(Dn)
            for _, slo, Do, Dslc, Drsh in rows:
                blk[tuple(slice(*x) for x in Dslc)] = data[_sl(slo)].reshape(Do).permute(order).reshape(Drsh)
        return out

    @staticmethod
    def setup_context(ctx, inputs, output):
        data, order, meta_new, meta_mrg, Dsize = inputs
        ctx.order, ctx.meta_new, ctx.meta_mrg, ctx.n = order, meta_new, meta_mrg, data.numel()

    @staticmethod
    def backward(ctx, g):
        inv = tuple(int(x) for x in np.argsort(ctx.order))
        out = torch.zeros((ctx.n,), dtype=g.dtype, device=g.device)
        for (tn, Dn, sln), rows in _groups(ctx.meta_new, ctx.meta_mrg):
            gb = g[_sl(sln)].view(Dn)
            for _, slo, Do, Dslc, _ in rows:
                shp = tuple(Do[k] for k in ctx.order)
                out[_sl(slo)].reshape(Do)[:] = gb[tuple(slice(*x) for x in Dslc)].reshape(shp).permute(inv)
        return out, None, None, None, None


class _Unmerge(torch.autograd.Function):
    @staticmethod
    def forward(data, meta):
        out = torch.empty_like(data)
        for sln, Dn, slo, Do, sub in meta:
            out[_sl(sln)].view(Dn)[:] = data[_sl(slo)].view(Do)[tuple(slice(*x) for x in sub)]
        return out

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.meta, ctx.n = inputs[1], inputs[0].numel()

    @staticmethod
    def backward(ctx, g):
        out = torch.empty(ctx.n, dtype=g.dtype, device=g.device)
        for sln, Dn, slo, Do, sub in ctx.meta:
            out[_sl(slo)].view(Do)[tuple(slice(*x) for x in sub)] = g[_sl(sln)].view(Dn)
        return out, None


class _Dot(torch.autograd.Function):
    @staticmethod
    def forward(A, B, meta_dot, Dsize):
        dtype = torch.promote_types(A.dtype, B.dtype)
        A, B = A.to(dtype), B.to(dtype)
        out = torch.zeros(Dsize, dtype=dtype, device=A.device)
        for slc, Dc, sla, Da, slb, Db in meta_dot:
            out[_sl(slc)].view(Dc)[:] = A[_sl(sla)].view(Da) @ B[_sl(slb)].view(Db)
        return out

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.save_for_backward(inputs[0], inputs[1])
        ctx.meta = inputs[2]

    @staticmethod
    def backward(ctx, gC):
        A0, B0 = ctx.saved_tensors
        dtype = torch.promote_types(A0.dtype, B0.dtype)
        A, B = A0.to(dtype), B0.to(dtype)
        gA = torch.zeros_like(A)
        gB = torch.zeros_like(B)
        for slc, Dc, sla, Da, slb, Db in ctx.meta:
            c = gC[_sl(slc)].view(Dc)
            gA[_sl(sla)].view(Da)[:] += c @ B[_sl(slb)].view(Db).adjoint()
            gB[_sl(slb)].view(Db)[:] += A[_sl(sla)].view(Da).adjoint() @ c
        if not A0.is_complex() and gA.is_complex():
            gA = gA.real
        if not B0.is_complex() and gB.is_complex():
            gB = gB.real
        return gA, gB, None, None


AD_EPS = 1e-12  # the reference's fixed Lorentzian broadening (torch_svd_gesdd.py:4-6)


def set_ad_broadening(eps):
    global AD_EPS
    AD_EPS = float(eps)


def _lor(x):
    return x / (x * x + AD_EPS)


def _svd_block_backward(u, s, vh, gu, gs, gvh):
    """Per-block SVD adjoint with the reference's Lorentzian broadening (torch_svd_gesdd.py:296-428)."""
    m, k = u.shape
    n = vh.shape[1]
    out = (u * gs.unsqueeze(-2)) @ vh
    sinv = _lor(s)
    F = _lor(s.unsqueeze(-2) - s.unsqueeze(-1))
    F.diagonal().fill_(0)
    G = _lor(s.unsqueeze(-2) + s.unsqueeze(-1))
    G.diagonal().fill_(0)
    uh = u.conj().transpose(-2, -1)
    t = u @ ((F + G) * (uh @ gu - gu.conj().transpose(-2, -1) @ u)) * 0.5
    if m > k:
        proj = -u @ uh
        proj.diagonal().add_(1)
        t = t + proj @ (gu * sinv.unsqueeze(-2))
    out = out + t @ vh
    v = vh.conj().transpose(-2, -1)
    gv = gvh.conj().transpose(-2, -1)
    t = ((F - G) * (vh @ gv - gvh @ v)) @ vh * 0.5
    if n > k:
        proj = -v @ vh
        proj.diagonal().add_(1)
        t = t + sinv.unsqueeze(-1) * (gvh @ proj)
    out = out + u @ t
    if u.is_complex():
        L = (uh @ gu).diagonal().imag * sinv
        out = out + (u * (1j * L).unsqueeze(-2)) @ vh
    return out


class _Svd(torch.autograd.Function):
    @staticmethod
    def forward(data, meta, sizes, fullrank_uv=False, ad_decomp_reg=1e-12, diagnostics=None):
        real = data.real.dtype if data.is_complex() else data.dtype
        U = torch.empty(sizes[0], dtype=data.dtype, device=data.device)
        S = torch.empty(sizes[1], dtype=real, device=data.device)
        V = torch.empty(sizes[2], dtype=data.dtype, device=data.device)
        for sl, D, slU, DU, slS, slV, DV in meta:
            u, s, vh = torch.linalg.svd(data[_sl(sl)].view(D), full_matrices=False)
            U[_sl(slU)].reshape(DU)[:] = u
            S[_sl(slS)] = s
            V[_sl(slV)].reshape(DV)[:] = vh
        return U, S, V

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.save_for_backward(*output)
        ctx.meta, ctx.n = inputs[1], inputs[0].numel()

    @staticmethod
    def backward(ctx, gU, gS, gV):
        U, S, V = ctx.saved_tensors
        out = torch.zeros(ctx.n, dtype=U.dtype, device=U.device)
        for sl, D, slU, DU, slS, slV, DV in ctx.meta:
            out[_sl(sl)].view(D)[:] = _svd_block_backward(U[_sl(slU)].view(DU), S[_sl(slS)], V[_sl(slV)].view(DV),
                                                          gU[_sl(slU)].view(DU), gS[_sl(slS)], gV[_sl(slV)].view(DV))
        return out, None, None, None, None, None


def transpose_and_merge(data, order, meta_new, meta_mrg, Dsize):
    return _TransposeAndMerge.apply(data, order, meta_new, meta_mrg, Dsize)


def unmerge(data, meta):
    return _Unmerge.apply(data, meta)


def dot(Adata, Bdata, meta_dot, Dsize):
    return _Dot.apply(Adata, Bdata, meta_dot, Dsize)


def svd(data, meta, sizes, fullrank_uv=False, ad_decomp_reg=1.0e-12, diagnostics=None, **kwargs):
    return _Svd.apply(data, meta, sizes, fullrank_uv, ad_decomp_reg, diagnostics)


def fix_svd_signs(Udata, Vhdata, meta):
    Ud = torch.empty_like(Udata)
    Vd = torch.empty_like(Vhdata)
    q = (Udata.detach().abs() * (2 ** 40)).to(torch.int64)
    for _, _, slU, DU, _, slV, DV in meta:
        u = Udata[_sl(slU)].reshape(DU)
        v = Vhdata[_sl(slV)].reshape(DV)
        idx = torch.argmax(q[_sl(slU)].reshape(DU), dim=0, keepdim=True)
        ph = torch.take_along_dim(u, idx, dim=0)
        ph = ph / ph.abs()
        Ud[_sl(slU)].reshape(DU)[:] = u * ph.conj().reshape(1, -1)
        Vd[_sl(slV)].reshape(DV)[:] = v * ph.reshape(-1, 1)
    return Ud, Vd


def eigh(data, meta=None, sizes=(1, 1), order_by_magnitude=False, ad_decomp_reg=1.0e-12):
    if meta is None:
        return torch.linalg.eigh(data)
    real = data.real.dtype if data.is_complex() else data.dtype
    S = torch.zeros((sizes[0],), dtype=real, device=data.device)
    U = torch.zeros((sizes[1],), dtype=data.dtype, device=data.device)
    for sl, D, slU, DU, slS in meta:
        s, u = torch.linalg.eigh(data[_sl(sl)].view(D))
        S[_sl(slS)] = s
        U[_sl(slU)].view(DU)[:] = u
    return S, U


def apply_mask(Adata, mask, meta, Dsize, axis, a_ndim):
    pre = (slice(None),) * axis
    out = torch.empty(Dsize, dtype=Adata.dtype, device=Adata.device)
    for sln, Dn, sla, Da, tm in meta:
        idx = pre + (mask[tm].to(Adata.device),)
        out[_sl(sln)].view(Dn)[:] = Adata[_sl(sla)].view(Da)[idx]
    return out


def embed_mask(Adata, mask, meta, Dsize, axis, a_ndim):
    pre = (slice(None),) * axis
    out = torch.zeros(Dsize, dtype=Adata.dtype, device=Adata.device)
    for sln, Dn, sla, Da, tm in meta:
        out[_sl(sln)].view(Dn)[pre + (mask[tm].to(Adata.device),)] = Adata[_sl(sla)].view(Da)
    return out


def transpose(data, axes, meta_transpose):
    out = torch.zeros_like(data)
    for sln, Dn, slo, Do in meta_transpose:
        out[_sl(sln)].view(Dn)[:] = data[_sl(slo)].view(Do).permute(axes)
    return out


def dot_diag(Adata, Bdata, meta, Dsize, axis, a_ndim):
    """The reference torch backend's dot_diag (backend_torch.py:557-564): one broadcast product per block,
    differentiated by torch."""
    dim = [1] * a_ndim
    dim[axis] = -1
    dtype = torch.promote_types(Adata.dtype, Bdata.dtype)
    newdata = torch.empty(Dsize, dtype=dtype, device=Adata.device)
    for sln, slb, Db, sla in meta:
        newdata[_sl(sln)].reshape(Db)[:] = Adata[_sl(sla)].reshape(dim) * Bdata[_sl(slb)].reshape(Db)
    return newdata
