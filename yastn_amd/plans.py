"""Compile the reference planner's meta tuples into device work lists for the C-ABI kernels.

The reference planner (yastn/tensor/_merging.py, _contractions.py, linalg.py) emits
nested tuples of ints that are lru-cached, so the same object comes back every
iteration of an algorithm.  Plans are therefore cached by the identity of the meta
object (a strong reference is kept so that the id cannot be reused), plus the
scalar arguments.  Each plan is validated once on the host by the library's
``yastn_hip_check_*_plan`` before it is uploaded, so no kernel is ever launched
with an out-of-bounds work list.
"""
import ctypes
from collections import OrderedDict
import os
import math
from itertools import product

import numpy as np
import torch

from . import _lib as L


# --------------------------------------------------------------------------- cache

class _Cache:
    """Bounded LRU plan cache.  Lookup by the identity of the meta objects first (the reference planner
    returns lru-cached tuples); on a miss, by their content (some metas, e.g. the SVD's, are rebuilt on
    every call by the reference), so a plan is compiled once per distinct structure.

    Both maps are LRU-ordered and bounded: an identity entry holds its meta objects strongly (so that
    their ids cannot be reused while the entry lives), and metas rebuilt on every call would otherwise
    accumulate one entry per call.  The least recently used entries are evicted one at a time; a plan
    evicted and rebuilt later gets the same tile choice, which the autotuner keeps per block-size
    histogram."""

    def __init__(self, maxsize=4096):
        self.d = OrderedDict()
        self.by_content = OrderedDict()
        self.maxsize = maxsize

    def _put(self, od, key, val):
        od[key] = val
        od.move_to_end(key)
        while len(od) > self.maxsize:
            od.popitem(last=False)

    def __len__(self):
        return len(self.d)

    def get(self, key, refs, build):
        hit = self.d.get(key)
        if hit is not None and all(a is b for a, b in zip(hit[0], refs)):
            self.d.move_to_end(key)
            return hit[1]
        ckey = (key[0],) + tuple(x for x in key[1:] if not isinstance(x, int) or x not in {id(r) for r in refs})
        try:
            chash = hash((ckey, refs))
        except TypeError:
            chash = None
        if chash is not None:
            hit2 = self.by_content.get(chash)
            if hit2 is not None and hit2[0] == (ckey, refs):
                self.by_content.move_to_end(chash)
                self._put(self.d, key, (refs, hit2[1]))
                return hit2[1]
        val = build()
        self._put(self.d, key, (refs, val))
        if chash is not None:
            self._put(self.by_content, chash, ((ckey, refs), val))
        return val


_CACHE = _Cache()


def _device_key(device):
    return (device.type, device.index)


# --------------------------------------------------------------------------- box gather

def _c_strides(shape):
    st, acc = [], 1
    for d in reversed(shape):
        st.append(acc)
        acc *= d
    return tuple(reversed(st))


def _coalesce(shape, ss, ds):
    """Drop unit dims and merge neighbours that are contiguous in both source and destination."""
    dims = [(n, s, d) for n, s, d in zip(shape, ss, ds) if n != 1]
    if not dims:
        return (1,), (0,), (1,)
    out = [list(dims[-1])]
    for n, s, d in reversed(dims[:-1]):
        n0, s0, d0 = out[-1]
        if s == s0 * n0 and d == d0 * n0:
            out[-1] = [n0 * n, s0, d0]
        else:
            out.append([n, s, d])
    out.reverse()
    return tuple(x[0] for x in out), tuple(x[1] for x in out), tuple(x[2] for x in out)


class GatherPlan:
    """Device work list of the box-gather engine (see include/yastn_hip.h)."""

    def __init__(self, jobs, src_numel, dst_numel, device):
        # jobs: list of (src_off, dst_off, shape, src_strides, dst_strides) in destination C order
        rows, tile_job = [], []
        for (src_off, dst_off, shape, ss, ds) in jobs:
            if any(n == 0 for n in shape):
                continue
            shape, ss, ds = _coalesce(shape, ss, ds)
            if len(shape) > L.YH_MAXR:
                raise ValueError(f"box of rank {len(shape)} after coalescing exceeds {L.YH_MAXR}")
            if ds[-1] != 1:  # destination inner dim must be unit stride: append a unit dim
                shape, ss, ds = shape + (1,), ss + (0,), ds + (1,)
            r = len(shape)
            kind, tdim = 0, 0
            if src_off >= 0 and ss[-1] != 1:
                cand = [k for k in range(r - 1) if ss[k] == 1]
                if cand and shape[cand[0]] >= 8 and shape[-1] >= 8:
                    kind, tdim = 1, cand[0]
            numel = math.prod(shape)
            if kind == 0:
                ntile = -(-numel // L.YH_GATHER_TILE)
            else:
                te = L.YH_TRANSPOSE_TILE
                ntile = (-(-shape[tdim] // te)) * (-(-shape[-1] // te)) * (numel // (shape[tdim] * shape[-1]))
            w = np.zeros(L.YH_JOB_WORDS, dtype=np.int64)
            w[0], w[1], w[2], w[3], w[4], w[5] = src_off, dst_off, r, kind, tdim, len(tile_job)
            M = L.YH_MAXR
            w[6:6 + M] = 1
            w[6:6 + r] = shape
            w[6 + M:6 + M + r] = ss
            w[6 + 2 * M:6 + 2 * M + r] = ds
            tile_job.extend([len(rows)] * ntile)
            rows.append(w)
        self.njobs = len(rows)
        self.ntiles = len(tile_job)
        self.h_jobs = np.ascontiguousarray(np.stack(rows) if rows else np.zeros((0, L.YH_JOB_WORDS), np.int64))
        self.h_tile_job = np.ascontiguousarray(np.array(tile_job, dtype=np.int32))
        self.src_numel, self.dst_numel = src_numel, dst_numel
        # algorithmic element counts (read, written) for roofline accounting
        self.elems_read = int(sum(math.prod(j[2]) for j in jobs if j[0] >= 0))
        self.elems_written = int(sum(math.prod(j[2]) for j in jobs))
        L.call('yastn_hip_check_gather_plan', self.h_jobs.ctypes.data, self.njobs,
               self.h_tile_job.ctypes.data, self.ntiles, src_numel, dst_numel)
        self.d_jobs = torch.from_numpy(self.h_jobs).to(device)
        self.d_tile_job = torch.from_numpy(self.h_tile_job).to(device)

    def args(self):
        return (self.d_jobs.data_ptr(), self.njobs, self.d_tile_job.data_ptr(), self.ntiles)


class GatherPairPlan:
    """Two gather plans in one launch: the jobs of ``pb`` follow those of ``pa`` with the buffer word set
    to 1 (see include/yastn_hip.h), tile indices shifted by ``pa``'s tile count."""

    def __init__(self, pa, pb, device):
        jb = pb.h_jobs.copy()
        jb[:, 5] += pa.ntiles
        jb[:, L.YH_JOB_BUFFER_WORD] = 1
        self.h_jobs = np.ascontiguousarray(np.concatenate([pa.h_jobs, jb]))
        self.h_tile_job = np.ascontiguousarray(np.concatenate([pa.h_tile_job, pb.h_tile_job + pa.njobs]).astype(np.int32))
        self.njobs, self.ntiles = pa.njobs + pb.njobs, pa.ntiles + pb.ntiles
        L.call('yastn_hip_check_gather_plan2', self.h_jobs.ctypes.data, self.njobs, self.h_tile_job.ctypes.data,
               self.ntiles, pa.src_numel, pa.dst_numel, pb.src_numel, pb.dst_numel)
        self.d_jobs = torch.from_numpy(self.h_jobs).to(device)
        self.d_tile_job = torch.from_numpy(self.h_tile_job).to(device)

    def args(self):
        return (self.d_jobs.data_ptr(), self.njobs, self.d_tile_job.data_ptr(), self.ntiles)


def pair_plan(pa, pb, device):
    """The launch of two gather plans together (cached on the pair)."""
    return _CACHE.get(('pair', id(pa), id(pb), _device_key(device)), (pa, pb), lambda: GatherPairPlan(pa, pb, device))


def _uncovered_boxes(Dn, boxes):
    """Boxes of the C-ordered block of shape Dn not covered by any of ``boxes`` ((start, stop) per dim)."""
    r = len(Dn)
    cuts = []
    for d in range(r):
        pts = {0, Dn[d]}
        for b in boxes:
            pts.add(b[d][0])
            pts.add(b[d][1])
        cuts.append(sorted(pts))
    covered = np.zeros(tuple(len(c) - 1 for c in cuts), dtype=bool)
    for b in boxes:
        idx = tuple(slice(cuts[d].index(b[d][0]), cuts[d].index(b[d][1])) for d in range(r))
        covered[idx] = True
    out = []
    for cell in zip(*np.nonzero(~covered)):
        out.append(tuple((cuts[d][c], cuts[d][c + 1]) for d, c in enumerate(cell)))
    return out


def _box_in_block(block_shape, box):
    """(offset, strides, shape) of a sub-box of a C-ordered block."""
    st = _c_strides(block_shape)
    off = sum(b[0] * s for b, s in zip(box, st))
    return off, st, tuple(b[1] - b[0] for b in box)


def _merge_groups(Drsh, P):
    """Group consecutive dims of the permuted source shape P so that group products give Drsh
    (the reference planner fuses consecutive legs after the permutation, _merging.py:355-372)."""
    groups, i = [], 0
    for target in Drsh:
        g, prod = [], 1
        while prod < target:
            if i >= len(P):
                raise ValueError(f"cannot split {P} into {Drsh}")
            g.append(i)
            prod *= P[i]
            i += 1
        if prod != target:
            raise ValueError(f"cannot split {P} into {Drsh}")
        groups.append(g)
    if any(P[j] != 1 for j in range(i, len(P))):
        raise ValueError(f"cannot split {P} into {Drsh}")
    return groups


def _merge_maps(order, Do, Dslc, Drsh, Dn):
    """For one meta_mrg row: destination offset (inside the block) and destination stride of each
    permuted source dim P_j = Do[order[j]], using that Drsh groups consecutive permuted dims."""
    P = tuple(Do[k] for k in order)
    dst_st = _c_strides(Dn)
    off = sum(sl[0] * s for sl, s in zip(Dslc, dst_st))
    groups = _merge_groups(Drsh, P)
    pst = [0] * len(P)
    for g, gstride in zip(groups, dst_st):
        acc = gstride
        for j in reversed(g):
            pst[j] = acc
            acc *= P[j]
    return off, P, tuple(pst)


def _group_meta(meta_new, meta_mrg):
    runs, i = [], 0
    for tn, Dn, sln in meta_new:
        j = i
        while j < len(meta_mrg) and tuple(meta_mrg[j][0]) == tuple(tn):
            j += 1
        runs.append(((tn, tuple(Dn), tuple(sln)), meta_mrg[i:j]))
        i = j
    if i != len(meta_mrg):
        raise ValueError("meta_mrg rows do not follow meta_new")
    return runs


def merge_plan(order, meta_new, meta_mrg, Dsize, src_numel, device):
    """A1 transpose_and_merge: copy jobs for every meta_mrg row plus zero-fill of uncovered boxes."""
    def build():
        jobs = []
        src_c = {}
        for (tn, Dn, sln), rows in _group_meta(meta_new, meta_mrg):
            boxes = []
            for _, slo, Do, Dslc, Drsh in rows:
                Do = tuple(Do)
                off, P, pst = _merge_maps(order, Do, Dslc, Drsh, Dn)
                sst = src_c.setdefault(Do, _c_strides(Do))
                jobs.append((slo[0], sln[0] + off, P, tuple(sst[k] for k in order), pst))
                boxes.append(tuple(tuple(x) for x in Dslc))
            for box in _uncovered_boxes(Dn, boxes):
                off, st, shp = _box_in_block(Dn, box)
                jobs.append((-1, sln[0] + off, shp, (0,) * len(shp), st))
        return GatherPlan(jobs, src_numel, Dsize, device)
    # meta_new is rebuilt by the caller on every call from the same cached structure that produced
    # meta_mrg, so meta_mrg's identity (plus cheap consistency fields) identifies the plan
    return _CACHE.get(('merge', id(meta_mrg), len(meta_new), tuple(meta_new[-1]) if meta_new else None,
                       tuple(order), Dsize, src_numel, _device_key(device)), (meta_mrg,), build)


def _fill_gaps(covered, total):
    """1-D zero-fill jobs for [0, total) minus the union of ``covered`` (start, stop) ranges."""
    jobs, pos = [], 0
    for a, b in sorted(covered):
        if a > pos:
            jobs.append((-1, pos, (a - pos,), (0,), (1,)))
        pos = max(pos, b)
    if pos < total:
        jobs.append((-1, pos, (total - pos,), (0,), (1,)))
    return jobs


def merge_adjoint_plan(order, meta_new, meta_mrg, Dsize, src_numel, device):
    """A6: grad of the source = gather from the merged gradient under the inverse permutation."""
    def build():
        jobs, cov = [], []
        inv = np.argsort(order)
        for (tn, Dn, sln), rows in _group_meta(meta_new, meta_mrg):
            for _, slo, Do, Dslc, Drsh in rows:
                Do = tuple(Do)
                off, P, pst = _merge_maps(order, Do, Dslc, Drsh, Dn)
                # destination = the source block in its own C order; source stride of dim d = pst[j], order[j] = d
                sst = tuple(pst[int(inv[d])] for d in range(len(Do)))
                jobs.append((sln[0] + off, slo[0], Do, sst, _c_strides(Do)))
                cov.append((slo[0], slo[1]))
        jobs.extend(_fill_gaps(cov, src_numel))
        return GatherPlan(jobs, Dsize, src_numel, device)
    return _CACHE.get(('merge_adj', id(meta_mrg), len(meta_new), tuple(meta_new[-1]) if meta_new else None,
                       tuple(order), Dsize, src_numel, _device_key(device)), (meta_mrg,), build)


def unmerge_plan(meta, numel, device):
    """A2 unmerge: newdata[sln] = data[slo].view(Do)[sub_slc]."""
    def build():
        jobs = []
        for sln, Dn, slo, Do, sub in meta:
            off, st, shp = _box_in_block(tuple(Do), sub)
            if shp != tuple(Dn):
                raise ValueError(f"unmerge sub-block {shp} != {Dn}")
            jobs.append((slo[0] + off, sln[0], shp, st, _c_strides(shp)))
        return GatherPlan(jobs, numel, numel, device)
    return _CACHE.get(('unmerge', id(meta), numel, _device_key(device)), (meta,), build)


def unmerge_adjoint_plan(meta, numel, device):
    """A8: grad_src[slo].view(Do)[sub_slc] = grad[sln].view(Dn)."""
    def build():
        jobs, cov = [], []
        for sln, Dn, slo, Do, sub in meta:
            off, st, shp = _box_in_block(tuple(Do), sub)
            jobs.append((sln[0], slo[0] + off, shp, _c_strides(shp), st))
            cov.append((slo[0], slo[1]))
        # the forward covers the whole source (no zero blocks), gaps only if the meta is partial
        jobs.extend(_fill_gaps(_union(cov), numel))
        return GatherPlan(jobs, numel, numel, device)
    return _CACHE.get(('unmerge_adj', id(meta), numel, _device_key(device)), (meta,), build)


def _union(ranges):
    out = []
    for a, b in sorted(set(ranges)):
        if out and a <= out[-1][1]:
            out[-1] = (out[-1][0], max(out[-1][1], b))
        else:
            out.append((a, b))
    return out


def transpose_plan(axes, meta, numel, device):
    """backend.transpose: newdata[sln].view(Dn) = data[slo].view(Do).permute(axes)."""
    def build():
        jobs = []
        for sln, Dn, slo, Do in meta:
            st = _c_strides(tuple(Do))
            jobs.append((slo[0], sln[0], tuple(Dn), tuple(st[a] for a in axes), _c_strides(tuple(Dn))))
        return GatherPlan(jobs, numel, numel, device)
    return _CACHE.get(('transpose', id(meta), tuple(axes), numel, _device_key(device)), (meta,), build)


def _runs(idx):
    """Split a sorted-or-not index list into runs of consecutive values: (position, start, length)."""
    out = []
    idx = [int(x) for x in idx]
    i = 0
    while i < len(idx):
        j = i
        while j + 1 < len(idx) and idx[j + 1] == idx[j] + 1:
            j += 1
        out.append((i, idx[i], j - i + 1))
        i = j + 1
    return out


def _axis_box(shape, axis, start, length):
    return tuple((start, start + length) if d == axis else (0, n) for d, n in enumerate(shape))


def mask_plan(mask_host, meta, Dsize, src_numel, axis, device, embed=False, mask_key=None):
    """A12 apply_mask (embed=False) / embed_mask (embed=True) as box-gather jobs over index runs.

    apply: out[sln].view(Dn) = A[sla].view(Da)[.., idx, ..]  (Dn has len(idx) on ``axis``)
    embed: out[sln].view(Dn)[.., idx, ..] = A[sla].view(Da), zero elsewhere.
    """
    def build():
        jobs = []
        for sln, Dn, sla, Da, tm in meta:
            Dn, Da = _as_shape(Dn), _as_shape(Da)
            idx = mask_host[tuple(tm) if isinstance(tm, (list, tuple)) else tm]
            if isinstance(idx, slice):  # resolved against the masked dimension of this block
                idx = range(*idx.indices((Dn if embed else Da)[axis]))
            runs = _runs(idx)
            if not embed:
                for pos, start, ln in runs:
                    so, sst, shp = _box_in_block(Da, _axis_box(Da, axis, start, ln))
                    do, dst, _ = _box_in_block(Dn, _axis_box(Dn, axis, pos, ln))
                    jobs.append((sla[0] + so, sln[0] + do, shp, sst, dst))
            else:
                boxes = []
                for pos, start, ln in runs:
                    so, sst, shp = _box_in_block(Da, _axis_box(Da, axis, pos, ln))
                    box = _axis_box(Dn, axis, start, ln)
                    do, dst, _ = _box_in_block(Dn, box)
                    jobs.append((sla[0] + so, sln[0] + do, shp, sst, dst))
                    boxes.append(box)
                for box in _uncovered_boxes(Dn, boxes):
                    off, st, shp = _box_in_block(Dn, box)
                    jobs.append((-1, sln[0] + off, shp, (0,) * len(shp), st))
        if embed:
            cov = [tuple(m[0]) for m in meta]
            jobs.extend(_fill_gaps(cov, Dsize))
        return GatherPlan(jobs, src_numel, Dsize, device)
    return _CACHE.get(('mask', embed, id(meta), mask_key, axis, Dsize, src_numel, _device_key(device)),
                      (meta,), build)


def _as_shape(D):
    return (D,) if isinstance(D, (int, np.integer)) else tuple(D)


# --------------------------------------------------------------------------- grouped GEMM

class GemmPlan:
    """Device work list of the grouped GEMM (see include/yastn_hip.h)."""

    def __init__(self, gemms, terms, in_numel, out_numel, device, cfg=None):
        g0 = np.array(gemms, dtype=np.int64).reshape(-1, L.YH_GEMM_WORDS)
        t0 = np.array(terms, dtype=np.int64).reshape(-1, L.YH_TERM_WORDS)
        # useful flops (real MACs x2) and coverage of the plan, independent of tiling and splitting
        k0 = np.array([int(t0[r[5]:r[5] + r[6], 6].sum()) if r[6] else 0 for r in g0], dtype=np.int64)
        self.macs = int(np.sum(g0[:, 0] * g0[:, 1] * k0)) if len(g0) else 0
        self.covered = int(np.sum(g0[:, 0] * g0[:, 1])) if len(g0) else 0
        self.accumulates = bool(np.any(g0[:, 7] & 1)) if len(g0) else False
        out_numel = list(out_numel)
        g0, t0, self.h_red, self.part_numel = _split_k(g0, t0, len(out_numel))
        self.part_buf = len(out_numel) if self.part_numel else None
        if self.part_numel:
            out_numel.append(self.part_numel)
            L.call('yastn_hip_check_splitk_plan', self.h_red.ctypes.data, len(self.h_red),
                   np.array(out_numel + [0] * (L.YH_MAX_BUFS - len(out_numel)), dtype=np.int64).ctypes.data,
                   len(out_numel), self.part_numel)
            self.d_red = torch.from_numpy(self.h_red).to(device)
        self.h_gemms = np.ascontiguousarray(g0)
        self.h_terms = np.ascontiguousarray(t0)
        self.ngemms, self.nterms = len(self.h_gemms), len(self.h_terms)
        self.in_numel = np.array(list(in_numel) + [0] * (L.YH_MAX_BUFS - len(in_numel)), dtype=np.int64)
        self.out_numel = np.array(out_numel + [0] * (L.YH_MAX_BUFS - len(out_numel)), dtype=np.int64)
        self.nin, self.nout = len(in_numel), len(out_numel)
        k_of = np.zeros(self.ngemms, dtype=np.int64)
        for g, row in enumerate(self.h_gemms):
            k_of[g] = int(self.h_terms[row[5]:row[5] + row[6], 6].sum()) if row[6] else 0
        self.k_of = k_of
        self.device = device
        self._tiles = {}
        self.cfg = choose_tile(self) if cfg is None else cfg
        self.d_gemms = torch.from_numpy(self.h_gemms).to(device)
        self.d_terms = torch.from_numpy(self.h_terms).to(device)
        self.set_cfg(self.cfg)

    def tiles_for(self, cfg):
        if cfg not in self._tiles:
            bm, bn = L.TILE_DIMS[cfg]
            tl = []
            for g, row in enumerate(self.h_gemms):
                M, N = int(row[0]), int(row[1])
                for m0 in range(0, M, bm):
                    for n0 in range(0, N, bn):
                        tl.append((int(self.k_of[g]), g, m0, n0))
            h = np.ascontiguousarray(np.array(_xcd_order(tl), dtype=np.int32).reshape(-1, 3))
            L.call('yastn_hip_check_gemm_plan', self.h_gemms.ctypes.data, self.ngemms, self.h_terms.ctypes.data,
                   self.nterms, h.ctypes.data, len(h), cfg, self.in_numel.ctypes.data, self.nin,
                   self.out_numel.ctypes.data, self.nout)
            self._tiles[cfg] = (h, torch.from_numpy(h).to(self.device))
        return self._tiles[cfg]

    def applicable(self, cfg):
        """The streaming small-M kernel takes only plans whose GEMMs all have M <= 16."""
        if cfg == L.YH_TILE_SMALLM:
            return self.ngemms == 0 or int(self.h_gemms[:, 0].max()) <= L.TILE_DIMS[cfg][0]
        return True

    def set_cfg(self, cfg):
        self.cfg = cfg
        self.h_tiles, self.d_tiles = self.tiles_for(cfg)
        self.ntiles = len(self.h_tiles)

    def args(self):
        return (self.d_gemms.data_ptr(), self.ngemms, self.d_terms.data_ptr(), self.nterms,
                self.d_tiles.data_ptr(), self.ntiles, self.cfg)

    def histogram_key(self):
        return tuple(sorted(zip(self.h_gemms[:, 0].tolist(), self.h_gemms[:, 1].tolist(), self.k_of.tolist())))

    def operand_traffic(self, cfg):
        """Operand elements the tiles of shape ``cfg`` request (A's row panel once per column tile, B's
        column panel once per row tile, C once): the L2 / HBM reads a tile shape causes beyond the
        algorithmic M K + K N + M N."""
        bm, bn = L.TILE_DIMS[cfg]
        M = self.h_gemms[:, 0].astype(np.float64)
        N = self.h_gemms[:, 1].astype(np.float64)
        K = self.k_of.astype(np.float64)
        return float(np.sum(M * K * np.ceil(N / bn) + K * N * np.ceil(M / bm) + M * N))

    def pick_tile(self, times, near=0.05):
        """Of the timed shapes ({cfg: ms}) within ``near`` of the fastest, the one with the least
        operand traffic."""
        fastest = min(times.values())
        return min((c for c in times if times[c] <= fastest * (1 + near)), key=self.operand_traffic)

    def autotune(self, run, reps=3, batches=3, near=0.05):
        """Pick a tile shape for this plan by timing ``run()`` (one launch of the plan) with every shape:
        each shape's time is the fastest of ``batches`` event-timed batches of ``reps`` launches (a single
        batch let a stray delay decide between shapes a few percent apart); of the shapes within ``near``
        (relative) of the fastest, the one whose tiles request the fewest operand elements
        (``operand_traffic``), so that near-ties go to the shape with fewer re-reads.  The choice is
        remembered per block-size histogram, so plans of other operands with the same sector shapes reuse
        it."""
        if getattr(self, 'tuned', False):
            return
        key = self.histogram_key()
        if key in _TUNED:
            self.set_cfg(_TUNED[key])
            self.tuned = True
            return
        times = {}
        for cfg in L.TILE_DIMS:
            if not self.applicable(cfg):
                continue
            self.set_cfg(cfg)
            run()
            best_t = None
            for _ in range(batches):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    run()
                e1.record()
                e1.synchronize()
                t = e0.elapsed_time(e1) / reps
                best_t = t if best_t is None else min(best_t, t)
            times[cfg] = best_t
        best = self.pick_tile(times, near)
        self.set_cfg(best)
        self.tune_times = times
        self.tuned = True
        _TUNED[key] = best


_TUNED = {}
NUM_CU = 256
NUM_XCD = 8  # workgroup i of a launch runs on XCD i % 8, each XCD with its own L2


def _xcd_order(tl):
    """Dispatch order of the tiles (k, g, m0, n0) of one launch: the tiles of one row panel (g, m0)
    share their A rows, so panels are dealt to the XCDs as units (largest work first, to the least
    loaded XCD), and position i of the list holds a tile of XCD i % NUM_XCD; each XCD's tiles go
    longest K first.  An XCD that runs out takes tiles from the fullest one."""
    units = {}
    for t in tl:
        units.setdefault((t[1], t[2]), []).append(t)
    load = [0] * NUM_XCD
    queues = [[] for _ in range(NUM_XCD)]
    for key, ts in sorted(units.items(), key=lambda kv: -sum(t[0] + 1 for t in kv[1])):
        x = min(range(NUM_XCD), key=lambda i: load[i])
        queues[x].extend(ts)
        load[x] += sum(t[0] + 1 for t in ts)
    for q in queues:
        q.sort(key=lambda t: (-t[0], t[1], t[2], t[3]))
        q.reverse()  # pop() from the end takes the longest first
    out = []
    while any(queues):
        for x in range(NUM_XCD):
            q = queues[x] if queues[x] else max(queues, key=len)
            if q:
                out.append(q.pop()[1:])
    return out
# split the K of a single-term GEMM beyond this length, when its output has at most SPLITK_MAX_MN
# elements, into chunks of at least SPLITK_MIN_KC (tuning knobs)
SPLITK_MIN_K = int(os.environ.get('YASTN_HIP_SPLITK_MIN_K', '8192'))
SPLITK_MAX_MN = int(os.environ.get('YASTN_HIP_SPLITK_MAX_MN', str(64 * 64)))
SPLITK_MIN_KC = int(os.environ.get('YASTN_HIP_SPLITK_MIN_KC', '2048'))
SPLITK_MAX_S = 512


def _split_k(g0, t0, out_buf):
    """Cut long-K single-term GEMMs with small outputs into S chunks writing partial products to the
    buffer `out_buf` (row-major M x N per chunk); returns the new gemm / term lists, the reduction
    records and the partials' size.  Other GEMMs are kept as they are."""
    gl, tl, red, part = [], [], [], 0
    for row in g0:
        M, N, c_off, c_rs, c_cs, tb, tc, fl = (int(x) for x in row)
        if tc == 1 and out_buf < L.YH_MAX_BUFS and M * N <= SPLITK_MAX_MN and int(t0[tb, 6]) >= SPLITK_MIN_K:
            a_off, a_rs, a_cs, b_off, b_rs, b_cs, K, tf = (int(x) for x in t0[tb])
            kc = max(SPLITK_MIN_KC, -(-K // SPLITK_MAX_S))
            kc = -(-kc // 16) * 16
            S = -(-K // kc)
            for sidx in range(S):
                ks = min(kc, K - sidx * kc)
                gl.append((M, N, part + sidx * M * N, N, 1, len(tl), 1, out_buf << 8))
                tl.append((a_off + sidx * kc * a_cs, a_rs, a_cs, b_off + sidx * kc * b_rs, b_rs, b_cs, ks, tf))
            red.append((M, N, c_off, c_rs, c_cs, part, S, fl))
            part += S * M * N
        else:
            gl.append((M, N, c_off, c_rs, c_cs, len(tl), tc, fl))
            tl.extend(tuple(int(x) for x in t) for t in t0[tb:tb + tc])
    g = np.array(gl, dtype=np.int64).reshape(-1, L.YH_GEMM_WORDS)
    t = np.array(tl, dtype=np.int64).reshape(-1, L.YH_TERM_WORDS)
    r = np.ascontiguousarray(np.array(red, dtype=np.int64).reshape(-1, L.YH_SPLITK_WORDS))
    return g, t, r, part


def choose_tile(plan):
    """Tile choice from the plan's block-size histogram.

    Estimated makespan = max(longest tile, padded MACs / CUs) / relative tile efficiency: tiles of one
    launch spread over the CUs, and a larger tile reuses operands better (higher efficiency) but pads
    small sectors and exposes fewer tiles.  Skinny tiles serve sectors with M or N <= 16.
    """
    if plan.ngemms == 0:
        return L.YH_TILE_64x64
    M = plan.h_gemms[:, 0].astype(np.float64)
    N = plan.h_gemms[:, 1].astype(np.float64)
    K = np.maximum(plan.k_of.astype(np.float64), 1.0)
    if plan.applicable(L.YH_TILE_SMALLM) and M.max() <= 8:
        return L.YH_TILE_SMALLM  # a few rows: streaming beats any MFMA tile, whose rows would be padding
    eff = {L.YH_TILE_32x32: 0.35, L.YH_TILE_64x64: 0.8, L.YH_TILE_128x64: 1.0, L.YH_TILE_16x64: 0.3,
           L.YH_TILE_64x16: 0.3, L.YH_TILE_128x128: 1.1, L.YH_TILE_SMALLM: 0.3}
    best, best_cost = None, None
    for cfg, (bm, bn) in L.TILE_DIMS.items():
        if not plan.applicable(cfg) or cfg == L.YH_TILE_RD32:  # the register-direct tile: by timing only
            continue
        tiles = np.ceil(M / bm) * np.ceil(N / bn)
        work = bm * bn * K
        padded = float(np.sum(tiles * work))
        cost = max(float(work.max()), padded / NUM_CU) / eff[cfg]
        if best_cost is None or cost < best_cost * 0.999:
            best, best_cost = cfg, cost
    return best


def dot_plan(meta_dot, a_numel, b_numel, Dsize, device):
    """A3 dot: one GEMM (one term) per output block."""
    def build():
        gemms, terms = [], []
        for slc, Dc, sla, Da, slb, Db in meta_dot:
            M, K = Da
            K2, N = Db
            if K != K2 or tuple(Dc) != (M, N):
                raise ValueError(f"dot block shapes {Da} x {Db} -> {Dc}")
            gemms.append((M, N, slc[0], N, 1, len(terms), 1, 0))
            terms.append((sla[0], K, 1, slb[0], N, 1, K, (0 << 8) | (1 << 16)))
        return GemmPlan(gemms, terms, (a_numel, b_numel), (Dsize,), device)
    return _CACHE.get(('dot', id(meta_dot), a_numel, b_numel, Dsize, _device_key(device)), (meta_dot,), build)


def dot_adjoint_plan(meta_dot, a_numel, b_numel, c_numel, device):
    """A7: gA[sla] = sum gC[slc] B[slb]^H,  gB[slb] = sum A[sla]^H gC[slc]; inputs (A, B, gC), outputs (gA, gB).
    Returns the plan and whether gA / gB have regions that no GEMM writes (then they start from zeros)."""
    def build():
        by_a, by_b = {}, {}
        for slc, Dc, sla, Da, slb, Db in meta_dot:
            by_a.setdefault((tuple(sla), tuple(Da)), []).append((slc, Dc, slb, Db))
            by_b.setdefault((tuple(slb), tuple(Db)), []).append((slc, Dc, sla, Da))
        gemms, terms = [], []
        for (sla, Da), lst in by_a.items():
            M, K = Da
            gemms.append((M, K, sla[0], K, 1, len(terms), len(lst), 0 << 8))
            for slc, Dc, slb, Db in lst:
                N = Dc[1]
                # op(A) = gC (M x N), op(B)[n][k] = conj(B[k][n])  (B stored K x N)
                terms.append((slc[0], N, 1, slb[0], 1, N, N, 2 | (2 << 8) | (1 << 16)))
        for (slb, Db), lst in by_b.items():
            K, N = Db
            gemms.append((K, N, slb[0], N, 1, len(terms), len(lst), 1 << 8))
            for slc, Dc, sla, Da in lst:
                M = Da[0]
                # op(A)[k][m] = conj(A[m][k])  (A stored M x K), op(B) = gC (M x N)
                terms.append((sla[0], 1, K, slc[0], N, 1, M, 1 | (0 << 8) | (2 << 16)))
        plan = GemmPlan(gemms, terms, (a_numel, b_numel, c_numel), (a_numel, b_numel), device)
        cov_a = sum(math.prod(Da) for (_, Da) in by_a)
        cov_b = sum(math.prod(Db) for (_, Db) in by_b)
        plan.zero_a = cov_a != a_numel
        plan.zero_b = cov_b != b_numel
        return plan
    return _CACHE.get(('dot_adj', id(meta_dot), a_numel, b_numel, c_numel, _device_key(device)), (meta_dot,), build)


def gemm_list_plan(gemms, terms, in_numel, out_numel, device, key=None, refs=()):
    if key is None:
        return GemmPlan(gemms, terms, in_numel, out_numel, device)
    return _CACHE.get(key + (_device_key(device),), refs, lambda: GemmPlan(gemms, terms, in_numel, out_numel, device))


# --------------------------------------------------------------------------- Jacobi SVD / eigh

def _circle_pair(ne, r, k):
    n1 = ne - 1
    if k == 0:
        i, j = r, n1
    else:
        i, j = (r + k) % n1, (r - k + n1) % n1
    return (i, j) if i < j else (j, i)


PRECOND = True  # QR-precondition the panel path of SVD blocks (off: plain one-sided Jacobi on W0)
# SVD blocks with more columns than this take the QR-preconditioned panel path even where they fit the
# LDS kernel, whose unpreconditioned sweeps on one workgroup outlast the panel rounds past this size.
SVD_LDS_MAX_Q = int(os.environ.get('YASTN_HIP_SVD_LDS_MAX_Q', '16'))
# Panel width b of the panel Jacobi (a workgroup rotates 2b columns per round): smaller panels give more
# workgroups per round but more rounds per sweep.  Tuning knob; the LDS limit may lower it further.
PANEL_WIDTH = int(os.environ.get('YASTN_HIP_PANEL_WIDTH', '16'))
PANEL_CHUNK_ROWS = 64  # rows per workgroup of the panel Gram kernel
# workgroups per panel pair in the inner kernel (0: about one 16-row stripe of W and V per wave); tuning knob
PANEL_SLICES = max(0, min(64, int(os.environ.get('YASTN_HIP_PANEL_SLICES', '0'))))


class SvdPlan:
    """Block list of the batched Jacobi SVD / eigh.

    Blocks whose working matrices (W and V, (p q + q^2) elements) fit the LDS limit run in the
    LDS-resident kernel, one workgroup per block, all in one launch.  Larger blocks use the panel
    path: W and V in a workspace, panels of width b, one launch per round of panel pairs.  For SVD
    blocks the panel path first factors W0 = Q R (blocked Householder QR: a panel kernel plus three
    grouped GEMMs per step; the p x YH_QR_NB panel is staged in LDS when every block's fits, else
    factored in place in the workspace) and iterates on X = R^H; the p-side vectors are multiplied
    by Q at the end (three grouped GEMMs per step).
    """

    def __init__(self, blocks, itemsize, mode, a_numel, u_numel, s_numel, v_numel, device):
        # blocks: list of (a_off, m, n, u_off, s_off, v_off)
        NB = L.YH_QR_NB
        words, small, big, pre = [], [], [], []
        work = 0
        lds_bytes = 0
        dev_lds = L.device_lds_per_block()
        # the LDS kernel also holds ~8 KiB of static shared memory
        lds_limit = min(L.YH_SVD_LDS_BYTES, dev_lds - 9 * 1024)
        qd = np.zeros((len(blocks), L.YH_QR_WORDS), dtype=np.int64)
        jwords = []
        for b, (a_off, m, n, u_off, s_off, v_off) in enumerate(blocks):
            tr = mode == 0 and m < n
            p, q = (n, m) if tr else (m, n)
            need = p * q + q * q
            jw = None
            if need * itemsize <= lds_limit and not (mode == 0 and PRECOND and q > SVD_LDS_MAX_Q):
                small.append(b)
                lds_bytes = max(lds_bytes, need * itemsize)
                w_off = 0
            elif mode == 0 and PRECOND:
                big.append(b)
                pre.append(b)
                nst = -(-q // NB)
                w_off = work
                x_off = work + p * q
                vall_off = x_off + 2 * q * q
                tm_off = vall_off + p * q
                y_off = tm_off + nst * NB * NB
                work = y_off + 2 * NB * q
                qd[b] = (p, q, w_off, vall_off, tm_off, x_off, y_off, 0)
                jw = (a_off, q, q, u_off, s_off, v_off, x_off, 0)
            else:
                big.append(b)
                w_off = work
                work += need
            words.append((a_off, m, n, u_off, s_off, v_off, w_off, 0))
            jwords.append(jw if jw is not None else words[-1])
        size = lambda b: blocks[b][1] * blocks[b][2] * min(blocks[b][1], blocks[b][2])
        small.sort(key=lambda b: -size(b))
        big.sort(key=lambda b: -size(b))
        self.h_blocks = np.ascontiguousarray(np.array(words, dtype=np.int64).reshape(-1, L.YH_SVD_WORDS))
        self.h_order = np.ascontiguousarray(np.array(small + big, dtype=np.int32))
        self.nblocks, self.n_lds, self.lds_bytes, self.work_numel = len(words), len(small), lds_bytes, work
        self.flops = sum(4 * m * n * min(m, n) + 8 * min(m, n) ** 3 for (_, m, n, *_r) in blocks)
        dt = L.YH_C128 if itemsize == 16 else L.YH_F64
        self.dt = dt
        L.call('yastn_hip_check_svd_plan', self.h_blocks.ctypes.data, self.nblocks, self.h_order.ctypes.data,
               self.n_lds, self.lds_bytes, mode, dt, a_numel, u_numel, s_numel, v_numel, max(work, 0))
        self.d_blocks = torch.from_numpy(self.h_blocks).to(device)
        self.d_order = torch.from_numpy(self.h_order).to(device)
        # panel path
        self.n_big = len(big)
        self.n_pre = len(pre)
        self.rounds = []
        if big:
            self.h_jblocks = np.ascontiguousarray(np.array(jwords, dtype=np.int64).reshape(-1, L.YH_SVD_WORDS))
            if pre:
                # the Jacobi stage of the preconditioned blocks works on X (q x q) in its own workspace
                L.call('yastn_hip_check_svd_plan', self.h_jblocks.ctypes.data, self.nblocks,
                       self.h_order.ctypes.data, 0, 0, mode, dt, a_numel, u_numel, s_numel, v_numel, max(work, 0))
            self.d_jblocks = torch.from_numpy(self.h_jblocks).to(device)
            limit = dev_lds - 2048
            width = max(1, min(PANEL_WIDTH, 32))
            # LDS of the inner sweep of a panel pair: its Gram matrix and the rotation accumulator,
            # 2b x (2b + 1) each; the columns themselves stay in the workspace
            while width > 1 and 2 * (2 * width) * (2 * width + 1) * itemsize > limit:
                width //= 2
            self.panel_width = width
            pmax = max(max(jwords[b][1], jwords[b][2]) for b in big)
            qmax = max(min(jwords[b][1], jwords[b][2]) for b in big)
            self.maxch_p = -(-pmax // PANEL_CHUNK_ROWS)
            # a sweep: every panel against itself (round -1), then the circle order of panel pairs
            per_round = {}
            for b in big:
                m, n = blocks[b][1], blocks[b][2]
                q = min(m, n) if mode == 0 else n
                npan = -(-q // width)
                ne = npan + (npan & 1)
                for i in range(npan):
                    per_round.setdefault(-1, []).append((b, i, i))
                for r in range(ne - 1):
                    for k in range(ne // 2):
                        i, j = _circle_pair(ne, r, k)
                        if j < npan:
                            per_round.setdefault(r, []).append((b, i, j))
            jobs, offs = [], []
            for r in sorted(per_round):
                offs.append((len(jobs), len(per_round[r])))
                jobs.extend(per_round[r])
            self.rounds = offs
            # workgroups per panel pair in the inner kernel: each sums the Gram partials and repeats the
            # inner sweep, then updates its slice of the 16-row stripes of W and V.  At most about one
            # stripe per wave (16 waves); and about one workgroup per CU over the largest round of cross
            # pairs, since every further slice repeats the sequential inner sweep (measured: rounds of
            # ~260 pairs of up to 1134-row complex blocks run 2x faster with 1 slice than with 9, rounds
            # of ~75 pairs of up to 440-row blocks fastest with 3).
            per_wave = -(-(-(-pmax // 16) - (-qmax // 16)) // 16)
            maxn = max(n for _, n in offs)
            cross = max((n for _, n in offs[1:]), default=maxn)
            self.panel_slices = PANEL_SLICES or max(1, min(64, per_wave, L.device_cu_count() // cross))
            self.h_rounds = np.ascontiguousarray(np.array(offs, dtype=np.int64).reshape(-1, 2))
            self.d_rounds = torch.from_numpy(self.h_rounds).to(device)
            self.max_round_jobs = maxn
            self.h_jobs = np.ascontiguousarray(np.array(jobs, dtype=np.int32).reshape(-1, 3))
            L.call('yastn_hip_check_panel_jobs', self.h_jblocks.ctypes.data, self.nblocks, self.h_jobs.ctypes.data,
                   len(jobs), width, mode, self.maxch_p)
            t16 = -(-2 * width // 16)
            self.part_numel = maxn * self.maxch_p * (t16 * (t16 + 1) // 2) * 256 * (itemsize // 8)
            self.d_jobs = torch.from_numpy(self.h_jobs).to(device)
            self.h_big = np.ascontiguousarray(np.array(big, dtype=np.int32))
            self.d_big = torch.from_numpy(self.h_big).to(device)
            plain = [b for b in big if b not in set(pre)]
            self.n_plain = len(plain)
            self.d_plain = torch.from_numpy(np.array(plain or [0], dtype=np.int32)).to(device)
            self.njobs = len(jobs)
        if pre:
            self._qr_plans(blocks, pre, qd, itemsize, work, u_numel, v_numel, device)

    def _qr_plans(self, blocks, pre, qd, itemsize, work, u_numel, v_numel, device):
        """Panel lists and grouped-GEMM plans of the Householder QR (trailing updates) and of the
        final multiplication of the p-side vectors by Q (buffers: 0 workspace, 1 U, 2 Vh)."""
        NB = L.YH_QR_NB
        self.h_qdesc = np.ascontiguousarray(qd)
        pmax = max(int(qd[b, 0]) for b in pre)
        # panel staged in LDS, or 0: factored in place in the workspace (the panel kernel's static LDS,
        # the T matrices, takes ~17.5 KiB at 32 columns)
        self.qr_lds = pmax * NB * itemsize
        if self.qr_lds > L.device_lds_per_block() - 20 * 1024:
            self.qr_lds = 0
        ids = np.array(pre, dtype=np.int32)
        L.call('yastn_hip_check_qr_plan', self.h_qdesc.ctypes.data, self.h_blocks.ctypes.data, ids.ctypes.data,
               len(ids), self.dt, max(work, 0), self.qr_lds)
        self.d_qdesc = torch.from_numpy(self.h_qdesc).to(device)
        self.d_pre = torch.from_numpy(ids).to(device)
        nsteps = max(-(-int(qd[b, 1]) // NB) for b in pre)
        bufs = (max(work, 1), max(u_numel, 1), max(v_numel, 1))
        self.qr_steps = []   # (d_ids, n, trailing plans)
        self.q_apply = []    # per step (reverse order): plans H1, H2, H3
        for s in range(nsteps):
            k0 = s * NB
            act = [b for b in pre if int(qd[b, 1]) > k0]
            g1, t1, g2, t2, g3, t3 = [], [], [], [], [], []
            h1, u1, h2, u2, h3, u3 = [], [], [], [], [], []
            for b in act:
                p, q, w_off, vall_off, tm_off, x_off, y_off, _ = (int(x) for x in qd[b])
                m, n = blocks[b][1], blocks[b][2]
                tr = m < n
                nbs = min(NB, q - k0)
                Lr = p - k0
                tm = tm_off + s * NB * NB
                y2 = y_off + NB * q
                rest = q - k0 - nbs
                v0 = vall_off + k0 + k0 * p
                if rest > 0:
                    # Y = V^H C ; Y2 = Tm^H Y ; C += V Y2   (Tm = -T: C <- (I - V T^H V^H) C)
                    g1.append((nbs, rest, y_off, q, 1, len(t1), 1, 0))
                    t1.append((v0, p, 1, w_off + k0 + (k0 + nbs) * p, 1, p, Lr, 1))
                    g2.append((nbs, rest, y2, q, 1, len(t2), 1, 0))
                    t2.append((tm, NB, 1, y_off, q, 1, nbs, 1))
                    g3.append((Lr, rest, w_off + k0 + (k0 + nbs) * p, 1, p, len(t3), 1, 1))
                    t3.append((v0, 1, p, y2, q, 1, nbs, 0))
                # Q application on the p-side vectors M (U for m >= n; conj(M) stored in Vh rows for m < n)
                if not tr:
                    mb, m_off, m_rs, m_cs, cj = 1, blocks[b][3], q, 1, 0
                else:
                    mb, m_off, m_rs, m_cs, cj = 2, blocks[b][5], 1, p, 1
                h1.append((nbs, q, y_off, q, 1, len(u1), 1, 0))
                u1.append((v0, p, 1, m_off + k0 * m_rs, m_rs, m_cs, Lr, (1 - cj) | (0 << 8) | (mb << 16)))
                h2.append((nbs, q, y2, q, 1, len(u2), 1, 0))
                u2.append((tm, 1, NB, y_off, q, 1, nbs, cj))
                h3.append((Lr, q, m_off + k0 * m_rs, m_rs, m_cs, len(u3), 1, 1 | (mb << 8)))
                u3.append((v0, 1, p, y2, q, 1, nbs, cj))
            d_ids = torch.from_numpy(np.array(act, dtype=np.int32)).to(device)
            trail = None
            if g1:
                trail = (GemmPlan(g1, t1, bufs[:1], bufs[:1], device), GemmPlan(g2, t2, bufs[:1], bufs[:1], device),
                         GemmPlan(g3, t3, bufs[:1], bufs[:1], device))
            self.qr_steps.append((d_ids, len(act), trail))
            self.q_apply.append((GemmPlan(h1, u1, bufs, bufs[:1], device), GemmPlan(h2, u2, bufs[:1], bufs[:1], device),
                                 GemmPlan(h3, u3, bufs[:1], bufs, device)))
        self.q_apply.reverse()


class QrPlan:
    """Batched Householder QR of the blocks of backend.qr (meta rows (sl, (m, n), slQ, (m, k), slR,
    (k, n))): per block the workspace holds W0 (m x n), the reflectors (m x k), -T per panel, E (m x k)
    and the GEMM scratch; the panel steps and the GEMM plans follow the preconditioning of the SVD
    (trailing updates of W0 after each panel, then E <- H_s E for the panels in reverse).  The
    32-column Householder panels are staged in LDS when every block's fits (``lds`` > 0), else
    factored in place in the workspace (``lds`` = 0)."""

    def __init__(self, meta, itemsize, a_numel, sizes, device):
        NB = L.YH_QR_NB
        self.dt = L.YH_C128 if itemsize == 16 else L.YH_F64
        qd, qb, work = [], [], 0
        for sl, D, slQ, DQ, slR, DR in meta:
            m, n = int(D[0]), int(D[1])
            k = min(m, n)
            if tuple(DQ) != (m, k) or tuple(DR) != (k, n):
                raise ValueError(f"qr meta shapes {D} {DQ} {DR}")
            npan = -(-k // NB)
            offs = []
            for size in (m * n, m * k, npan * NB * NB, m * k, 2 * NB * n):
                offs.append(work)
                work += size
            qd.append((m, n, *offs, k))
            qb.append((int(sl[0]), int(slQ[0]), int(slR[0])))
        self.nblocks = len(qd)
        self.work_numel = max(work, 1)
        self.h_qdesc = np.ascontiguousarray(np.array(qd, dtype=np.int64).reshape(-1, L.YH_QR_WORDS))
        self.h_qblocks = np.ascontiguousarray(np.array(qb, dtype=np.int64).reshape(-1, L.YH_QROP_WORDS))
        pmax = max((r[0] for r in qd), default=1)
        self.lds = pmax * NB * itemsize
        if self.lds > L.device_lds_per_block() - 20 * 1024:  # static LDS of the panel kernel: ~17.5 KiB
            self.lds = 0
        self.native = True
        L.call('yastn_hip_check_qr_op_plan', self.h_qdesc.ctypes.data, self.h_qblocks.ctypes.data, self.nblocks,
               self.dt, a_numel, sizes[0], sizes[1], self.work_numel, self.lds)
        self.d_qdesc = torch.from_numpy(self.h_qdesc).to(device)
        self.d_qblocks = torch.from_numpy(self.h_qblocks).to(device)
        nsteps = max((-(-r[7] // NB) for r in qd), default=0)
        bufs = (self.work_numel,)
        self.steps, self.q_apply = [], []
        for s in range(nsteps):
            k0 = s * NB
            act = [b for b in range(self.nblocks) if qd[b][7] > k0]
            g1, t1, g2, t2, g3, t3 = [], [], [], [], [], []
            h1, u1, h2, u2, h3, u3 = [], [], [], [], [], []
            for b in act:
                m, n, w_off, vall_off, tm_off, e_off, y_off, k = qd[b]
                nbs = min(NB, k - k0)
                Lr = m - k0
                tm = tm_off + s * NB * NB
                y2 = y_off + NB * n
                rest = n - k0 - nbs
                v0 = vall_off + k0 + k0 * m  # reflectors: m x k column-major
                if rest > 0:  # W0(:, right of the panel) <- (I - V T^H V^H) W0(...), Tm = -T
                    g1.append((nbs, rest, y_off, n, 1, len(t1), 1, 0))
                    t1.append((v0, m, 1, w_off + k0 + (k0 + nbs) * m, 1, m, Lr, 1))
                    g2.append((nbs, rest, y2, n, 1, len(t2), 1, 0))
                    t2.append((tm, NB, 1, y_off, n, 1, nbs, 1))
                    g3.append((Lr, rest, w_off + k0 + (k0 + nbs) * m, 1, m, len(t3), 1, 1))
                    t3.append((v0, 1, m, y2, n, 1, nbs, 0))
                # E(k0:, :) <- (I - V T V^H) E(k0:, :)
                h1.append((nbs, k, y_off, n, 1, len(u1), 1, 0))
                u1.append((v0, m, 1, e_off + k0, 1, m, Lr, 1))
                h2.append((nbs, k, y2, n, 1, len(u2), 1, 0))
                u2.append((tm, 1, NB, y_off, n, 1, nbs, 0))
                h3.append((Lr, k, e_off + k0, 1, m, len(u3), 1, 1))
                u3.append((v0, 1, m, y2, n, 1, nbs, 0))
            d_ids = torch.from_numpy(np.array(act, dtype=np.int32)).to(device)
            trail = None
            if g1:
                trail = tuple(GemmPlan(g, tt, bufs, bufs, device) for g, tt in ((g1, t1), (g2, t2), (g3, t3)))
            self.steps.append((d_ids, len(act), trail))
            self.q_apply.append(tuple(GemmPlan(g, tt, bufs, bufs, device) for g, tt in ((h1, u1), (h2, u2),
                                                                                           (h3, u3))))
        self.q_apply.reverse()


class QrAdjointPlan:
    """Grouped-GEMM plans and words of the QR adjoint (see yastn_hip_qr_adjoint_core), one plan per
    product, each with its own (outputs; inputs):
      gq: gQ' += A(:, k:) gR(:, k:)^H   (gQ'; A, gR)     only blocks with m < n
      gy: gA(:, k:) = Q gR(:, k:)        (gA; Q, gR)      only blocks with m < n
      m1: M1 = R1 gR1^H                  (M1; R, gR)
      m2: M2 = gQ'^H Q                   (M2; gQ', Q)
      b:  B += Q C, B starting as gQ'    (B; Q, C)"""

    def __init__(self, meta, itemsize, a_numel, sizes, device):
        self.dt = L.YH_C128 if itemsize == 16 else L.YH_F64
        words, mo = [], 0
        gq, tq, gy, ty, gm1, tm1, gm2, tm2, gb, tb = ([] for _ in range(10))
        self.max_rows = self.max_k = 0
        for sl, D, slQ, DQ, slR, DR in meta:
            m, n = int(D[0]), int(D[1])
            k = min(m, n)
            a, qo, ro = int(sl[0]), int(slQ[0]), int(slR[0])
            words.append((m, n, k, mo, qo, a, ro, 0))
            self.max_rows, self.max_k = max(self.max_rows, m), max(self.max_k, k)
            if n > k:
                gq.append((m, k, qo, k, 1, len(tq), 1, 1))
                tq.append((a + k, n, 1, ro + k, 1, n, n - k, 2 | (0 << 8) | (1 << 16)))
                gy.append((m, n - k, a + k, n, 1, len(ty), 1, 0))
                ty.append((qo, k, 1, ro + k, n, 1, k, (0 << 8) | (1 << 16)))
            gm1.append((k, k, mo, k, 1, len(tm1), 1, 0))
            tm1.append((ro, n, 1, ro, 1, n, k, 2 | (0 << 8) | (1 << 16)))
            gm2.append((k, k, mo, k, 1, len(tm2), 1, 0))
            tm2.append((qo, 1, k, qo, k, 1, m, 1 | (0 << 8) | (1 << 16)))
            gb.append((m, k, qo, k, 1, len(tb), 1, 1))
            tb.append((qo, k, 1, mo, k, 1, k, (0 << 8) | (1 << 16)))
            mo += k * k
        self.nblocks = len(words)
        self.m_numel = max(mo, 1)
        self.h_words = np.ascontiguousarray(np.array(words, dtype=np.int64).reshape(-1, L.YH_QRADJ_WORDS))
        L.call('yastn_hip_check_qr_adjoint_plan', self.h_words.ctypes.data, self.nblocks, self.dt, self.m_numel,
               sizes[0], a_numel, sizes[1])
        self.d_words = torch.from_numpy(self.h_words).to(device)
        nq, nr = max(sizes[0], 1), max(sizes[1], 1)
        self.gq = GemmPlan(gq, tq, (a_numel, nr), (nq,), device) if gq else None
        self.gy = GemmPlan(gy, ty, (nq, nr), (a_numel,), device) if gy else None
        self.m1 = GemmPlan(gm1, tm1, (nr, nr), (self.m_numel,), device)
        self.m2 = GemmPlan(gm2, tm2, (nq, nq), (self.m_numel,), device)
        self.b = GemmPlan(gb, tb, (nq, self.m_numel), (nq,), device)


class TdsPlan:
    """transpose_dot_sum (no-fusion policy, reference backend_np.py:590-601,
    _backend_torch_backwards.py:141-226): the participating blocks of A and B are permuted into packed
    matrices At (Dl x Dr each) and Bt by box gathers (with their inverse maps for the adjoint), then
    every output block is one GEMM whose terms are its (ta, tb) products, summed in order.  The adjoint
    is two grouped GEMMs whose outputs are the packed matrices: gAt[ta] = sum gC Bt[tb]^H and
    gBt[tb] = sum At[ta]^H gC over the outputs that use them (no atomics)."""

    def __init__(self, meta_dot, Areshape, Breshape, Aorder, Border, a_numel, b_numel, Dsize, device):
        def pack(reshape, order, numel):
            inv = tuple(int(x) for x in np.argsort(order))
            fwd, bwd, offs, o = [], [], [], 0
            for sl, Di, Dl, Dr in reshape:
                Di = tuple(int(x) for x in Di)
                Dn = tuple(Di[i] for i in order)
                st = _c_strides(Di)
                fwd.append((int(sl[0]), o, Dn, tuple(st[i] for i in order), _c_strides(Dn)))
                sn = _c_strides(Dn)
                bwd.append((o, int(sl[0]), Di, tuple(sn[i] for i in inv), st))
                offs.append(o)
                o += int(Dl) * int(Dr)
            packed = max(o, 1)
            return (GatherPlan(fwd, numel, packed, device), GatherPlan(bwd, packed, numel, device), offs, packed)

        self.a_fwd, self.a_bwd, a_off, self.at_numel = pack(Areshape, Aorder, a_numel)
        self.b_fwd, self.b_bwd, b_off, self.bt_numel = pack(Breshape, Border, b_numel)
        g, tm = [], []
        ua = [[] for _ in Areshape]
        ub = [[] for _ in Breshape]
        for sl, Dslc, tab in meta_dot:
            M, N = int(Dslc[0]), int(Dslc[1])
            g.append((M, N, int(sl[0]), N, 1, len(tm), len(tab), 0))
            for ta, tb in tab:
                Dla, K = int(Areshape[ta][2]), int(Areshape[ta][3])
                Klb, Nb = int(Breshape[tb][2]), int(Breshape[tb][3])
                if Dla != M or Klb != K or Nb != N:
                    raise ValueError(f"transpose_dot_sum shapes ({M}, {N}) from ({Dla}, {K}) x ({Klb}, {Nb})")
                tm.append((a_off[ta], K, 1, b_off[tb], N, 1, K, (0 << 8) | (1 << 16)))
                ua[ta].append((int(sl[0]), M, N, tb))
                ub[tb].append((int(sl[0]), M, N, ta))
        self.dot = GemmPlan(g, tm, (self.at_numel, self.bt_numel), (max(Dsize, 1),), device)
        ga, ta_ = [], []
        for ia, uses in enumerate(ua):
            M, K = int(Areshape[ia][2]), int(Areshape[ia][3])
            ga.append((M, K, a_off[ia], K, 1, len(ta_), len(uses), 0))
            for c0, _, N, tb in uses:  # gC (M x N) Bt[tb]^H (N x K)
                ta_.append((c0, N, 1, b_off[tb], 1, N, N, 2 | (0 << 8) | (1 << 16)))
        gb, tb_ = [], []
        for ib, uses in enumerate(ub):
            K, N = int(Breshape[ib][2]), int(Breshape[ib][3])
            gb.append((K, N, b_off[ib], N, 1, len(tb_), len(uses), 0))
            for c0, M, _, ia in uses:  # At[ia]^H (K x M) gC (M x N)
                tb_.append((a_off[ia], 1, K, c0, N, 1, M, 1 | (0 << 8) | (1 << 16)))
        self.dot_a = GemmPlan(ga, ta_, (max(Dsize, 1), self.bt_numel), (self.at_numel,), device) if ga else None
        self.dot_b = GemmPlan(gb, tb_, (self.at_numel, max(Dsize, 1)), (self.bt_numel,), device) if gb else None


def tds_plan(meta_dot, Areshape, Breshape, Aorder, Border, a_numel, b_numel, Dsize, device):
    refs = (meta_dot, Areshape, Breshape)
    key = ('tds', id(meta_dot), id(Areshape), id(Breshape), tuple(Aorder), tuple(Border), a_numel, b_numel, Dsize,
           _device_key(device))
    return _CACHE.get(key, refs, lambda: TdsPlan(meta_dot, Areshape, Breshape, Aorder, Border, a_numel, b_numel,
                                                  Dsize, device))


class DotDiagPlan:
    """dot_diag work list: one job per output block (out_off, a_off, b_off, n, inner, D), and the jobs
    grouped by the diagonal block they read, for the adjoint with respect to the diagonal."""

    def __init__(self, meta, axis, a_ndim, a_numel, b_numel, Dsize, device):
        jobs, groups = [], {}
        self.covered = 0
        for sln, slb, Db, sla in meta:
            Db = _as_shape(Db)
            if len(Db) != a_ndim:
                raise ValueError(f"dot_diag block {Db} of rank != {a_ndim}")
            n, D = math.prod(Db), Db[axis]
            inner = math.prod(Db[axis + 1:])
            if sln[1] - sln[0] != n or slb[1] - slb[0] != n or sla[1] - sla[0] != D:
                raise ValueError(f"dot_diag slices {sln} {slb} {sla} for block {Db}")
            if n == 0:
                continue
            groups.setdefault((sla[0], D), []).append(len(jobs))
            jobs.append((sln[0], sla[0], slb[0], n, inner, D, 0, 0))
            self.covered += n
        ids, grp = [], []
        for (a_off, D), js in groups.items():
            grp.append((a_off, D, len(ids), len(js)))
            ids.extend(js)
        self.h_jobs = np.array(jobs, dtype=np.int64).reshape(-1, L.YH_DIAG_WORDS)
        self.h_groups = np.array(grp, dtype=np.int64).reshape(-1, 4)
        self.h_ids = np.array(ids, dtype=np.int64)
        L.call('yastn_hip_check_dot_diag_plan', self.h_jobs.ctypes.data, len(self.h_jobs), self.h_groups.ctypes.data,
               len(self.h_groups), self.h_ids.ctypes.data, len(self.h_ids), Dsize, a_numel, b_numel)
        self.njobs, self.ngroups = len(self.h_jobs), len(self.h_groups)
        self.max_n = int(self.h_jobs[:, 3].max()) if self.njobs else 0
        self.max_D = int(self.h_groups[:, 1].max()) if self.ngroups else 0
        self.a_covered = sum(D for _, D, _, _ in grp)
        # the adjoint with respect to B runs the same kernel with the output and dense slices swapped
        self.h_jobs_t = self.h_jobs[:, [2, 1, 0, 3, 4, 5, 6, 7]].copy()
        L.call('yastn_hip_check_dot_diag_plan', self.h_jobs_t.ctypes.data, len(self.h_jobs_t), None, 0, None, 0,
               b_numel, a_numel, Dsize)
        self.b_covered = int(self.h_jobs[:, 3].sum()) if self.njobs else 0
        self.d_jobs = torch.from_numpy(self.h_jobs).to(device)
        self.d_jobs_t = torch.from_numpy(self.h_jobs_t).to(device)
        self.d_groups = torch.from_numpy(self.h_groups).to(device)
        self.d_ids = torch.from_numpy(self.h_ids).to(device)


def dot_diag_plan(meta, axis, a_ndim, a_numel, b_numel, Dsize, device):
    return _CACHE.get(('dot_diag', id(meta), axis, a_ndim, a_numel, b_numel, Dsize, _device_key(device)), (meta,),
                      lambda: DotDiagPlan(meta, axis, a_ndim, a_numel, b_numel, Dsize, device))


def qr_plan(meta, itemsize, a_numel, sizes, device):
    return _CACHE.get(('qr', id(meta), itemsize, a_numel, tuple(sizes), _device_key(device)), (meta,),
                      lambda: QrPlan(meta, itemsize, a_numel, sizes, device))


def qr_adjoint_plan(meta, itemsize, a_numel, sizes, device):
    return _CACHE.get(('qradj', id(meta), itemsize, a_numel, tuple(sizes), _device_key(device)), (meta,),
                      lambda: QrAdjointPlan(meta, itemsize, a_numel, sizes, device))


def svd_plan(meta, itemsize, a_numel, sizes, device):
    def build():
        blocks = [(sl[0], D[0], D[1], slU[0], slS[0], slV[0]) for sl, D, slU, DU, slS, slV, DV in meta]
        for sl, D, slU, DU, slS, slV, DV in meta:
            k = min(D)
            if tuple(DU) != (D[0], k) or tuple(DV) != (k, D[1]) or slS[1] - slS[0] != k:
                raise ValueError(f"svd meta shapes {D} {DU} {DV} {slS}")
        return SvdPlan(blocks, itemsize, 0, a_numel, sizes[0], sizes[1], sizes[2], device)
    return _CACHE.get(('svd', id(meta), itemsize, a_numel, tuple(sizes), _device_key(device)), (meta,), build)


def eigh_plan(meta, itemsize, a_numel, sizes, device):
    def build():
        blocks = []
        for sl, D, slU, DU, slS in meta:
            if D[0] != D[1] or tuple(DU) != tuple(D) or slS[1] - slS[0] != D[0]:
                raise ValueError(f"eigh meta shapes {D} {DU} {slS}")
            blocks.append((sl[0], D[0], D[1], slU[0], slS[0], 0))
        return SvdPlan(blocks, itemsize, 1, a_numel, sizes[1], sizes[0], 1, device)
    return _CACHE.get(('eigh', id(meta), itemsize, a_numel, tuple(sizes), _device_key(device)), (meta,), build)


def as_ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)
