"""Batched GMG-preconditioned flexible CG: many voxel problems, each with several load cases,
solved together with ONE launch per pass of the iteration.

Every (problem, load case) pair is an instance with its own vectors; the instances of a
problem share its moduli, masks, hierarchy, Chebyshev bounds and coarsest inverse.  Vectors
of all instances at one hierarchy depth live in one concatenated buffer, so elementwise work
is a single launch over the buffer, and every operator pass (products with their fused
epilogues, transfers, coarsest solves, segmented dots and the per-instance scalar recurrences)
is one launch of a descriptor-table kernel (include/voxelhip.h, "batched problems").  Problems
of different sizes have different hierarchy depths: an instance whose coarsest level is depth d
takes the dense solve there while deeper instances keep smoothing, and all coarsest solves of
the batch are one launch between the down and the up sweep.  With graph=True the iteration is
captured in two HIP graphs (buffer parity), replayed once per iteration.  Convergence is
decided on the device: each iteration's residual pass ends in a per-instance scalar step that
freezes a converged instance (its recurrence yields alpha = beta = 0, so its iterate stops
moving); the host reads the residual norms asynchronously while the next iteration runs.

The algorithm per instance is that of voxelhip.engine.GMGSolver (flexible CG with the true
residual every iteration, src/voxel_mg.py:181-217; Chebyshev-Jacobi V-cycle), precision "fp64"
or "mixed" (FP32 V-cycle and FP32 search direction in the FP64 outer iteration: x, b and the
true residual's arithmetic stay FP64).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from . import ops
from .engine import build_hierarchy

import voxel_mg as mg

_DT = {torch.float64: L.F64, torch.float32: L.F32}


def _ptr(t, offset=0):
    """Device address of element `offset` of tensor t (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr() + int(offset) * t.element_size()


class _Handle:
    def __init__(self, h):
        self.h = h

    def launch(self):
        L.call("vh_batch_launch", self.h, ops._stream())

    def offsets(self):
        p, n, tot = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
        L.call("vh_batch_offsets", self.h, ctypes.byref(p), ctypes.byref(n), ctypes.byref(tot))
        return p.value, n.value, tot.value

    def __del__(self):
        lib = L._LIB
        if lib is not None and self.h:
            lib.vh_batch_destroy(self.h)
            self.h = None


def product_batch(dtype, mask_kind, tile, items):
    arr = (L.ProductItem * len(items))(*items)
    h = ctypes.c_void_p()
    tiles = (ctypes.c_int * len(items))()
    L.call("vh_product_batch_create", _DT[dtype], mask_kind, tile, len(items), arr, ctypes.byref(h), tiles)
    return _Handle(h.value), list(tiles)


def segments(lengths):
    arr = (ctypes.c_longlong * len(lengths))(*[int(v) for v in lengths])
    h = ctypes.c_void_p()
    L.call("vh_segments_create", len(lengths), arr, ctypes.byref(h))
    return _Handle(h.value)


def transfer_batch(dtype, restrict_op, items, assign=False):
    arr = (L.TransferItem * len(items))(*items)
    h = ctypes.c_void_p()
    L.call("vh_transfer_batch_create", _DT[dtype], int(restrict_op), int(assign), len(items), arr, ctypes.byref(h))
    return _Handle(h.value)


def dense_batch(a_dtype, v_dtype, items):
    arr = (L.DenseItem * len(items))(*items)
    h = ctypes.c_void_p()
    L.call("vh_dense_batch_create", _DT[a_dtype], _DT[v_dtype], len(items), arr, ctypes.byref(h))
    return _Handle(h.value)


def _epilogue(mode, **kw):
    kd = kw.pop("kdiag", None)
    ep = L.Epilogue()
    ep.mode = mode
    ep.fixed_zero = int(bool(kw.pop("fixed_zero", False)))
    for name in ("b", "dinv", "d", "partial", "w", "sdev", "u_out", "y32"):
        setattr(ep, name, kw.pop(name, None))
    ep.omega = float(kw.pop("omega", 0.0))
    ep.load_scale = float(kw.pop("load_scale", 1.0))
    ep.cheb_a = float(kw.pop("cheb_a", 0.0))
    ep.cheb_c = float(kw.pop("cheb_c", 0.0))
    ep.lx = int(kw.pop("lx", 0))
    ep.tile = 0
    ep.flags = int(kw.pop("flags", 0))
    ep.sidx = int(kw.pop("sidx", 0))
    if kd is not None:
        for c in range(3):
            ep.kdiag[c] = float(kd[c])
    if kw:
        raise TypeError(f"unknown epilogue fields {sorted(kw)}")
    return ep


class BatchProblem:
    """One voxel problem: grid shape, element moduli E (n_el), mask values (3 n_nodes, 1 = free,
    or a voxelhip.ops.Mask) and its load cases (a list of force vectors, 3 n_nodes each)."""

    def __init__(self, shape, E, mask, loads):
        self.shape = tuple(int(s) for s in shape)
        self.E = E
        self.mask = mask
        self.loads = list(loads)


def coarse_dof_shapes(shape, max_coarse_dofs=3000):
    """Hierarchy of a batched problem: halve while every extent is even and the level has more
    than max_coarse_dofs DOFs.  (src/voxel_mg.py:220-225 halves while x > max_coarse_x, which
    suits long cantilevers; the problems of a batch range from cubes to slabs, and bounding the
    coarsest DOF count bounds every problem's dense coarsest inverse.)"""
    shapes = [tuple(int(s) for s in shape)]
    while all(s % 2 == 0 for s in shapes[-1]) and 3 * ops.n_nodes(shapes[-1]) > max_coarse_dofs:
        shapes.append(tuple(s // 2 for s in shapes[-1]))
    return shapes


class BatchedGMGSolver:
    def __init__(self, problems, ke, *, max_coarse_x=None, max_coarse_dofs=3000, degree=3, precision="mixed",
                 graph=True, lower_frac=1.0 / 30.0, omega_iterations=30, tile=L.TILE_16x16, lx=0):
        if precision not in ("fp64", "mixed"):
            raise ValueError(precision)
        if int(degree) < 2:
            raise ValueError("the batched Chebyshev smoother needs degree >= 2")
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.ke = np.asarray(ke, dtype=np.float64)
        self.precision = precision
        self.pdtype = torch.float64 if precision == "fp64" else torch.float32
        self.degree, self.lower_frac = int(degree), float(lower_frac)
        self.graph, self.tile, self.lx = bool(graph), int(tile), int(lx)
        self.problems = list(problems)
        # ---- per-problem hierarchies
        self.hier = []
        mask_kinds = set()
        for pb in self.problems:
            shapes = (mg.hierarchy_shapes(pb.shape, max_coarse_x) if max_coarse_x is not None
                      else coarse_dof_shapes(pb.shape, max_coarse_dofs))
            mask = pb.mask if isinstance(pb.mask, ops.Mask) else ops.Mask.from_values(pb.mask, pb.shape, self.dev)
            mask_kinds.add(mask.kind)
            E64 = voxelhip_device(pb.E, torch.float64, self.dev)
            fine, M = build_hierarchy(pb.shape, self.ke, mask, E64, shapes, self.pdtype, smoother="chebyshev",
                                      degree=self.degree, lower_frac=self.lower_frac,
                                      omega_iterations=omega_iterations)
            self.hier.append((shapes, fine, M))
        if len(mask_kinds) != 1:
            raise ValueError("all problems of a batch need the same mask kind (x = 0 clamp or mask arrays)")
        if any(len(h[0]) < 2 for h in self.hier):
            raise ValueError("every problem of a batch needs at least two levels")
        self.mask_kind = mask_kinds.pop()
        # ---- instances (problem, load case)
        self.inst = [(p, l) for p, pb in enumerate(self.problems) for l in range(len(pb.loads))]
        self.n = len(self.inst)
        self.depth = [len(self.hier[p][0]) - 1 for p, _ in self.inst]     # depth of each coarsest level
        self.D = max(self.depth)
        self._alloc()
        self._build()
        self._graphs = None
        self._start_graph = None

    # ------------------------------------------------------------------ buffers
    def _alloc(self):
        dev, n = self.dev, self.n
        # per depth: element offsets of each present instance in the concatenated buffers
        self.off = []
        for d in range(self.D + 1):
            offs, tot = {}, 0
            for k, (p, _) in enumerate(self.inst):
                if self.depth[k] >= d:
                    offs[k] = tot
                    tot += self.hier[p][2].levels[d].n
            self.off.append((offs, tot))
        f64 = dict(dtype=torch.float64, device=dev)
        pd = dict(dtype=self.pdtype, device=dev)
        nf = self.off[0][1]
        z = lambda: torch.zeros(nf, **f64)  # noqa: E731
        # mixed precision forms the search direction in FP32, as GMGSolver does: z and p are FP32,
        # the DOT pass runs on the FP32 fine level, the FP64 RESID pass reads the FP32 p and
        # writes only the FP32 copy of r (the V-cycle right-hand side; the Polak-Ribiere dots
        # take r from it), so no FP64 r exists
        self.p32 = self.pdtype != torch.float64
        self.b = z()
        self.r = None if self.p32 else z()
        self.xs = [z(), z()]
        self.ps = [torch.zeros(nf, **pd), torch.zeros(nf, **pd)]
        self.zs = [torch.zeros(nf, **pd), torch.zeros(nf, **pd)]
        self.scal = torch.zeros(16 * n, **f64)
        self.tol2 = torch.zeros(1, **f64)         # tol^2 of the running solve
        self.X, self.X2, self.RES, self.RHS, self.DT = [], [], [], [], []
        for d in range(self.D + 1):
            tot = self.off[d][1]
            self.X.append(torch.zeros(tot, **pd))
            self.X2.append(torch.zeros(tot, **pd))
            self.RES.append(torch.zeros(tot, **pd))
            self.RHS.append(torch.zeros(tot, **pd) if (d > 0 or self.pdtype != torch.float64) else self.r)
            self.DT.append(torch.zeros(tot, **pd))
        # D^-1 / theta per instance (the first Chebyshev step), copied per load case
        for d in range(self.D + 1):
            offs, _ = self.off[d]
            for k, o in offs.items():
                p = self.inst[k][0]
                M = self.hier[p][2]
                if d < len(M.levels) - 1:
                    theta, _ = mg._cheb_params(M.lambdas[d], self.lower_frac)
                    lv = M.levels[d]
                    self.DT[d][o:o + lv.n] = lv.inv_diag / theta
        self.segs = segments([self.hier[p][1].n for p, _ in self.inst])
        self.seg_offs, _, self.seg_total = self.segs.offsets()

    def _build(self):
        n, D = self.n, self.D
        f64 = torch.float64
        mk = self.mask_kind
        # ---- fine level (FP64): the two parities of the DOT and RESID passes.  Tile counts are
        # known once the tables exist: build against a placeholder partial array, size the real
        # one, and build again.
        self.partial = torch.zeros(1, dtype=f64, device=self.dev)
        for attempt in range(2):
            base = self.partial.data_ptr()
            self.dot_h, self.res_h = [], []
            for i in (0, 1):
                o = 1 - i
                items_d, items_r = [], []
                for k in range(n):
                    p = self.inst[k][0]
                    fine = self.hier[p][1]
                    fo = self.off[0][0][k]
                    sdev = _ptr(self.scal, 16 * k)
                    items_d.append(self._item(self.hier[p][2].levels[0] if self.p32 else fine, _ptr(self.zs[i], fo),
                                              None, L.MODE_DOT, partial=base, w=_ptr(self.ps[o], fo), sdev=sdev,
                                              sidx=L.S_BETA, u_out=_ptr(self.ps[i], fo)))
                    items_r.append(self._item(fine, _ptr(self.xs[o], fo), None if self.p32 else _ptr(self.r, fo),
                                              L.MODE_RESID, b=_ptr(self.b, fo), partial=base,
                                              w=_ptr(self.ps[i], fo), sdev=sdev, sidx=L.S_ALPHA,
                                              u_out=_ptr(self.xs[i], fo),
                                              y32=_ptr(self.RHS[0], fo) if self.p32 else None,
                                              flags=L.W_F32 if self.p32 else 0))
                self.dot_h.append(product_batch(self.pdtype, mk, self.tile, items_d))
                self.res_h.append(product_batch(f64, mk, self.tile, items_r))
            if attempt == 0:
                need = max(max(sum(t) for _, t in self.dot_h + self.res_h), 2 * self.seg_total)
                self.partial = torch.zeros(need, dtype=f64, device=self.dev)
        self.dot_offs = [h.offsets() for h, _ in self.dot_h]
        self.res_offs = [h.offsets() for h, _ in self.res_h]
        # ---- V-cycle tables per depth
        pd = self.pdtype
        self.vc = []                      # per depth: dict of handles
        for d in range(D):
            ks = [k for k in range(n) if self.depth[k] > d]
            lv = lambda k: self.hier[self.inst[k][0]][2].levels[d]  # noqa: E731
            o = lambda k: self.off[d][0][k]  # noqa: E731
            bufs = {"X": self.X[d], "X2": self.X2[d]}
            steps = []
            # pre: steps 1 and 2 in one pass (x1 = D^-1 r / theta formed at load time from the
            # rhs and the scaled inverse diagonal DT, written to X), then the CHEB3 recurrence
            cur, oth = "X", "X2"
            alpha = {k: None for k in ks}
            th = {k: mg._cheb_params(self._lam(k, d), self.lower_frac) for k in ks}
            for k in ks:
                alpha[k] = 2.0 / th[k][0]
            for s in range(1, self.degree):
                items = []
                for k in ks:
                    theta, delta = th[k]
                    c = 0.25 * delta * delta * alpha[k]
                    alpha[k] = 1.0 / (theta - c)
                    if s == 1:
                        items.append(self._item(lv(k), _ptr(self.RHS[d], o(k)), _ptr(bufs[oth], o(k)),
                                                L.MODE_CHEB3, b=_ptr(self.RHS[d], o(k)), cheb_a=alpha[k], cheb_c=c,
                                                kdiag=lv(k).product.kdiag, flags=1, w=_ptr(self.DT[d], o(k)),
                                                load_scale=1.0, u_out=_ptr(bufs[cur], o(k))))
                    else:
                        items.append(self._item(lv(k), _ptr(bufs[cur], o(k)), _ptr(bufs[oth], o(k)), L.MODE_CHEB3,
                                                b=_ptr(self.RHS[d], o(k)), cheb_a=alpha[k], cheb_c=c,
                                                kdiag=lv(k).product.kdiag))
                steps.append(product_batch(pd, mk, self.tile, items)[0])
                cur, oth = oth, cur
            pre_out = cur
            res_items = [self._item(lv(k), _ptr(bufs[cur], o(k)), _ptr(self.RES[d], o(k)), L.MODE_RESID,
                                    b=_ptr(self.RHS[d], o(k)), fixed_zero=True) for k in ks]
            resid = product_batch(pd, mk, self.tile, res_items)[0]
            r_items = []
            for k in ks:
                M = self.hier[self.inst[k][0]][2]
                it = L.TransferItem()
                it.src, it.dst = _ptr(self.RES[d], o(k)), _ptr(self.RHS[d + 1], self.off[d + 1][0][k])
                it.nx, it.ny, it.nz = M.levels[d].shape
                it.fmask_kind, it.fmask = M.levels[d].maskobj.kind, M.levels[d].maskobj.ptr
                it.cmask_kind, it.cmask = M.levels[d + 1].maskobj.kind, M.levels[d + 1].maskobj.ptr
                r_items.append(it)
            restr = transfer_batch(pd, True, r_items)
            # post: first step with c = 0 from the prolongated iterate (the damped Jacobi sweep with
            # omega = 1 / theta, which does not read the previous iterate), then the recurrence
            post = []
            items = []
            for k in ks:
                items.append(self._item(lv(k), _ptr(bufs[cur], o(k)), _ptr(bufs[oth], o(k)), L.MODE_JACOBI,
                                        b=_ptr(self.RHS[d], o(k)), omega=1.0 / th[k][0],
                                        kdiag=lv(k).product.kdiag))
            post.append(product_batch(pd, mk, self.tile, items)[0])
            post_in = cur
            cur, oth = oth, cur
            for k in ks:
                alpha[k] = 2.0 / th[k][0]
            for s in range(1, self.degree):
                items = []
                for k in ks:
                    theta, delta = th[k]
                    c = 0.25 * delta * delta * alpha[k]
                    alpha[k] = 1.0 / (theta - c)
                    items.append(self._item(lv(k), _ptr(bufs[cur], o(k)), _ptr(bufs[oth], o(k)), L.MODE_CHEB3,
                                            b=_ptr(self.RHS[d], o(k)), cheb_a=alpha[k], cheb_c=c,
                                            kdiag=lv(k).product.kdiag))
                post.append(product_batch(pd, mk, self.tile, items)[0])
                cur, oth = oth, cur
            self.vc.append(dict(ks=ks, pre=steps, pre_out=pre_out, resid=resid, restrict=restr, post=post,
                                post_in=post_in, out=cur))
        # the output buffer of every instance at every depth (coarsest: X of that depth)
        def out_buf(k, d):
            return self.X[d] if self.depth[k] == d else (self.X[d] if self.vc[d]["out"] == "X" else self.X2[d])
        for d in range(D):
            ks = self.vc[d]["ks"]
            p_items = []
            for k in ks:
                M = self.hier[self.inst[k][0]][2]
                it = L.TransferItem()
                it.src = _ptr(out_buf(k, d + 1), self.off[d + 1][0][k])
                tgt = self.X[d] if self.vc[d]["post_in"] == "X" else self.X2[d]
                it.dst = _ptr(tgt, self.off[d][0][k])
                it.nx, it.ny, it.nz = M.levels[d].shape
                it.fmask_kind, it.fmask = M.levels[d].maskobj.kind, M.levels[d].maskobj.ptr
                it.cmask_kind, it.cmask = M.levels[d + 1].maskobj.kind, M.levels[d + 1].maskobj.ptr
                p_items.append(it)
            self.vc[d]["prolong"] = transfer_batch(pd, False, p_items)
        # coarsest solves of every instance in one launch
        items = []
        for k in range(n):
            d = self.depth[k]
            M = self.hier[self.inst[k][0]][2]
            it = L.DenseItem()
            it.A = M.coarse_inverse.data_ptr()
            it.ld, it.n = M.coarse_inverse.stride(0), M.levels[-1].n
            it.r = _ptr(self.RHS[d], self.off[d][0][k])
            it.y = _ptr(self.X[d], self.off[d][0][k])
            it.mask = M._coarse_mask.array.data_ptr()
            items.append(it)
        cdt = self.hier[0][2].coarse_inverse.dtype
        self.coarse = dense_batch(cdt, pd, items)
        self.zl = self.X[0] if self.vc[0]["out"] == "X" else self.X2[0]

    def _lam(self, k, d):
        return self.hier[self.inst[k][0]][2].lambdas[d]

    def _item(self, level, u, y, mode, **ep):
        it = L.ProductItem()
        blocks = np.ascontiguousarray(level.product.blocks, dtype=np.float64).reshape(72)
        self._keep = getattr(self, "_keep", [])
        self._keep.append(blocks)
        it.blocks72 = blocks.ctypes.data
        it.E = level.E.data_ptr()
        it.u, it.y = u, y
        it.nx, it.ny, it.nz = level.shape
        it.mask = level.maskobj.ptr
        it.ep = _epilogue(mode, lx=self.lx, **ep)
        return it

    # ------------------------------------------------------------------ one V-cycle (all instances)
    def vcycle(self):
        for d in range(self.D):
            v = self.vc[d]
            for h in v["pre"]:
                h.launch()
            v["resid"].launch()
            v["restrict"].launch()
        self.coarse.launch()
        for d in range(self.D - 1, -1, -1):
            v = self.vc[d]
            v["prolong"].launch()
            for h in v["post"]:
                h.launch()
        return self.zl

    def _finish(self, offs, count, dst, op):
        p, n, tot = offs
        L.call("vh_seg_finish", self.partial.data_ptr(), tot, p, n, count, self.scal.data_ptr(), 16, dst, op,
               ops._stream())

    def _iteration(self, i):
        o = 1 - i
        self.dot_h[i][0].launch()
        self._finish(self.dot_offs[i], 1, L.S_PQ, L.OP_ALPHA)
        self.res_h[i][0].launch()
        self._finish(self.res_offs[i], 1, 8, L.OP_RESNORM)
        zl = self.vcycle()
        vd = _DT[self.zs[i].dtype]
        L.call("vh_seg_dotdiff", self.segs.h, _DT[zl.dtype], vd, zl.data_ptr(), self.RHS[0].data_ptr(),
               self.zs[i].data_ptr(), self.zs[o].data_ptr(), self.partial.data_ptr(), ops._stream())
        self._finish((self.seg_offs, self.n, self.seg_total), 2, 5, L.OP_BETA_PR)

    def _start(self):
        """The start of a solve on the device: the scalar blocks (tol^2 |b|^2 per instance), x = 0,
        p = 0, z = M r with r = b, and r.z per instance (captured in a graph of its own)."""
        sv = self.scal.view(self.n, 16)
        sv.zero_()
        torch.mul(self.nb2, self.tol2, out=sv[:, L.S_TOL2])
        self.xs[1].zero_()
        self.ps[1].zero_()
        if self.p32:
            ops.convert(self.b, self.RHS[0])
        else:
            self.r.copy_(self.b)
        zl = self.vcycle()
        self.zs[0].copy_(zl)
        L.call("vh_seg_reduce", self.segs.h, _DT[self.zs[0].dtype], L.RED_DOT, self.RHS[0].data_ptr(),
               self.zs[0].data_ptr(), self.partial.data_ptr(), ops._stream())
        self._finish((self.seg_offs, self.n, self.seg_total), 1, L.S_RZ, L.OP_NONE)

    def _capture_start(self):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):          # warm-up on a side stream (the replay redoes its work)
            self._start()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._start()
        return g

    def _capture(self):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._iteration(0)
            self._iteration(1)
        torch.cuda.current_stream().wait_stream(s)
        graphs = []
        for i in (0, 1):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._iteration(i)
            graphs.append(g)
        return graphs

    # ------------------------------------------------------------------ solve
    def _upload_loads(self):
        """Every load case into the resident right-hand side b (once), and |b|^2 per instance."""
        for k, (p, l) in enumerate(self.inst):
            fo, nf = self.off[0][0][k], self.hier[p][1].n
            self.b[fo:fo + nf] = voxelhip_device(self.problems[p].loads[l], torch.float64, self.dev)
        self.scal.zero_()
        L.call("vh_seg_reduce", self.segs.h, L.F64, L.RED_NORM2, self.b.data_ptr(), None, self.partial.data_ptr(),
               ops._stream())
        self._finish((self.seg_offs, self.n, self.seg_total), 1, 9, L.OP_NONE)
        self.nb2 = self.scal.view(self.n, 16)[:, 9].clone()      # |b|^2 per instance, on the device
        self.nb = self.nb2.sqrt().cpu().numpy()
        self._loaded = True

    def set_load(self, k, force):
        """Replaces the load vector of instance k (a device or host vector of its problem's size)."""
        if not getattr(self, "_loaded", False):
            self._upload_loads()            # the other instances' loads first, so this one is kept
        p, _ = self.inst[k]
        fo, nf = self.off[0][0][k], self.hier[p][1].n
        self.b[fo:fo + nf] = voxelhip_device(force, torch.float64, self.dev)
        self.nb2[k] = self.b[fo:fo + nf].square().sum()
        self.nb[k] = float(self.nb2[k].sqrt())

    def solve(self, tol=1e-8, maxiter=200):
        """Solves every instance from x = 0 to ||b - A x|| / ||b|| <= tol.  Returns
        (x list per instance (views), iterations per instance, converged flags, rel residuals).

        The load vectors stay resident in HBM between solves (uploaded once; set_load replaces
        one).  Convergence is tested on the device: each iteration's residual pass ends in
        OP_RESNORM per instance, which counts the instance's iterations and freezes its recurrence
        (alpha = beta = 0) once |r|^2 <= tol^2 |b|^2.  The host reads the scalar block
        asynchronously and keeps the next iteration queued while some instance's residual
        history shows it cannot converge in the iteration being read, so the GPU does not wait
        for the host between iterations; an iteration past an instance's convergence leaves its
        iterate unchanged."""
        if self.graph and self._graphs is None:
            self._graphs = self._capture()
        if not getattr(self, "_loaded", False):
            self._upload_loads()
        n = self.n
        self.tol2.fill_(float(tol) ** 2)
        if self.graph:
            if self._start_graph is None:
                self._start_graph = self._capture_start()
            self._start_graph.replay()
        else:
            self._start()
        if not hasattr(self, "_hbuf"):
            self._hbuf = [torch.empty(16 * n, dtype=torch.float64, pin_memory=True) for _ in range(2)]
            self._hev = [torch.cuda.Event() for _ in range(2)]
        pending, hist = [], []
        nb = np.where(self.nb > 0, self.nb, 1.0)

        def launch(k):
            i = (k - 1) % 2
            if self._graphs is not None:
                self._graphs[i].replay()
            else:
                self._iteration(i)
            self._hbuf[k % 2].copy_(self.scal, non_blocking=True)
            self._hev[k % 2].record()
            pending.append(k)

        def next_cannot_converge():
            # some unconverged instance's residual, shrunk at its recent rate, stays above twice
            # the tolerance: the batch needs the iteration after the next one as well
            if not hist:
                return False
            h1 = hist[-1]
            h2 = hist[-2] if len(hist) > 1 else np.ones_like(h1)
            rate = h1 / np.maximum(h2, 1e-300)
            if len(hist) > 2:
                rate = np.maximum(rate, h2 / np.maximum(hist[-3], 1e-300))
            live = h1 > tol
            return bool((h1[live] * rate[live] > 2.0 * tol).any())

        launch(1)
        while True:
            if len(pending) == 1 and pending[0] < maxiter and next_cannot_converge():
                launch(pending[0] + 1)
            k = pending.pop(0)
            self._hev[k % 2].synchronize()
            v = self._hbuf[k % 2].view(n, 16).numpy()
            rel = np.sqrt(v[:, 8]) / nb
            hist.append(rel)
            done = v[:, L.S_DONE] != 0.0
            if done.all() or k >= maxiter:
                for q in pending:                     # iterations queued past convergence
                    self._hev[q % 2].synchronize()
                its = np.where(done, v[:, L.S_ITS], k).astype(int)
                last = pending[-1] if pending else k
                x = self.xs[(last - 1) % 2]
                out = []
                for j, (p, _) in enumerate(self.inst):
                    fo, nf = self.off[0][0][j], self.hier[p][1].n
                    out.append(x[fo:fo + nf])
                return out, its, done, rel
            if not pending:
                launch(k + 1)


def random_cantilevers(seed=4, n=48):
    """A batch workload: n cantilever-type problems with grids drawn from the seed between
    16 x 16 x 16 and 96 x 48 x 48 elements (every extent a multiple of 8), densities iid
    U[0.01, 1] smoothed by the radius-1.5 cone filter (E = 1e-9 + (1 - 1e-9) rho^3), x = 0 face
    clamped, and 1-4 load cases of one unit point load each at a random node of a random face
    other than x = 0.  Returns a list of BatchProblem."""
    from .simp import ConeFilter, modulus
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda", torch.cuda.current_device())
    out = []
    for _ in range(n):
        nx = 8 * int(rng.integers(2, 13))
        ny = 8 * int(rng.integers(2, min(6, nx // 8) + 1))
        nz = 8 * int(rng.integers(2, min(6, nx // 8) + 1))
        shape = (nx, ny, nz)
        rho = torch.from_numpy(rng.uniform(0.01, 1.0, nx * ny * nz)).to(dev)
        E = modulus(ConeFilter(shape, 1.5, device=dev).forward(rho), 3.0, 1.0, 1e-9)
        mask = np.ones((nx + 1, ny + 1, nz + 1, 3))
        mask[0] = 0
        loads = []
        for _ in range(int(rng.integers(1, 5))):
            f = np.zeros((nx + 1, ny + 1, nz + 1, 3))
            face = int(rng.integers(0, 5))
            ix, iy, iz = int(rng.integers(1, nx + 1)), int(rng.integers(0, ny + 1)), int(rng.integers(0, nz + 1))
            ix = nx if face == 0 else ix
            iy = 0 if face == 1 else (ny if face == 2 else iy)
            iz = 0 if face == 3 else (nz if face == 4 else iz)
            d = rng.normal(size=3)
            f[ix, iy, iz] = d / np.linalg.norm(d)
            loads.append(f.reshape(-1))
        out.append(BatchProblem(shape, E, mask.reshape(-1), loads))
    return out


def voxelhip_device(a, dtype, dev):
    if isinstance(a, torch.Tensor):
        return a.to(dev, dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)
