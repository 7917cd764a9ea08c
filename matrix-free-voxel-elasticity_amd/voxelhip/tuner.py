"""Empirical launch tuning of the parity product.

The product kernel takes two launch parameters: the node tile of a workgroup (16 x 32, 8 x 32,
4 x 8, or the wide tile of one workgroup per CU: 22 x 32 nodes in FP64, 30 x 32 in FP32, in y, z)
and the x-chunk length lx (0 lets the occupancy model choose).  Their best values depend on the
grid, the precision and the fused pass (a smoothing pass streams two more vectors than a plain
product), and the occupancy model does not rank them reliably.  `LaunchTuner` times every
candidate with HIP events on the operands of the pass being tuned and keeps the fastest; results
are cached per process and, when a path is given (or the VOXELHIP_TUNE_CACHE environment variable
names one), in a JSON file keyed by device name, grid, precision, mask kind and pass.

Tuning is a set-up step: it launches each candidate a few times on scratch vectors and
leaves the level's own vectors untouched.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import torch

from . import _lib as L
from . import ops

TILES = (L.TILE_16x32, L.TILE_8x32, L.TILE_WIDE)
LX_CANDIDATES = (0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64)
# one-wave 4 x 8 tiles, for grids small enough to be latency-bound (elements)
SMALL_GRID = 64 * 32 * 32


def _candidates(nx, n_el=None):
    seen, out = set(), []
    tiles = TILES + ((L.TILE_4x8,) if n_el is not None and n_el <= SMALL_GRID else ())
    for t in tiles:
        for lx in LX_CANDIDATES:
            lx_eff = min(lx, nx)
            if (t, lx_eff) not in seen:
                seen.add((t, lx_eff))
                out.append((t, lx_eff))
    return out


class LaunchTuner:
    def __init__(self, path=None, reps=8, rounds=3):
        self.path = Path(path) if path else (Path(os.environ["VOXELHIP_TUNE_CACHE"])
                                             if os.environ.get("VOXELHIP_TUNE_CACHE") else None)
        self.reps, self.rounds = int(reps), int(rounds)
        self.cache = {}
        if self.path is not None and self.path.exists():
            try:
                self.cache = {k: tuple(v) for k, v in json.loads(self.path.read_text()).items()}
            except (OSError, ValueError):
                self.cache = {}
        self.timings = {}

    @staticmethod
    def key(level, mode, combine, dinv_array):
        dev = torch.cuda.get_device_name(level.E.device)
        st = level.product.storage
        return "|".join(str(v) for v in (dev, *level.shape, str(level.dtype)[6:], st if st is not None else "-",
                                         level.maskobj.kind, mode, int(combine), int(dinv_array)))

    def _save(self):
        if self.path is None:
            return
        self.path.parent.mkdir(parents=True, exist_ok=True)
        tmp = self.path.with_suffix(".tmp")
        tmp.write_text(json.dumps(self.cache, indent=0, sort_keys=True))
        os.replace(tmp, self.path)

    def _time(self, fn):
        """Mean duration of one launch, from a HIP graph of `reps` launches replayed `rounds`
        times: the graph removes the host's per-launch cost, which otherwise hides the
        differences between the plans of the small coarse-level products."""
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(self.reps):
                fn()
        graph.replay()
        s = torch.cuda.current_stream()
        best = float("inf")
        for _ in range(self.rounds):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            graph.replay()
            e1.record(s)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) / self.reps)
        del graph
        return best * 1e3                                    # microseconds

    def tune(self, level, mode, combine=False, dinv_array=False):
        """Best (tile, lx) of `level`'s product for one pass kind (cached)."""
        k = self.key(level, mode, combine, dinv_array)
        if k in self.cache:
            return self.cache[k]
        n, dt, dev = level.n, level.dtype, level.E.device
        g = torch.Generator(device=dev).manual_seed(1234)
        vec = lambda: torch.randn(n, generator=g, device=dev, dtype=torch.float64).to(dt)  # noqa: E731
        u, y, b, w, uo = vec(), vec(), vec(), vec(), vec()
        if level.product.storage == L.BF16:
            u = u.to(torch.bfloat16)
        sdev = torch.zeros(16, dtype=torch.float64, device=dev)
        cands = _candidates(level.shape[0], level.shape[0] * level.shape[1] * level.shape[2])
        most = max(ops.matvec_tiles(level.shape, lx=lx, tile=t, dtype=dt, blocks=level.product.blocks,
                                    storage=level.product.storage)[0] for t, lx in cands)
        partial = torch.zeros(2 * max(most, 1), dtype=torch.float64, device=dev)
        kw = {"mode": mode}
        smoothing = (L.MODE_JACOBI, L.MODE_CHEB, L.MODE_CHEB3, L.MODE_CHEB3D)
        if mode in (L.MODE_RESID,) + smoothing:
            kw["b"] = b
        if mode in smoothing:
            if dinv_array:
                kw["dinv"] = level.inv_diag
            else:
                kw["kdiag"] = level.product.kdiag
            kw.update(omega=0.5, cheb_a=0.5, cheb_c=0.25)
        if mode == L.MODE_CHEB:
            kw["d"] = vec()
        if mode in (L.MODE_DOT, L.MODE_RESID, L.MODE_CHEB3D):
            kw["partial"] = partial
        if mode == L.MODE_CHEB3D:
            kw.update(zold=vec(), prev=vec())
        if combine:
            kw.update(w=w, sdev=sdev, sidx=0, u_out=uo)
            if mode == L.MODE_CHEB3:         # the first Chebyshev step starts from a zero iterate
                kw["prev_zero"] = True
        res = {}
        for t, lx in cands:
            res[(t, lx)] = self._time(lambda: ops.matvec(level.product.blocks, level.E, u, y, level.shape,
                                                         mask=level.maskobj, storage=level.product.storage,
                                                         tile=t, lx=lx, **kw))
        best = min(res, key=res.get)
        self.timings[k] = {f"t{t}l{lx}": round(v, 2) for (t, lx), v in res.items()}
        self.cache[k] = best
        self._save()
        return best
