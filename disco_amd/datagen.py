"""Batched sample generation with the host's work hidden behind the device's.

A datagen loop (many samples per batch, many batches) has host work per batch:
the loop draws, their contours (the C++ producer, include/disco_geometry.h, restating
Code/loop.cpp:536-580) and the packing of the batch plan (disco_hip_batch_plan). It
also has device work: the samples' lattices (disco_hip_lattice, Code/lattice.hpp:35-51)
and the sum (one persistent launch per pass, replayed as a HIP graph).
``OverlappedBatches`` runs batch k's sum on a non-blocking stream while the host
prepares batch k+1. Batch k+1's lattices are queued on the device behind batch k,
into the second of two atom buffers. The host's synchronous copies go to the null
stream, which a non-blocking stream does not wait for, so neither side waits for
the other except at the end of each batch.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from . import backend as B


@dataclass
class BatchTimes:
    """Per batch: wall time since the previous batch ended, device time (the
    lattices of the next batch and this batch's sum), host time spent preparing
    the next batch while this one ran."""
    wall_ms: list = field(default_factory=list)
    device_ms: list = field(default_factory=list)
    host_ms: list = field(default_factory=list)
    atoms: list = field(default_factory=list)


class OverlappedBatches:
    """make_specs(k) -> list of geometry.SampleSpec (lattice samples: meta holds
    "lattice", "cells", "a0"); batches k = 0 .. nbatches-1 are summed over images
    +-img at precision prec.  on_batch(k, specs, batch, (x, y, z), (ux, uy, uz)),
    if given, is called once batch k is done, with its device arrays (valid until
    batch k + 2 starts: copy out what must outlive that)."""

    def __init__(self, ctx, make_specs, nbatches, img=1, prec=B.PREC_FP32C, tuning=None,
                 capacity=None):
        self.ctx, self.make_specs, self.nbatches = ctx, make_specs, int(nbatches)
        self.img, self.prec, self.tuning = int(img), int(prec), tuning
        self.capacity = capacity          # atoms per buffer (default 1.3 x batch 0)
        self.times = BatchTimes()

    def _prepare(self, k):
        specs = self.make_specs(k)
        return specs, B.Batch(self.ctx, specs, self.img, self.prec)

    def _lattices(self, specs, buf, stream):
        import torch
        with torch.cuda.stream(stream):
            off = 0
            for sp in specs:
                N = sp.meta["cells"]
                if isinstance(N, int):
                    N = (N, N, N)
                B.lattice_torch(self.ctx, sp.meta["lattice"], *N, sp.meta["a0"],
                                out=[t[off:off + sp.n] for t in buf[:3]])
                off += sp.n

    def run(self, on_batch=None):
        import torch
        dev = torch.device(f"cuda:{self.ctx.device}")
        s = torch.cuda.Stream(device=dev)     # PyTorch creates it non-blocking
        specs, bt = self._prepare(0)
        cap = self.capacity or int(bt.n * 1.3) + 1
        bufs = [[torch.empty(cap, dtype=torch.float64, device=dev) for _ in range(6)]
                for _ in range(2)]
        if bt.n > cap:
            raise ValueError("batch 0 larger than the buffer capacity")
        self._lattices(specs, bufs[0], s)
        s.synchronize()
        if self.tuning is None:
            self.tuning, _ = bt.autotune(*[t[:bt.n] for t in bufs[0]], stride=8,
                                         stream=s.cuda_stream)
            s.synchronize()
        cur = (specs, bt)
        t_prev = time.perf_counter()
        for k in range(self.nbatches):
            specs, bt = cur
            buf = bufs[k % 2]
            n = bt.n
            bt.set_tuning(self.tuning["block"], self.tuning["tile"], self.tuning["groups"])
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            bt.run_torch(*[t[:n] for t in buf], stream=s.cuda_stream)
            nxt, host = None, 0.0
            if k + 1 < self.nbatches:         # the next batch's host work, overlapped
                t0 = time.perf_counter()
                nxt = self._prepare(k + 1)
                if nxt[1].n > cap:
                    raise ValueError(f"batch {k + 1} larger than the buffer capacity")
                self._lattices(nxt[0], bufs[(k + 1) % 2], s)
                host = time.perf_counter() - t0
            e1.record(s)
            e1.synchronize()
            t_now = time.perf_counter()
            self.times.wall_ms.append((t_now - t_prev) * 1e3)
            self.times.device_ms.append(e0.elapsed_time(e1))
            self.times.host_ms.append(host * 1e3)
            self.times.atoms.append(n)
            t_prev = t_now
            if on_batch is not None:
                on_batch(k, specs, bt, [t[:n] for t in buf[:3]], [t[:n] for t in buf[3:]])
                t_prev = time.perf_counter()  # the caller's time is not the pipeline's
            bt.free()
            cur = nxt
        return self.times
