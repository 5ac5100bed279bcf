"""Continuous-batching serving of elastic solves on one GPU.

Requests (meshes) arrive over time.  A request is admitted when the HBM it
will hold fits the budget.  A pool of host threads then prepares it: the
reference's set-up (RCM, numeq, CSRC), the device layout, GPU assembly and
the upload of b.  Prepared systems join a running batch (libelastic_hip's
elastic_batch_*) as soon as a slot is free.  Every step advances all active
systems by a fixed number of PCG iterations in one HIP-graph replay, and
finished systems leave the batch at the next poll.  So requests that arrive
while others are iterating start without waiting for them, and their host
and device work overlap with the running batch.  Latency of a request =
completion (x on the host) minus arrival.

There is no counterpart in the reference: its solver call serves one system
(Solv.f:13-440).  Each instance runs the reference's PCG (Iters.f:252-524)
in FP64 or with the mixed-precision operator, with Jacobi or node-block
Jacobi (block_precond / op_block_precond, Precond.f:163-540).  The server
picks both per request (`request_policy`) unless it is given a fixed
precision.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import FP64, MIXED, ElasticError, Stats


class Batch:
    """elastic_batch_*: a multi-instance PCG on one device."""

    def __init__(self, device: int = 0, max_slots: int = 64, iters_per_step: int = 16,
                 use_graph: bool = True):
        self._L = _lib.lib()
        self._h = C.c_void_p()
        rc = self._L.elastic_batch_create(C.byref(self._h), device, max_slots, iters_per_step,
                                          1 if use_graph else 0)
        if rc != _lib.OK:
            raise ElasticError(rc, "elastic_batch_create")
        self.max_slots = max_slots
        self._tags = np.zeros(max_slots, np.int64)
        self._stats = (Stats * max_slots)()

    def _chk(self, rc, allow=()):
        if rc != _lib.OK and rc not in allow:
            raise ElasticError(rc, self._L.elastic_batch_last_error(self._h).decode())
        return rc

    def submit(self, ctx, precision: int = MIXED, tol: float = 1e-10, maxit: int = 20000,
               tag: int = 0) -> bool:
        """Admits a prepared context; False when every slot is taken."""
        rc = self._L.elastic_batch_submit(self._h, ctx._h, precision, tol, maxit, tag)
        return self._chk(rc, allow=(_lib.BUSY,)) == _lib.OK

    def step(self):
        self._chk(self._L.elastic_batch_step(self._h))

    def poll(self):
        """Waits for the last step; returns [(tag, stats dict)] of finished instances."""
        nd = C.c_int32()
        self._chk(self._L.elastic_batch_poll(self._h, self._tags, C.cast(self._stats, C.c_void_p),
                                             self.max_slots, C.byref(nd)))
        return [(int(self._tags[i]), self._stats[i].as_dict()) for i in range(nd.value)]

    def active(self) -> int:
        n = C.c_int32()
        self._chk(self._L.elastic_batch_active(self._h, C.byref(n)))
        return n.value

    def close(self):
        if self._h:
            self._L.elastic_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# 3-dof systems of at least this many nodes take the mixed-precision operator
POLICY_MIXED_NODES = 32768


def request_policy(ctx) -> tuple[int, int]:
    """(precision, preconditioner) for a prepared request, from its layout:

    * 3-dof node blocks and at least POLICY_MIXED_NODES nodes: MIXED (FP32
      values, and FP32 sums by the context's default) with Jacobi.  On such
      systems the FP32 operator costs few iterations (hex8 100^3: 1,436
      against 1,329) and halves the bytes of each product.
    * 2-dof node blocks (plane problems): FP64 with the 2x2 node-block
      Jacobi.  FP32 values cost these ill-conditioned systems 1.2-3.6 times
      the FP64 iterations (condition numbers of 1e6-1e7 amplify the
      rounding of the values), and the 2x2 blocks save about 10% of the
      iterations on the tria3 requests (none on quad4).
    * anything else: FP64 with Jacobi.

    Returns the batch_precond option value as the preconditioner (2 Jacobi,
    6 node blocks)."""
    info = ctx.layout_info()
    nb = info["nb"]
    if nb == 3 and ctx.neq // 3 >= POLICY_MIXED_NODES:
        return MIXED, 2
    if nb == 2:
        return FP64, 6
    return FP64, 2


def footprint_bytes(mesh) -> int:
    """Device bytes a prepared request holds, from its mesh alone (admission
    happens before anything is allocated): CSRC (ia, ja, al, ad), the blocked
    FP64 and FP32 planes, layout words and merge lists, the external partials
    and 9 FP64 vectors, with nad estimated from the element type's stencil."""
    ndf = mesh.ndf
    neq = int(mesh.nnode) * ndf
    per_node_lower = {3: 3.5, 4: 4.5, 6: 7.5, 8: 13.5}.get(mesh.nen, 13.5)  # lower neighbours
    nad = int(neq * (per_node_lower * ndf + 1))
    return int(4 * neq + 12 * nad + 12 * nad + nad // 2 + 16 * neq + 9 * 8 * neq + 64 * 1024 * 8)


@dataclass
class Request:
    k: int
    arrival: float                 # seconds after the stream starts
    make_mesh: object              # callable -> Mesh
    footprint: int = 0
    neq: int = 0
    t_prep: float = 0.0
    t_ready: float = 0.0
    t_submit: float = 0.0
    t_done: float = 0.0
    precision: int = 0
    precond: int = 0
    tuned: bool = False
    stats: dict = field(default_factory=dict)
    x: np.ndarray | None = None


def _prepare(req: Request, device: int):
    from .system import System

    mesh = req.make_mesh()
    s = System(mesh, device=device)
    s.ctx.upload_rhs(s.b)
    return s


class Server:
    """Continuous batching over a request stream (see the module text).

    mode "batched": prepared systems share one Batch.
    mode "sequential": each prepared system is solved alone with
    elastic_solve_resident, in arrival order (the plain drop-in path).
    precision None (the default): precision and preconditioner per request
    by request_policy; FP64 or MIXED: that precision with Jacobi for every
    request.
    autotune True: the first request of each size class (node block size,
    precision and the bit length of neq) runs elastic_autotune before its
    solve; later requests of the class take the tile size and lanes it chose.
    """

    def __init__(self, device: int = 0, max_slots: int = 64, iters_per_step: int = 16,
                 hbm_budget: float = 64e9, prep_workers: int = 8, precision: int | None = None,
                 tol: float = 1e-10, maxit: int = 50000, mode: str = "batched",
                 keep_x: bool = False, progress=None, autotune: bool = False):
        self.device = device
        self.max_slots = max_slots
        self.iters_per_step = iters_per_step
        self.hbm_budget = hbm_budget
        self.prep_workers = prep_workers
        self.precision = precision
        self.tol = tol
        self.maxit = maxit
        self.mode = mode
        self.keep_x = keep_x
        self.progress = progress  # a file (e.g. sys.stderr) for a status line every 5 s
        self.autotune = autotune
        self.tuned = {}           # size class -> (tile size, lanes)

    def _tune(self, ctx, prec: int):
        """Layout of the request's size class (see the class text)."""
        nb = ctx.layout_info()["nb"]
        if not self.autotune or nb == 0:    # the gather product has no tiles
            return
        tile_opt = "tile_blocks" if nb > 1 else "tile_nnz"
        lanes_opt = "lanes32" if prec == MIXED else "lanes64"
        key = (nb, prec, int(ctx.neq).bit_length())
        hit = self.tuned.get(key)
        if hit is None:
            tile, _ = ctx.autotune(precision=prec, reps=16)
            self.tuned[key] = (tile, ctx.get_option(lanes_opt))
            return
        tile, lanes = hit
        if ctx.get_option(tile_opt) != tile:
            ctx.set_option(tile_opt, tile)
        ctx.set_option(lanes_opt, lanes)

    def run(self, requests: list[Request], realtime: bool = True) -> dict:
        """Serves the stream; realtime=False makes every request arrive at 0."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if not realtime:
                r.arrival = 0.0
        pending = list(reqs)            # not yet admitted (arrival order)
        preparing = {}                  # future -> request
        ready = []                      # (request, system) prepared, waiting for a slot
        running = {}                    # tag -> (request, system)
        done = []
        reserved = 0
        pool = cf.ThreadPoolExecutor(max_workers=self.prep_workers)
        batch = Batch(self.device, self.max_slots, self.iters_per_step) if self.mode == "batched" else None
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def finish(req, sysm, st):
            req.t_done = now()
            req.stats = st
            if self.keep_x:
                req.x = sysm.ctx.download_x()
            sysm.ctx.close()
            done.append(req)

        last = 0.0
        try:
            while len(done) < len(reqs):
                t = now()
                if self.progress is not None and t - last >= 5.0:
                    last = t
                    print("serving %.1f s: done %d/%d, running %d, ready %d, preparing %d, "
                          "waiting %d" % (t, len(done), len(reqs), len(running), len(ready),
                                          len(preparing), len(pending)),
                          file=self.progress, flush=True)
                # admission by HBM footprint, in arrival order
                while pending and pending[0].arrival <= t:
                    r = pending[0]
                    if r.footprint == 0:
                        r.footprint = footprint_bytes(r.make_mesh.mesh_hint()) \
                            if hasattr(r.make_mesh, "mesh_hint") else 0
                    if reserved + r.footprint > self.hbm_budget and (preparing or ready or running):
                        break
                    pending.pop(0)
                    reserved += r.footprint
                    r.t_prep = t
                    preparing[pool.submit(_prepare, r, self.device)] = r
                # prepared systems
                for f in [f for f in preparing if f.done()]:
                    r = preparing.pop(f)
                    s = f.result()
                    r.neq = s.neq
                    r.t_ready = now()
                    ready.append((r, s))
                if self.mode == "sequential":
                    if ready:
                        r, s = ready.pop(0)
                        r.t_submit = now()
                        prec = self.precision if self.precision is not None else FP64
                        self._tune(s.ctx, prec)
                        st = s.ctx.solve_resident(precision=prec, tol=self.tol,
                                                  maxit=self.maxit)
                        finish(r, s, st)
                        reserved -= r.footprint
                        continue
                else:
                    while ready:
                        r, s = ready[0]
                        if self.precision is None:
                            prec, pre = request_policy(s.ctx)
                            s.ctx.set_option("batch_precond", pre)
                        else:
                            prec, pre = self.precision, 2
                        if not r.tuned:
                            self._tune(s.ctx, prec)
                            r.tuned = True
                        if not batch.submit(s.ctx, prec, self.tol, self.maxit, r.k):
                            break
                        ready.pop(0)
                        r.t_submit = now()
                        r.precision, r.precond = prec, pre
                        running[r.k] = (r, s)
                    if running:
                        batch.step()
                        for tag, st in batch.poll():
                            r, s = running.pop(tag)
                            finish(r, s, st)
                            reserved -= r.footprint
                        continue
                # idle: wait for the next arrival or a prepared system
                waits = [pending[0].arrival - now()] if pending else []
                if preparing:
                    cf.wait(list(preparing), timeout=max(0.0, min(waits)) if waits else None,
                            return_when=cf.FIRST_COMPLETED)
                elif waits and waits[0] > 0:
                    time.sleep(waits[0])
        finally:
            pool.shutdown(wait=True)
            if batch is not None:
                batch.close()
        makespan = now()
        lat = np.array([r.t_done - r.arrival for r in done])
        svc = np.array([r.t_done - r.t_prep for r in done])
        slv = np.array([r.t_done - r.t_submit for r in done])
        neq = np.array([r.neq for r in done], dtype=np.float64)
        return {
            "requests": len(done),
            "makespan_s": makespan,
            "solves_per_s": len(done) / makespan,
            "dof_per_s": float(neq.sum()) / makespan,
            # from arrival (queueing included), from admission (set-up,
            # assembly, waiting for a slot and the solve), from submission
            "latency_ms_p50": 1e3 * float(np.percentile(lat, 50)),
            "latency_ms_p99": 1e3 * float(np.percentile(lat, 99)),
            "service_ms_p50": 1e3 * float(np.percentile(svc, 50)),
            "service_ms_p99": 1e3 * float(np.percentile(svc, 99)),
            "solve_ms_p50": 1e3 * float(np.percentile(slv, 50)),
            "solve_ms_p99": 1e3 * float(np.percentile(slv, 99)),
            "precision_mix": {str(p): int(sum(getattr(r, "precision", None) == p for r in done))
                              for p in (FP64, MIXED)},
            "iters_mean": float(np.mean([r.stats["iters"] for r in done])),
            "maxit_hits": int(sum(r.stats["status"] != 0 for r in done)),
            "done": done,
        }


class MeshRequest:
    """Callable making request k of the plane-strain stream, with a cheap
    size hint for admission."""

    def __init__(self, k: int, nmin: int = 32, nmax: int = 512):
        self.k, self.nmin, self.nmax = k, nmin, nmax

    def __call__(self):
        from .mesh import serving_request

        return serving_request(self.k, nmin=self.nmin, nmax=self.nmax)

    def mesh_hint(self):
        from .mesh import BASE_SEED, uniform01

        u = uniform01(BASE_SEED + 4, np.arange(5 * self.k, 5 * self.k + 1, dtype=np.uint64))
        n = self.nmin + min(int(u[0] * (self.nmax - self.nmin + 1)), self.nmax - self.nmin)

        class Hint:
            pass

        h = Hint()
        h.nnode, h.ndf, h.nen = (n + 1) ** 2, 2, 3 if self.k % 2 else 4
        return h


def request_stream(nreq: int = 512, rate: float = 200.0, nmin: int = 32, nmax: int = 512,
                   seed: int = 20261206) -> list[Request]:
    """nreq requests of the plane-strain stream with exponential inter-arrival
    times of mean 1/rate seconds (a Poisson process)."""
    from .mesh import uniform01

    u = uniform01(seed, np.arange(nreq, dtype=np.uint64))
    gaps = -np.log1p(-u) / rate
    t = np.cumsum(gaps) - gaps[0]
    return [Request(k=k, arrival=float(t[k]), make_mesh=MeshRequest(k, nmin, nmax))
            for k in range(nreq)]
