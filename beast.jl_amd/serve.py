"""Continuous-batching solve service over the resident-matrix C ABI.

Requests (geometry, operator, wavenumber, plane-wave right-hand sides) arrive
over time.  Worker threads each own a context; a worker takes the queued requests
it can start now, assembles the new ones' matrices together in shared launches
(bh_matrix_assemble_batch), then factors and solves each, while the other worker's
batch runs beside it on the device.  A request is admitted when its HBM footprint
fits the budget next to what is resident:
- 16 N^2 bytes for the FP64 matrix (the GMRES-IR residual operator);
- 8 N^2 bytes for its FP32 LU factors.

Factored matrices stay resident.  A later request on the same geometry,
operator and wavenumber only solves.  The least recently used matrices are
evicted when the budget requires it.  Each request's arithmetic runs in its
own buffers with a fixed reduction order (the batched assembly gives every
request the bits it gets alone), so its result is the same alone and in the mix.

The reference has no counterpart.  It assembles and factors per call
(src/solvers/lusolver.jl:9-39).
"""
import math
import threading
import time
from collections import OrderedDict

import numpy as np

from . import assembly as asm
from ._native import BHError
from .mesh import meshcube, meshellipsoid, meshsphere
from .operators import Identity, Maxwell3D, cross, curl, n as nvec
from .rt import raviartthomas


class Request:
    def __init__(self, rid, arrival, shape, dofs, kind, wavenumber, rhs):
        self.rid = rid
        self.arrival = arrival          # seconds after the stream start
        self.shape = shape              # "sphere" | "ellipsoid" | "cube"
        self.dofs = dofs                # requested RT dof count (nearest achievable mesh is used)
        self.kind = kind                # "efie" | "mfie"
        self.wavenumber = wavenumber
        self.rhs = rhs                  # list of (direction, polarization)
        self.n = None                   # dofs of the mesh used (set on arrival)
        self.result = None
        self.start = None
        self.done = None
        self.cache_hit = False
        self.info = {}

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def geometry(shape, dofs):
    """Mesh with about ``dofs`` RT functions: spheres and ellipsoids have N = 3(nv - 2), cubes N = 18 n^2."""
    if shape == "cube":
        nside = max(1, int(round(math.sqrt(dofs / 18.0))))
        return meshcube(nside)
    nv = max(8, int(round(dofs / 3.0 + 2)))
    if shape == "ellipsoid":
        return meshellipsoid(2.0, 1.0, 1.0, nv=nv)
    return meshsphere(1.0, nv=nv)


def random_rhs(rng, count):
    out = []
    for _ in range(count):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p = rng.normal(size=3)
        p -= p.dot(d) * d
        p /= np.linalg.norm(p)
        out.append((d, p))
    return out


def batched_requests(seed=4, count=64, nmin=500, nmax=8000, kinds=("efie",)):
    """The batched workload: all requests at t = 0, N log-uniform, k ~ U[0.5, 8], 1-4 right-hand sides."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        shape = ["sphere", "ellipsoid", "cube"][rng.integers(3)]
        dofs = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        k = float(rng.uniform(0.5, 8.0))
        kind = kinds[rng.integers(len(kinds))]
        reqs.append(Request(i, 0.0, shape, dofs, kind, k, random_rhs(rng, int(rng.integers(1, 5)))))
    return reqs


def poisson_requests(seed=5, rate=20.0, count=200, nmin=500, nmax=30000, kinds=("efie", "mfie"), repeat_prob=0.3):
    """Serving workload: Poisson arrivals; a share of requests repeat an earlier (geometry, k) with new RHS."""
    rng = np.random.default_rng(seed)
    reqs, t, seen = [], 0.0, []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        if seen and rng.uniform() < repeat_prob:
            shape, dofs, kind, k = seen[rng.integers(len(seen))]
        else:
            shape = ["sphere", "ellipsoid", "cube"][rng.integers(3)]
            dofs = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
            kind = kinds[rng.integers(len(kinds))]
            k = float(rng.uniform(0.5, 8.0))
            seen.append((shape, dofs, kind, k))
        reqs.append(Request(i, t, shape, dofs, kind, k, random_rhs(rng, int(rng.integers(1, 5)))))
    return reqs


class Server:
    """Continuous-batching solve service.

    ``workers`` threads each own a context (a HIP stream, and a second one for the FP32
    factorisation's look-ahead): two workers use four streams, the hardware queues of one
    GPU.  A worker takes, in arrival order, the queued requests it can start now: requests
    whose matrix is resident (cache hits, solved at once) and new requests admitted against
    the memory budget, up to ``max_batch`` of them and ``batch_dofs2`` summed N^2.  The new
    requests' matrices are assembled together (bh_matrix_assemble_batch: shared launches
    for their classification, near blocks and far tiles) and factored together
    (bh_matrix_factor_batch: shared launches for their panels, interchanges, block rows and
    trailing updates), then each is solved.
    While one worker assembles a batch the other factors and solves its own, so the device
    overlaps requests that arrived at different times.

    Admission counts ``footprint(N) = 24 N^2`` bytes per resident request: the FP64 matrix
    (16 N^2, the GMRES-IR residual operator) and its FP32 factors (8 N^2), against
    ``budget_bytes`` (250 GB of the 288 GB HBM by default: the rest is left to the
    factorisation workspace and the batches' near-field data).  Idle resident matrices are
    evicted least recently used first.
    """

    FOOTPRINT_BYTES_PER_N2 = 24.0

    def __init__(self, workers=2, budget_bytes=250e9, tol=1e-10, maxit=10, precision="fp32split", max_batch=16,
                 batch_dofs2=6.4e8, contexts=None):
        self.budget = budget_bytes
        self.tol = tol
        self.maxit = maxit
        self.precision = precision
        self.max_batch = max_batch
        self.batch_dofs2 = batch_dofs2
        self._lock = threading.Condition()
        self._geo = {}
        self._resident = OrderedDict()     # key -> (DeviceMatrix, bytes)
        self._busy = {}                    # key -> requests using or building it
        self._reserved = 0.0
        self._own = contexts is None
        self._contexts = list(contexts) if contexts is not None else [asm.Context(0) for _ in range(workers)]
        for c in self._contexts:   # evicted matrices free their memory (the budget counts only residents)
            c.pool(0)
        self.batches = []          # (builds, hits) per batch taken

    @classmethod
    def footprint(cls, n):
        return cls.FOOTPRINT_BYTES_PER_N2 * n * n

    def _space(self, shape, dofs):
        key = (shape, dofs)
        with self._lock:
            if key in self._geo:
                return self._geo[key]
        m = geometry(shape, dofs)
        sp = (m, raviartthomas(m))
        with self._lock:
            return self._geo.setdefault(key, sp)

    def _resident_bytes(self):
        return sum(b for _, b in self._resident.values())

    def _admit(self, nbytes):
        """Reserve memory for a new matrix, evicting idle resident ones (LRU); False if it cannot fit now."""
        if nbytes > self.budget:
            raise MemoryError(f"request needs {nbytes / 1e9:.1f} GB > budget {self.budget / 1e9:.1f} GB")
        while self._resident_bytes() + self._reserved + nbytes > self.budget:
            victim = next((k for k in self._resident if self._busy.get(k, 0) == 0), None)
            if victim is None:
                return False
            del self._resident[victim]
        self._reserved += nbytes
        return True

    @staticmethod
    def _key(req):
        return (req.shape, req.dofs, req.kind, req.wavenumber)

    @staticmethod
    def _operator(kind, k):
        """EFIE: T.  MFIE: n x K - Id/2 tested with RT, the reference's RT-tested MFIE
        (examples/nxmfie.jl:9-22; n x K = DoubleLayerRotatedMW3D, its default quadstrat
        DoubleNumQStrat(6,7), nxdbllayer.jl:23)."""
        if kind == "efie":
            return Maxwell3D.singlelayer(wavenumber=k)
        if kind == "mfie":
            return cross(nvec, Maxwell3D.doublelayer(wavenumber=k)) - 0.5 * Identity()
        raise ValueError(f"request kind {kind!r} is not served")

    @staticmethod
    def _rhs(X, kind, k, d, p, ctx):
        """EFIE: (n x E) x n.  MFIE: n x H with H = -1/(i mu omega) curl E (examples/nxmfie.jl:12-16),
        eps = mu = 1, omega = k."""
        E = Maxwell3D.planewave(direction=d, polarization=p, wavenumber=k)
        if kind == "efie":
            return asm.assemble_rhs(cross(cross(nvec, E), nvec), X, ctx=ctx)
        H = (-1.0 / (1j * k)) * curl(E)
        return asm.assemble_rhs(cross(nvec, H), X, ctx=ctx)

    def _take(self, queue):
        """Under the lock: the queued requests this worker starts now (see the class notes)."""
        builds, hits, work = [], [], 0.0
        claimed = set()
        for r in list(queue):
            if len(builds) + len(hits) >= self.max_batch:
                break
            key = self._key(r)
            if key in self._resident:
                self._resident.move_to_end(key)
                self._busy[key] = self._busy.get(key, 0) + 1
                r.cache_hit = True
                hits.append(r)
                queue.remove(r)
                continue
            if self._busy.get(key, 0) > 0 or key in claimed:   # being built: wait for it
                continue
            n = r.n
            if builds and work + float(n) * n > self.batch_dofs2:
                break
            if not self._admit(self.footprint(n)):
                break   # memory is full: later requests wait too (arrival order)
            self._busy[key] = 1
            claimed.add(key)
            work += float(n) * n
            builds.append(r)
            queue.remove(r)
        return builds, hits

    def _release(self, req, built=None, reserved=0.0):
        with self._lock:
            key = self._key(req)
            if built is not None:
                self._resident[key] = (built, reserved)
            self._reserved -= reserved
            self._busy[key] -= 1
            self._lock.notify_all()

    def _solve(self, req, M, ctx):
        m, X = self._space(req.shape, req.dofs)
        B = np.stack([self._rhs(X, req.kind, req.wavenumber, d, p, ctx) for d, p in req.rhs], axis=1)
        x = M.solve(B, tol=self.tol, maxit=self.maxit) if self.precision != "fp64" else M.solve(B)
        req.info = dict(M.stats or {})
        req.result = x
        req.done = time.perf_counter()

    def process_batch(self, builds, hits, ctx):
        """Build the new requests' matrices in shared launches, factor and solve each, and
        solve the cache hits."""
        t0 = time.perf_counter()
        for r in builds + hits:
            r.start = t0
        for r in hits:
            try:
                with self._lock:
                    M = self._resident[self._key(r)][0]
                self._solve(r, M, ctx)
            finally:
                self._release(r)
        if not builds:
            return
        spaces = [self._space(r.shape, r.dofs)[1] for r in builds]
        nbytes = [self.footprint(X.numfunctions()) for X in spaces]
        mats = None
        try:
            mats = asm.DeviceMatrix.assemble_batch(
                [(self._operator(r.kind, r.wavenumber), X) for r, X in zip(builds, spaces)], ctx=ctx)
        except BaseException:
            for r, nb in zip(builds, nbytes):   # give the reservations back
                self._release(r, None, nb)
            raise
        errors = []
        try:
            # shared launches for the batch's factorisations; a singular matrix is left
            # unfactored (its solve below reports it) while the others are factored
            asm.DeviceMatrix.factor_batch(mats, self.precision)
        except BHError:
            pass
        for r, M, nb in zip(builds, mats, nbytes):
            try:
                self._solve(r, M, ctx)
                self._release(r, M, nb)
            except BaseException as exc:
                self._release(r, None, nb)
                errors.append(exc)
        if errors:
            raise errors[0]

    def run(self, requests, realtime=True):
        """Serve ``requests`` (arrival times in seconds); returns the summary statistics."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        queue = []
        state = {"arrived": 0, "closed": False}
        finished, errors = [], []

        def dispatcher():
            for r in pending:
                if realtime:
                    delay = t0 + r.arrival - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                r.arrival = t0 + r.arrival if realtime else time.perf_counter()
                # the mesh and its RT space are prepared as the request arrives
                r.n = self._space(r.shape, r.dofs)[1].numfunctions()
                with self._lock:
                    queue.append(r)
                    state["arrived"] += 1
                    self._lock.notify_all()
            with self._lock:
                state["closed"] = True
                self._lock.notify_all()

        def worker(ctx):
            while True:
                with self._lock:
                    while True:
                        if errors:
                            return
                        try:
                            builds, hits = self._take(queue)
                        except BaseException as exc:   # e.g. a request larger than the budget
                            errors.append(exc)
                            self._lock.notify_all()
                            return
                        if builds or hits:
                            break
                        if state["closed"] and not queue:
                            return
                        self._lock.wait(0.05)
                    self.batches.append((len(builds), len(hits)))
                try:
                    self.process_batch(builds, hits, ctx)
                except BaseException as exc:   # reported to the caller after the join
                    with self._lock:
                        errors.append(exc)
                        self._lock.notify_all()
                    return
                with self._lock:
                    finished.extend(builds + hits)

        th = [threading.Thread(target=dispatcher)] + [threading.Thread(target=worker, args=(c,))
                                                       for c in self._contexts]
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t0
        if errors:
            raise errors[0]
        lat = np.array([r.latency for r in finished])
        served_at = np.sort([r.done for r in finished])
        return {"requests": len(finished), "wall_s": wall, "throughput_rps": len(finished) / wall,
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "cache_hits": sum(r.cache_hit for r in finished),
                "max_resident_gb": self._resident_bytes() / 1e9, "batches": len(self.batches),
                "mean_batch": float(np.mean([b + h for b, h in self.batches])) if self.batches else 0.0,
                "service_window_s": float(served_at[-1] - served_at[0]) if len(served_at) > 1 else 0.0}

    def close(self):
        self._resident.clear()
        if self._own:
            for c in self._contexts:
                c.close()
