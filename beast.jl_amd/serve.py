"""Continuous-batching solve service over the resident-matrix C ABI.

Requests (geometry, operator, wavenumber, plane-wave right-hand sides) arrive
over time.  Worker threads each own a context (one HIP stream), so the
assembly, factorisation and solves of different requests overlap on the device.
A request is admitted when its HBM footprint fits the budget next to what is
resident:
- 16 N^2 bytes for the FP64 matrix (the GMRES-IR residual operator);
- 8 N^2 bytes for its FP32 LU factors.

Factored matrices stay resident.  A later request on the same geometry,
operator and wavenumber only solves.  The least recently used matrices are
evicted when the budget requires it.  Each request's arithmetic runs in its
own buffers with a fixed reduction order, so its result is the same alone and
in the mix.

The reference has no counterpart.  It assembles and factors per call
(src/solvers/lusolver.jl:9-39).
"""
import math
import threading
import time
from collections import OrderedDict

import numpy as np

from . import assembly as asm
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
    def __init__(self, workers=3, budget_bytes=250e9, tol=1e-10, maxit=10, precision="fp32"):
        self.nworkers = workers
        self.budget = budget_bytes
        self.tol = tol
        self.maxit = maxit
        self.precision = precision
        self._lock = threading.Condition()
        self._geo = {}
        self._resident = OrderedDict()     # key -> (DeviceMatrix, bytes)
        self._busy = {}                    # key -> refcount of requests using it
        self._reserved = 0.0
        self._contexts = [asm.Context(0) for _ in range(workers)]
        for c in self._contexts:   # evicted matrices free their memory (the budget counts only residents)
            c.pool(0)

    @staticmethod
    def footprint(n):
        return 16.0 * n * n + 8.0 * n * n

    def _space(self, shape, dofs):
        key = (shape, dofs)
        with self._lock:
            if key not in self._geo:
                m = geometry(shape, dofs)
                self._geo[key] = (m, raviartthomas(m))
            return self._geo[key]

    def _resident_bytes(self):
        return sum(b for _, b in self._resident.values())

    def _admit(self, key, nbytes):
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
    def _operator(kind, k):
        """EFIE: T.  MFIE: n x K - Id/2 tested with RT, the reference's RT-tested MFIE
        (examples/nxmfie.jl:9-22; n x K = DoubleLayerRotatedMW3D, its default quadstrat
        DoubleNumQStrat(6,7), nxdbllayer.jl:23)."""
        if kind == "efie":
            return Maxwell3D.singlelayer(wavenumber=k)
        if kind == "mfie":
            return cross(nvec, Maxwell3D.doublelayer(wavenumber=k)) - 0.5 * Identity()
        raise ValueError(f"request kind {kind!r} is not served")

    def process(self, req, ctx):
        m, X = self._space(req.shape, req.dofs)
        N = X.numfunctions()
        key = (req.shape, req.dofs, req.kind, req.wavenumber)
        nbytes = self.footprint(N)
        with self._lock:
            while True:
                if key in self._resident:
                    M = self._resident[key][0]
                    self._resident.move_to_end(key)
                    self._busy[key] = self._busy.get(key, 0) + 1
                    req.cache_hit = True
                    break
                if self._busy.get(key, 0) > 0:      # someone is building it: wait for it
                    self._lock.wait(0.01)
                    continue
                if self._admit(key, nbytes):
                    self._busy[key] = 1
                    M = None
                    break
                self._lock.wait(0.01)
        req.start = time.perf_counter()
        try:
            if M is None:
                op = self._operator(req.kind, req.wavenumber)
                try:
                    M = asm.DeviceMatrix.assemble(op, X, ctx=ctx)
                    M.factor(self.precision)
                except BaseException:
                    # give the reservation back, or later admissions wait for it forever
                    with self._lock:
                        self._reserved -= nbytes
                        self._lock.notify_all()
                    raise
                with self._lock:
                    self._reserved -= nbytes
                    self._resident[key] = (M, nbytes)
                    self._lock.notify_all()
            B = np.stack([self._rhs(X, req.kind, req.wavenumber, d, p, ctx) for d, p in req.rhs], axis=1)
            x = M.solve(B, tol=self.tol, maxit=self.maxit) if self.precision == "fp32" else M.solve(B)
            req.info = dict(M.stats or {})
            req.result = x
        finally:
            with self._lock:
                self._busy[key] -= 1
                self._lock.notify_all()
        req.done = time.perf_counter()

    @staticmethod
    def _rhs(X, kind, k, d, p, ctx):
        """EFIE: (n x E) x n.  MFIE: n x H with H = -1/(i mu omega) curl E (examples/nxmfie.jl:12-16),
        eps = mu = 1, omega = k."""
        E = Maxwell3D.planewave(direction=d, polarization=p, wavenumber=k)
        if kind == "efie":
            return asm.assemble_rhs(cross(cross(nvec, E), nvec), X, ctx=ctx)
        H = (-1.0 / (1j * k)) * curl(E)
        return asm.assemble_rhs(cross(nvec, H), X, ctx=ctx)

    def run(self, requests, realtime=True):
        """Serve ``requests`` (arrival times in seconds); returns the summary statistics."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        queue = []
        qlock = threading.Condition()
        finished = []

        def dispatcher():
            for r in pending:
                if realtime:
                    delay = t0 + r.arrival - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                r.arrival = t0 + r.arrival if realtime else time.perf_counter()
                with qlock:
                    queue.append(r)
                    qlock.notify()
            with qlock:
                queue.append(None)
                qlock.notify_all()

        errors = []

        def worker(ctx):
            while True:
                with qlock:
                    while not queue:
                        qlock.wait()
                    r = queue[0]
                    if r is None:
                        return
                    queue.pop(0)
                try:
                    self.process(r, ctx)
                except BaseException as exc:   # reported to the caller after the join
                    errors.append(exc)
                    continue
                finished.append(r)

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
        return {"requests": len(finished), "wall_s": wall, "throughput_rps": len(finished) / wall,
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "cache_hits": sum(r.cache_hit for r in finished),
                "max_resident_gb": self._resident_bytes() / 1e9}

    def close(self):
        self._resident.clear()
        for c in self._contexts:
            c.close()
