"""Host logic of the solve service (no GPU: contexts and the matrix build are stubbed)."""
import importlib

import pytest


class _FakeContext:
    def __init__(self, device=0):
        import threading
        self.device = device
        self.lock = threading.RLock()

    def close(self):
        pass

    def pool(self, limit_bytes=-1):
        return 0


@pytest.mark.timeout(60)
def test_failed_build_returns_its_reservation(bj, monkeypatch):
    """A batch whose assembly raises must give its memory reservations back, and run()
    ends with the error instead of waiting forever (room for one matrix at a time)."""
    sv = importlib.import_module("beast_jl_amd.serve")
    asm = importlib.import_module("beast_jl_amd.assembly")
    monkeypatch.setattr(asm, "Context", _FakeContext)
    calls = []

    def failing_assemble(requests, quadstrat=None, ctx=None):
        calls.extend(X.numfunctions() for _, X in requests)
        raise MemoryError("device allocation failed (stub)")

    monkeypatch.setattr(asm.DeviceMatrix, "assemble_batch", staticmethod(failing_assemble))
    reqs = sv.batched_requests(seed=3, count=3, nmin=300, nmax=400)
    n_max = max(r.dofs for r in reqs) + 10
    srv = sv.Server(workers=2, budget_bytes=1.5 * sv.Server.footprint(n_max))
    with pytest.raises(MemoryError):
        srv.run(reqs, realtime=False)
    assert 1 <= len(calls) <= 2      # the workers stop at the first error
    assert srv._reserved == 0 and all(v == 0 for v in srv._busy.values())


class _FakeMatrix:
    def __init__(self, n):
        self.n = n
        self.stats = {"iterations": 1, "inner": 1, "residual": 0.0}

    def factor(self, precision):
        return self

    def solve(self, B, tol=0.0, maxit=0):
        return B * 2.0


@pytest.mark.timeout(60)
def test_batches_share_builds_and_respect_the_budget(bj, monkeypatch):
    """The scheduler's batches (device calls stubbed): new requests are built together up to
    max_batch, every request is served once, repeats of a resident (geometry, k) are solved
    without a build, and resident bytes never exceed the budget."""
    import numpy as np
    sv = importlib.import_module("beast_jl_amd.serve")
    asm = importlib.import_module("beast_jl_amd.assembly")
    monkeypatch.setattr(asm, "Context", _FakeContext)
    built = []

    def fake_batch(requests, quadstrat=None, ctx=None):
        built.append(len(requests))
        return [_FakeMatrix(X.numfunctions()) for _, X in requests]

    monkeypatch.setattr(asm.DeviceMatrix, "assemble_batch", staticmethod(fake_batch))
    factored = []
    monkeypatch.setattr(asm.DeviceMatrix, "factor_batch",
                        staticmethod(lambda mats, precision: factored.append(len(mats)) or mats))
    monkeypatch.setattr(sv.Server, "_rhs", staticmethod(lambda X, kind, k, d, p, ctx: np.ones(X.numfunctions())))
    reqs = sv.poisson_requests(seed=5, rate=1000.0, count=40, nmin=300, nmax=600, repeat_prob=0.4)
    budget = 6 * sv.Server.footprint(620)
    srv = sv.Server(workers=2, budget_bytes=budget, max_batch=8)
    peak = []
    orig = srv._admit

    def admit(nbytes):
        ok = orig(nbytes)
        peak.append(srv._resident_bytes() + srv._reserved)
        return ok

    srv._admit = admit
    st = srv.run(reqs, realtime=False)
    assert st["requests"] == 40 and all(r.result is not None for r in reqs)
    distinct = len({(r.shape, r.dofs, r.kind, r.wavenumber) for r in reqs})
    assert sum(built) >= distinct and sum(built) + st["cache_hits"] == 40
    assert max(built) > 1 and max(built) <= 8
    assert factored == built            # each batch's matrices are factored together
    assert max(peak) <= budget
    assert all(np.array_equal(r.result, 2.0 * np.ones((r.n, len(r.rhs)))) for r in reqs)
