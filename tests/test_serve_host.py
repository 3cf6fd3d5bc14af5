"""Host logic of the solve service (no GPU: contexts and the matrix build are stubbed)."""
import importlib

import pytest


class _FakeContext:
    def __init__(self, device=0):
        self.device = device

    def close(self):
        pass

    def pool(self, limit_bytes=-1):
        return 0


@pytest.mark.timeout(60)
def test_failed_build_returns_its_reservation(bj, monkeypatch):
    """A request whose assembly raises must give its memory reservation back: with room for
    one matrix at a time, the second request is admitted after the first fails, and run()
    ends with the first error instead of waiting forever."""
    sv = importlib.import_module("beast_jl_amd.serve")
    asm = importlib.import_module("beast_jl_amd.assembly")
    monkeypatch.setattr(asm, "Context", _FakeContext)
    calls = []

    def failing_assemble(op, X, quadstrat=None, ctx=None):
        calls.append(X.numfunctions())
        raise MemoryError("device allocation failed (stub)")

    monkeypatch.setattr(asm.DeviceMatrix, "assemble", staticmethod(failing_assemble))
    reqs = sv.batched_requests(seed=3, count=3, nmin=300, nmax=400)
    n_max = max(r.dofs for r in reqs) + 10
    srv = sv.Server(workers=2, budget_bytes=1.5 * sv.Server.footprint(n_max))
    with pytest.raises(MemoryError):
        srv.run(reqs, realtime=False)
    assert len(calls) == 3
    assert srv._reserved == 0
