"""Continuous-batching service: results alone vs in the mix, residency, admission (needs an MI355X)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _serve(bj):
    import importlib
    return importlib.import_module("beast_jl_amd.serve")


def test_same_alone_and_in_mix(bj):
    sv = _serve(bj)
    reqs = sv.batched_requests(seed=11, count=10, nmin=300, nmax=1500)
    srv = sv.Server(workers=3, tol=1e-12)
    stats = srv.run(reqs, realtime=False)
    srv.close()
    assert stats["requests"] == 10
    # each request again, alone, on a fresh single-worker server
    for r in reqs[:4]:
        alone = sv.Request(r.rid, 0.0, r.shape, r.dofs, r.kind, r.wavenumber, r.rhs)
        s1 = sv.Server(workers=1, tol=1e-12)
        s1.run([alone], realtime=False)
        s1.close()
        assert np.array_equal(alone.result, r.result), r.rid
        assert alone.info["residual"] <= 1e-12


def test_residency_and_percentiles(bj):
    sv = _serve(bj)
    reqs = sv.poisson_requests(seed=5, rate=50.0, count=24, nmin=300, nmax=1200, repeat_prob=0.5)
    srv = sv.Server(workers=3, tol=1e-10)
    stats = srv.run(reqs)
    srv.close()
    assert stats["requests"] == 24 and stats["cache_hits"] > 0
    assert 0 < stats["p50_s"] <= stats["p99_s"]
    for r in reqs:
        assert r.info["residual"] <= 1e-10


def test_admission_budget(bj):
    sv = _serve(bj)
    reqs = sv.batched_requests(seed=3, count=6, nmin=800, nmax=900)
    n_max = max(r.dofs for r in reqs) + 10
    # room for about two matrices at a time: requests wait for memory, none fails
    srv = sv.Server(workers=3, budget_bytes=2.2 * sv.Server.footprint(n_max), tol=1e-10)
    stats = srv.run(reqs, realtime=False)
    srv.close()
    assert stats["requests"] == 6
    assert stats["max_resident_gb"] * 1e9 <= 2.2 * sv.Server.footprint(n_max)
    with pytest.raises(MemoryError):
        sv.Server(workers=1, budget_bytes=1e3).run(sv.batched_requests(seed=1, count=1), realtime=False)


def test_mfie_requests_alone_and_in_mix(bj):
    """MFIE requests (K + N/2, RHS from the magnetic plane wave) mixed with EFIE ones: each
    converges and gives the same bits alone and in the mix."""
    sv = _serve(bj)
    reqs = sv.batched_requests(seed=12, count=8, nmin=300, nmax=1200, kinds=("efie", "mfie"))
    assert {r.kind for r in reqs} == {"efie", "mfie"}
    srv = sv.Server(workers=3, tol=1e-12)
    srv.run(reqs, realtime=False)
    srv.close()
    for r in [q for q in reqs if q.kind == "mfie"][:2]:
        assert r.info["residual"] <= 1e-12
        alone = sv.Request(r.rid, 0.0, r.shape, r.dofs, r.kind, r.wavenumber, r.rhs)
        s1 = sv.Server(workers=1, tol=1e-12)
        s1.run([alone], realtime=False)
        s1.close()
        assert np.array_equal(alone.result, r.result), r.rid
