"""Serving loop: continuous batching of A / AT / SIRT requests on the GPU.

Each request's result must be bitwise identical alone and in the mix, and
equal (bitwise) to the plain one-request-at-a-time operators.
"""
import pytest
import torch

from voxelproj_amd import serve

pytestmark = pytest.mark.gpu


def small_mix(n=18, seed=3):
    return serve.mixed_workload(n, rate=2000.0, seed=seed, sizes=(24, 40, 64),
                                heights=(8, 33, 64, 130), n_angles=(20, 45, 90), n_iter=4)


def test_mix_equals_alone_and_plain():
    reqs = small_mix()
    server = serve.Server()
    server.run(reqs)
    assert server.in_use == 0 and not server.active
    assert {r.kind for r in reqs} == {"A", "AT", "SIRT"}
    plain, _ = serve.run_plain(reqs)
    for r, p in zip(reqs, plain):
        assert r.t_done is not None and r.t_done >= r.arrival
        alone = serve.run_alone(r)
        assert torch.equal(alone, r.result), (r.rid, r.kind)
        assert torch.equal(p, r.result), (r.rid, r.kind)


def test_memory_admission_limits_concurrency():
    reqs = small_mix(10, seed=4)
    biggest = max(r.bytes_needed() for r in reqs)
    server = serve.Server(mem_budget=biggest)   # room for about one request at a time
    peak = []
    orig = server.tick

    def tick():
        peak.append(server.in_use)
        return orig()

    server.tick = tick
    server.run(reqs)
    assert max(peak) <= biggest
    assert all(r.t_done is not None for r in reqs)


def test_sirt_request_converges():
    reqs = [r for r in small_mix(30, seed=5) if r.kind == "SIRT"][:2]
    for r in reqs:
        r.n_iter = 20
    serve.Server().run(reqs)
    from voxelproj_amd import ops
    for r in reqs:
        res = ops.project(r.result, r.angles, r.PX) - r.data
        assert float(res.norm()) < 0.5 * float(r.data.norm())


def test_latency_and_throughput_reporting():
    reqs = small_mix(12, seed=6)
    server = serve.Server()
    server.run(reqs)
    st = serve.latency_stats(reqs)
    assert 0 <= st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert serve.throughput(reqs, server.makespan) > 0


def test_fifo_plain_baseline():
    """The unbatched first-in-first-out baseline serves the same arrivals: no
    request completes before it arrives, and the makespan covers the last
    arrival and all the work."""
    reqs = small_mix(8, seed=7)
    fifo = serve.run_plain_fifo(reqs)
    assert 0 <= fifo["p50_ms"] <= fifo["p99_ms"]
    assert fifo["makespan_s"] >= max(r.arrival for r in reqs)


@pytest.mark.parametrize("n_iter", [0, 1])
def test_sirt_with_zero_and_one_iteration_equals_plain(n_iter):
    # zero iterations: the weights are computed and the result is the zero
    # start volume, as recon.SIRT.iterate gives; one iteration: one update
    reqs = [r for r in small_mix(30, seed=9) if r.kind == "SIRT"][:2]
    for r in reqs:
        r.n_iter = n_iter
    serve.Server().run(reqs)
    plain, _ = serve.run_plain(reqs)
    for r, p in zip(reqs, plain):
        assert torch.equal(r.result, p), (r.rid, n_iter)
        if n_iter == 0:
            assert not bool(r.result.any())


def test_negative_iteration_count_is_rejected():
    r = small_mix(30, seed=9)[0]
    with pytest.raises(ValueError):
        serve.Request(r.rid, "SIRT", 0.0, r.data, r.angles, r.SY, r.SX, n_iter=-1)
