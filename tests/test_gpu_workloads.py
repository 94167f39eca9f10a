"""The batched, differentiable and served forms at their benchmark workloads.

* uneven batch: 48 volumes, SX = SY in {96 .. 384}, H in {32 .. 256},
  n_angles in {60 .. 1024} with jittered angles, half N(0,1), half ellipsoid
  phantoms (batch.uneven_workload), through grouped persistent launches and
  a graph-captured A -> AT pair;
* least squares: a 256^3 ellipsoid phantom, 360 angles, b = A x_true + 1 %
  Gaussian noise, 100 iterations of gradient descent and of SIRT;
* serving: a Poisson stream of A / AT / SIRT-10 requests with SX = SY in
  {128, 256, 512}, H in {64 .. 512} and 90 .. 1440 angles
  (serve.mixed_workload).

Each checks its results against the plain operator (bitwise) and, on a few
H slices per instance, against the FP64 oracle (rel-L2 <= 1e-6; slices are
independent, so a slice of the oracle is exact for that slice).
"""
import numpy as np
import pytest
import torch

import oracle as vpo
from voxelproj_amd import autograd, batch, ops, recon, serve

pytestmark = pytest.mark.gpu
DEV = "cuda"
REL_L2 = 1e-6


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    d = np.linalg.norm(np.asarray(a, np.float64) - b)
    return float(d / nb) if nb > 0 else (0.0 if d == 0 else float("inf"))


def _slices(H):
    # the first, middle and last slice (a phantom's outer slices are empty:
    # then the result must be exactly zero)
    return sorted({0, H // 2, H - 1})


def _check_forward_slices(x, ang, y, what):
    xn, an, yn = x.cpu().numpy(), ang.cpu().numpy(), y.cpu().numpy()
    for h in _slices(x.shape[2]):
        ref = vpo.forward_f64(xn, an, y.shape[1], 0, h, h + 1)[:, :, h]
        e = rel_l2(yn[:, :, h], ref)
        assert e <= REL_L2, f"{what} A slice {h}: rel-L2 {e:.2e}"


def _check_backward_slices(y, ang, x, what):
    yn, an, xn = y.cpu().numpy(), ang.cpu().numpy(), x.cpu().numpy()
    for h in _slices(y.shape[2]):
        ref = vpo.backward_f64(yn, an, x.shape[0], x.shape[1], 0, h, h + 1)[:, :, h]
        e = rel_l2(xn[:, :, h], ref)
        assert e <= REL_L2, f"{what} AT slice {h}: rel-L2 {e:.2e}"


def test_uneven_batch_of_48():
    inst = batch.uneven_workload(48, seed=0, device=DEV)
    xs = [x for x, _ in inst]
    angs = [a for _, a in inst]
    ys = [torch.empty((a.numel(), max(x.shape[0], x.shape[1]), x.shape[2]), device=DEV)
          for x, a in inst]
    xb = [torch.empty_like(x) for x in xs]
    pa = batch.GroupedPlan("A", xs, ys, angs)
    pb = batch.GroupedPlan("AT", ys, xb, angs)
    graph = batch.capture_pair(pa, pb)
    for rep in range(2):
        for t in ys + xb:
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for i, (x, a, y, b) in enumerate(zip(xs, angs, ys, xb)):
            assert torch.equal(y, ops.project(x, a, y.shape[1])), (rep, i)
            assert torch.equal(b, ops.backproject(y, a, x.shape[0], x.shape[1])), (rep, i)
    for i, (x, a, y, b) in enumerate(zip(xs, angs, ys, xb)):
        _check_forward_slices(x, a, y, f"instance {i} {tuple(x.shape)} x {a.numel()}")
        _check_backward_slices(y, a, b, f"instance {i} {tuple(x.shape)} x {a.numel()}")


@pytest.fixture(scope="module")
def least_squares_256():
    N, A = 256, 360
    ang = torch.tensor(np.linspace(0, np.pi, A, endpoint=False).astype(np.float32), device=DEV)
    x_true = recon.ellipsoid_phantom(N, N, N, seed=11, device=DEV)
    b = ops.project(x_true, ang, N)
    g = torch.Generator(device=DEV).manual_seed(12)
    sigma = 0.01 * float(b.pow(2).mean().sqrt())
    b = b + sigma * torch.randn(b.shape, device=DEV, generator=g)
    return N, ang, x_true, b


def test_least_squares_gradient_at_256(least_squares_256):
    N, ang, x_true, b = least_squares_256
    x = (0.5 * x_true).requires_grad_(True)
    loss = 0.5 * (autograd.project(x, ang, N) - b).pow(2).sum()
    loss.backward()
    loss2, g = autograd.least_squares_gradient(x.detach(), b, ang, N)
    # the autograd backward is the same hand-written AT of the same residual
    assert torch.equal(x.grad, g)
    assert float(loss.detach()) == pytest.approx(float(loss2), rel=1e-5)
    _check_backward_slices(ops.project(x.detach(), ang, N) - b, ang, g, "gradient")


def test_gradient_descent_100_iterations_at_256(least_squares_256):
    N, ang, x_true, b = least_squares_256
    step = recon.lipschitz_step(ang, N, N, N, N, device=DEV)
    x = torch.zeros_like(x_true)
    losses = recon.gradient_descent(x, b, ang, 100, step, N)
    print(f"gradient descent 256^3 x 360: loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert all(l1 <= l0 for l0, l1 in zip(losses, losses[1:]))
    assert losses[-1] < 0.05 * losses[0]


def test_sirt_100_iterations_at_256(least_squares_256):
    N, ang, x_true, b = least_squares_256
    solver = recon.SIRT(ang, N, N, N, N, device=DEV)
    x = torch.zeros_like(x_true)
    res = [solver.residual(x, b)]
    for _ in range(10):
        solver.iterate(x, b, 10)
        res.append(solver.residual(x, b))
    print(f"SIRT 256^3 x 360, 100 iterations: residual {res[0]:.4e} -> {res[-1]:.4e}", flush=True)
    assert all(r1 < r0 for r0, r1 in zip(res, res[1:]))
    assert res[-1] < 0.1 * res[0]
    err0 = float(x_true.norm())
    assert float((x - x_true).norm()) < 0.5 * err0


def test_serving_mix_at_full_sizes():
    reqs = serve.mixed_workload(12, rate=200.0, seed=39, device=DEV)
    kinds = {r.kind for r in reqs}
    assert kinds == {"A", "AT", "SIRT"}, kinds
    assert max(r.SX for r in reqs) == 512 and max(r.H for r in reqs) == 512
    assert max(r.n_angles for r in reqs) == 1440
    server = serve.Server(device=DEV)
    server.run(reqs)
    assert server.in_use == 0 and not server.active
    plain, _ = serve.run_plain(reqs)
    for r, p in zip(reqs, plain):
        what = f"request {r.rid} {r.kind} {r.SY}x{r.SX}x{r.H} x {r.n_angles}"
        print(what, flush=True)
        assert r.t_done is not None and r.t_done >= r.arrival, what
        assert torch.equal(p, r.result), what
        assert torch.equal(serve.run_alone(r), r.result), what
        if r.kind == "A":
            _check_forward_slices(r.data, r.angles, r.result, what)
        elif r.kind == "AT":
            _check_backward_slices(r.data, r.angles, r.result, what)
        else:
            res = ops.project(r.result, r.angles, r.PX) - r.data
            assert float(res.norm()) < 0.5 * float(r.data.norm()), what
