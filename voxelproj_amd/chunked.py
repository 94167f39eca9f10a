"""Projection of host-resident data larger than device memory, in H slabs.

The reference lists "data chunking for larger datasets" as future work
(README.md:65-68).  The operator acts on every H slice independently (the
in-plane footprint does not depend on H), so a volume or sinogram that lives
in host memory can be streamed through the device one slab of slices at a
time.  A detector translation dv along H couples neighbouring slices: a slab
of output slices then needs a halo of input slices, which the slab plan adds
(``vp_vshift`` reads slabs at global slice offsets).

Pipeline: two HIP streams at default priority alternate over the slabs.
Per slab the input slices are gathered into a pinned host buffer, copied to
the device, projected by the same HIP kernels as the in-memory path, and
copied back into a pinned buffer that the host adds into the output once the
stream's event has fired; the copy of one slab overlaps the projection of
the other.  Results equal the in-memory operator's: each output slice is
computed from the same input slices by the same kernels.

Semantics follow ``voxelproj.forward`` / ``backward``: outputs are
accumulated into (``out += op(in)``); with ``out=None`` a zero-filled numpy
array is returned.  Inputs are numpy arrays or CPU tensors, float32.
"""
import numpy as np
import torch

from . import _lib, ops


def _as_host(a, name):
    if isinstance(a, torch.Tensor):
        if a.is_cuda:
            raise ValueError(f"{name} must be host memory (numpy or CPU tensor)")
        a = a.numpy()
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name} must be a numpy array or CPU tensor")
    if a.ndim != 3:
        raise ValueError(f"{name} must be 3-D, got shape {a.shape}")
    return a


def _slab(a, z_order, i0, i1):
    """Slices [i0, i1) of a host array along H (a view)."""
    return a[:, :, i0:i1] if z_order == 0 else a[i0:i1]


def _halo(sv):
    """(floor(min dv), floor(max dv)) of a dv Shift, or (0, 0)."""
    if sv is None:
        return 0, 0
    v = sv.tensor.cpu().double()
    return int(np.floor(float(v.min()))), int(np.floor(float(v.max())))


def plan_slabs(H, slab, lo, hi):
    """Output slabs [h0, h1) of at most `slab` slices, each with the input range
    [i0, i1) it reads: output slice h reads input slices h + lo .. h + hi
    (clipped to [0, H)).  Slabs whose input range is empty are skipped."""
    out = []
    for h0 in range(0, H, slab):
        h1 = min(H, h0 + slab)
        i0, i1 = max(0, h0 + lo), min(H, h1 + hi)
        if i1 > i0:
            out.append((h0, h1, i0, i1))
    return out


def _slab_height(H, per_slice, halo, budget):
    """Largest slab height whose device footprint (two streams) fits budget."""
    if budget is None:
        free, _ = torch.cuda.mem_get_info()
        budget = free // 2
    hs = int(budget // (2 * 3 * per_slice)) - halo
    if hs < 1:
        raise RuntimeError(f"device budget of {budget} bytes is too small for one slice")
    return min(H, hs)


class _Pipeline:
    """Two-stream slab pipeline with pinned staging buffers."""

    def __init__(self, device):
        self.device = device
        cur = torch.cuda.current_stream(device)
        self.streams = [torch.cuda.Stream(device), torch.cuda.Stream(device)]
        for s in self.streams:
            s.wait_stream(cur)  # angles and shifts were made on the current stream
        self.pin_in = [None, None]
        self.pin_out = [None, None]
        self.pending = [None, None]

    @staticmethod
    def _pinned(buf, shape):
        n = int(np.prod(shape))
        if buf is None or buf.numel() < n:
            buf = torch.empty(n, dtype=torch.float32, pin_memory=True)
        return buf, buf[:n].view(shape)

    def _drain(self, b, add):
        if self.pending[b] is not None:
            ev, view, key = self.pending[b]
            ev.synchronize()
            add(key, view.numpy())
            self.pending[b] = None

    def run(self, jobs, fetch, compute, add):
        """jobs: list of keys; fetch(key) -> host array slab; compute(dev) ->
        device result; add(key, host result) accumulates into the output."""
        for i, key in enumerate(jobs):
            b = i % 2
            self._drain(b, add)
            src = fetch(key)
            self.pin_in[b], pin = self._pinned(self.pin_in[b], src.shape)
            pin.copy_(torch.from_numpy(src))
            st = self.streams[b]
            with torch.cuda.stream(st):
                dev = pin.to(self.device, non_blocking=True)
                res = compute(key, dev)
                self.pin_out[b], pout = self._pinned(self.pin_out[b], tuple(res.shape))
                pout.copy_(res, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(st)
            self.pending[b] = (ev, pout, key)
        for b in (0, 1):
            self._drain(b, add)
        for s in self.streams:
            s.synchronize()


def _vshift(z_order, transpose, src, dst, A, PX, in_off, out_off, sv):
    H_in = src.shape[2] if z_order == 0 else src.shape[0]
    H_out = dst.shape[2] if z_order == 0 else dst.shape[0]
    lib = _lib.load()
    rc = lib.vp_vshift(z_order, transpose, src.data_ptr(), dst.data_ptr(), A, PX, H_in, in_off,
                       H_out, out_off, sv.tensor.data_ptr(), 0,
                       torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, "vp_vshift")


def forward(x, angles, y=None, PX=None, z_order=0, du=None, dv=None, max_device_bytes=None,
            device="cuda"):
    """y (+)= A x for a host-resident volume x; returns y (numpy, float32)."""
    x = _as_host(x, "x")
    SY, SX, H = ops.vol_dims(x.shape, z_order)
    PX = max(SY, SX) if PX is None else int(PX)
    device = torch.device(device)
    a = ops.device_angles(angles, device)
    A = a.numel()
    shape = ops.sino_shape(A, PX, H, z_order)
    if y is None:
        y = np.zeros(shape, dtype=np.float32)
    else:
        y = _as_host(y, "y")
        if tuple(y.shape) != shape:
            raise ValueError(f"y must have shape {shape}, got {y.shape}")
    su, sv = ops.make_shifts(du, dv, A, device)
    kmin, kmax = _halo(sv)
    lo, hi = (kmin, kmax + 1) if sv is not None else (0, 0)
    hs = _slab_height(H, 4 * (SY * SX + 2 * A * PX), hi - lo, max_device_bytes)
    jobs = plan_slabs(H, hs, lo, hi)

    def fetch(key):
        return _slab(x, z_order, key[2], key[3])

    def compute(key, dev):
        h0, h1, i0, i1 = key
        y0 = ops.project(dev, a, PX, z_order, du=su)
        if sv is None:
            return y0
        ys = torch.empty(ops.sino_shape(A, PX, h1 - h0, z_order), device=device)
        _vshift(z_order, 0, y0, ys, A, PX, i0, h0, sv)
        return ys

    def add(key, res):
        _slab(y, z_order, key[0], key[1])[...] += res

    _Pipeline(device).run(jobs, fetch, compute, add)
    return y


def backward(y, angles, x=None, SY=None, SX=None, z_order=0, du=None, dv=None,
             max_device_bytes=None, device="cuda"):
    """x (+)= AT y for a host-resident sinogram y; returns x (numpy, float32)."""
    y = _as_host(y, "y")
    A, PX, H = ops.sino_dims(y.shape, z_order)
    device = torch.device(device)
    a = ops.device_angles(angles, device)
    if a.numel() != A:
        raise ValueError(f"{a.numel()} angles for a sinogram with {A} angle rows")
    SY = PX if SY is None else int(SY)
    SX = PX if SX is None else int(SX)
    shape = ops.vol_shape(SY, SX, H, z_order)
    if x is None:
        x = np.zeros(shape, dtype=np.float32)
    else:
        x = _as_host(x, "x")
        if tuple(x.shape) != shape:
            raise ValueError(f"x must have shape {shape}, got {x.shape}")
    su, sv = ops.make_shifts(du, dv, A, device)
    kmin, kmax = _halo(sv)
    # volume slice j reads sinogram slices j - k - 1 .. j - k
    lo, hi = (-kmax - 1, -kmin + 1) if sv is not None else (0, 0)
    hs = _slab_height(H, 4 * (2 * SY * SX + 2 * A * PX), hi - lo, max_device_bytes)
    jobs = plan_slabs(H, hs, lo, hi)

    def fetch(key):
        return _slab(y, z_order, key[2], key[3])

    def compute(key, dev):
        h0, h1, i0, i1 = key
        g = dev
        if sv is not None:
            g = torch.empty(ops.sino_shape(A, PX, h1 - h0, z_order), device=device)
            _vshift(z_order, 1, dev, g, A, PX, i0, h0, sv)
        return ops.backproject(g, a, SY, SX, z_order, du=su)

    def add(key, res):
        _slab(x, z_order, key[0], key[1])[...] += res

    _Pipeline(device).run(jobs, fetch, compute, add)
    return x
