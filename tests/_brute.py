"""Independent brute-force definition of the projector weights (test helper).

W_a[p, v] is the area of the unit voxel square v that falls inside the unit
detector strip of pixel p, i.e. the strip integral of the piecewise-constant
image.  It is computed here by clipping the square against the two
half-planes of the strip (Sutherland-Hodgman) and taking the polygon area
(shoelace formula) -- no trapezoid CDF, no support shortcut.  This pins the
oracle's closed-form footprint (projection_common.cuh:61-85) without
restating it.
"""
import numpy as np


def _clip(poly, nx, ny, off):
    """Keep the part of a convex polygon with nx*x + ny*y >= off."""
    out = []
    n = len(poly)
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        fp = nx * p[0] + ny * p[1] - off
        fq = nx * q[0] + ny * q[1] - off
        if fp >= 0:
            out.append(p)
        if (fp >= 0) != (fq >= 0):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def _area(poly):
    if len(poly) < 3:
        return 0.0
    s = 0.0
    for i in range(len(poly)):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % len(poly)]
        s += x0 * y1 - x1 * y0
    return abs(s) / 2.0


def weights(phi, SY, SX, PX, du=0.0):
    """Dense W of shape (PX, SY, SX) for one float32 angle phi; du translates
    the detector along its axis by du pixels (float32-rounded, as the library
    receives it)."""
    phi = float(np.float32(phi))
    c, s = np.cos(phi), np.sin(phi)
    xc, yc, pc = (SX - 1) / 2.0, (SY - 1) / 2.0, (PX - 1) / 2.0
    # detector coordinate of a point: (qx - xc) c + (qy - yc) s + pc - du
    shift = -xc * c - yc * s + pc - float(np.float32(du))
    W = np.zeros((PX, SY, SX))
    for iy in range(SY):
        for ix in range(SX):
            sq = [(ix - 0.5, iy - 0.5), (ix + 0.5, iy - 0.5), (ix + 0.5, iy + 0.5), (ix - 0.5, iy + 0.5)]
            for p in range(PX):
                lo = p - 0.5 - shift
                hi = p + 0.5 - shift
                poly = _clip(sq, c, s, lo)
                poly = _clip(poly, -c, -s, -hi)
                W[p, iy, ix] = _area(poly)
    return W


def forward(x_z0, angles, PX):
    """A x for a z_order-0 volume (SY, SX, H) in FP64, by dense weights."""
    SY, SX, H = x_z0.shape
    y = np.zeros((len(angles), PX, H))
    xf = x_z0.astype(np.float64)
    for a, phi in enumerate(angles):
        y[a] = np.einsum("pyx,yxh->ph", weights(phi, SY, SX, PX), xf)
    return y
