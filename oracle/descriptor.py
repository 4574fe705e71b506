"""CPU restatement of the learned bracket's descriptor (TEST INFRASTRUCTURE
ONLY: tests compare the device's descriptors with it; no product code may
import it).

Checks csrc/desc.hip (raster_build, k_raster_tris, descriptors,
k_descriptor) and the raster pixel size chosen in capi.cpp predict_cells.
There is no reference counterpart: the reference always runs the full
elevation search (mo_lradtopo_horayzon.cpp:490-555).

Every step is FP32 arithmetic in the device's operation order (numpy float32
operations are IEEE single precision, and the library is built with
-ffp-contract=off), so the raster and the sample positions are reproduced
exactly.  The one step that is not is the elevation angle: the device takes
atan2f of the bin's best sample (ROCm's atan2f, within a few ulp), this
restatement the FP64 atan2 rounded to FP32, so a BF16 descriptor value can
differ by one BF16 ulp where the FP32 angle sits at a BF16 rounding boundary.

Pinned by tests/test_descriptor_oracle.py against closed forms: the raster of
a plane is the plane at the pixel centres; seen from a point on a tilted
plane, every bin holds the arctangent of the plane's slope along the
azimuth; over a paraboloid bowl / dome the bin's value is the elevation of
its farthest / nearest sample, which fixes the ring radii; bins whose
samples all leave the raster hold -90 deg.
"""

from __future__ import annotations

import math

import numpy as np

N_AZ = 24
N_RINGS = 8
N_SAMPLES = 3
RING_INNER = 250.0
F = np.float32


def ring_radius(k: int) -> float:
    """Outer radius [m] of ring k (0 .. 7); ring -1 is the inner radius."""
    return RING_INNER if k < 0 else 500.0 * 80.0 ** (k / 7.0)


def tables():
    """Sample distances [24] (ring-major) and the azimuths' sin / cos [24],
    evaluated in FP64 and rounded to FP32 (desc.hip descriptors())."""
    dist = np.array([ring_radius(k - 1) * (ring_radius(k) / ring_radius(k - 1)) ** (s / N_SAMPLES)
                     for k in range(N_RINGS) for s in range(1, N_SAMPLES + 1)], np.float64)
    phi = np.arange(N_AZ) * (2.0 * 3.141592653589793 / N_AZ)
    return dist.astype(F), np.sin(phi).astype(F), np.cos(phi).astype(F)


def raster(vert32: np.ndarray, tri: np.ndarray) -> dict:
    """Max-height raster of the FP32 scene (capi.cpp predict_cells pixel
    size, desc.hip raster_build / k_raster_tris).  Returns z [ny][nx]
    (float32, -inf where no triangle covers a pixel centre), x0, y0, pix."""
    vert32 = np.asarray(vert32, F)
    tri = np.asarray(tri, np.int64)
    ntri = tri.shape[0]
    if vert32.shape[0]:
        b = [vert32[:, 0].min(), vert32[:, 1].min(), vert32[:, 0].max(), vert32[:, 1].max()]
    else:
        b = [F(0)] * 4
    w, h = F(b[2] - b[0]), F(b[3] - b[1])
    area = max(1.0, float(w) * float(h))
    pix = F(math.sqrt(area / max(1, ntri)))
    pix = max(pix, F(max(w, h) / F(4095.0)))
    pix = max(pix, F(1.0))
    nx, ny = int(w / pix) + 2, int(h / pix) + 2
    inv = F(F(1.0) / pix)
    z = np.full(nx * ny, -np.inf, F)
    if ntri:
        x0, y0 = F(b[0]), F(b[1])
        px = (vert32[tri, 0] - x0) * inv
        py = (vert32[tri, 1] - y0) * inv
        pz = vert32[tri, 2]
        det = (px[:, 1] - px[:, 0]) * (py[:, 2] - py[:, 0]) - (px[:, 2] - px[:, 0]) * (py[:, 1] - py[:, 0])
        with np.errstate(divide="ignore"):
            idet = F(1.0) / det
        ix0 = np.maximum(0, np.ceil(px.min(1)).astype(np.int64))
        ix1 = np.minimum(nx - 1, np.floor(px.max(1)).astype(np.int64))
        iy0 = np.maximum(0, np.ceil(py.min(1)).astype(np.int64))
        iy1 = np.minimum(ny - 1, np.floor(py.max(1)).astype(np.int64))
        keep = (det != 0) & (ix1 - ix0 <= 256) & (iy1 - iy0 <= 256)
        sel = np.nonzero(keep & (ix1 >= ix0) & (iy1 >= iy0))[0]
        e = F(-1e-4)
        wmax = int((ix1[sel] - ix0[sel]).max(initial=-1)) + 1
        hmax = int((iy1[sel] - iy0[sel]).max(initial=-1)) + 1
        for oy in range(hmax):
            for ox in range(wmax):
                t = sel[(ix0[sel] + ox <= ix1[sel]) & (iy0[sel] + oy <= iy1[sel])]
                if t.size == 0:
                    continue
                ix, iy = ix0[t] + ox, iy0[t] + oy
                qx, qy = ix.astype(F), iy.astype(F)
                p0x, p1x, p2x = px[t, 0], px[t, 1], px[t, 2]
                p0y, p1y, p2y = py[t, 0], py[t, 1], py[t, 2]
                b1 = ((qx - p0x) * (p2y - p0y) - (p2x - p0x) * (qy - p0y)) * idet[t]
                b2 = ((p1x - p0x) * (qy - p0y) - (qx - p0x) * (p1y - p0y)) * idet[t]
                b0 = F(1.0) - b1 - b2
                ok = (b0 >= e) & (b1 >= e) & (b2 >= e)
                hh = b0 * pz[t, 0] + b1 * pz[t, 1] + b2 * pz[t, 2]
                np.maximum.at(z, (iy * nx + ix)[ok], hh[ok])
    return dict(z=z.reshape(ny, nx), x0=F(b[0]), y0=F(b[1]), pix=pix, nx=nx, ny=ny)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 bit patterns, round to nearest even."""
    u = np.asarray(x, F).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def descriptors(ras: dict, org: np.ndarray, frame_n: np.ndarray, frame_north: np.ndarray,
                cells=None, chunk: int = 4096) -> np.ndarray:
    """BF16 descriptors [len(cells)][192] (desc.hip k_descriptor): for
    azimuth j and ring k at [j * 8 + k], the largest elevation angle (deg,
    times 0.1) of the ring's raster samples seen from the cell's origin
    along its own up vector; -90 deg where every sample is off the raster
    or on an empty pixel."""
    cells = np.arange(org.shape[0]) if cells is None else np.asarray(cells)
    dist, sn, cs = tables()
    z, nx, ny = ras["z"], ras["nx"], ras["ny"]
    x0, y0 = ras["x0"], ras["y0"]
    inv = F(F(1.0) / ras["pix"])
    out = np.empty((cells.size, N_AZ * N_RINGS), np.uint16)
    for a in range(0, cells.size, chunk):
        c = cells[a:a + chunk]
        o = np.asarray(org[c], F)
        n = np.asarray(frame_n[c]).astype(F)
        N = np.asarray(frame_north[c]).astype(F)
        ox, oy, oz = o[:, 0, None, None], o[:, 1, None, None], o[:, 2, None, None]
        nx3, ny3, nz3 = n[:, 0, None, None], n[:, 1, None, None], n[:, 2, None, None]
        Ex = N[:, 1] * n[:, 2] - N[:, 2] * n[:, 1]
        Ey = N[:, 2] * n[:, 0] - N[:, 0] * n[:, 2]
        hx = (cs[None, :] * N[:, 0, None] + sn[None, :] * Ex[:, None])[:, :, None]
        hy = (cs[None, :] * N[:, 1, None] + sn[None, :] * Ey[:, None])[:, :, None]
        qx = ox + dist[None, None, :] * hx
        qy = oy + dist[None, None, :] * hy
        ix = np.floor((qx - x0) * inv + F(0.5)).astype(np.int64)
        iy = np.floor((qy - y0) * inv + F(0.5)).astype(np.int64)
        inside = (ix >= 0) & (iy >= 0) & (ix < nx) & (iy < ny)
        zz = np.where(inside, z[np.where(inside, iy, 0), np.where(inside, ix, 0)], F(-np.inf))
        valid = np.isfinite(zz)
        vx, vy = qx - ox, qy - oy
        vz = np.where(valid, zz, F(0)) - oz
        vn = vx * nx3 + vy * ny3 + vz * nz3
        q = np.maximum(vx * vx + vy * vy + vz * vz - vn * vn, F(1e-6))
        ang = np.arctan2(vn.astype(np.float64), np.sqrt(q).astype(np.float64))
        ang = np.where(valid, ang, -np.inf).reshape(c.size, N_AZ, N_RINGS, N_SAMPLES).max(3)
        deg = np.where(np.isfinite(ang), ang.astype(F) * F(57.29577951308232), F(-90.0))
        out[a:a + c.size] = bf16_bits(deg * F(0.1)).reshape(c.size, N_AZ * N_RINGS)
    return out


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)
