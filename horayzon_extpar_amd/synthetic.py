"""Synthetic ICON-like regional grids and seeded terrains.

The reference consumes ICON grid files (clon, clat, vlon, vlat, cells_of_vertex
with 1-based indices and -1 fill, shifted to 0-based / -2 fill by the caller,
verification.py:82-104, development.py:188-201) plus the EXTPAR field
``topography_c``.  No such files ship with the reference, so every workload of
this package is generated here, deterministically from a seed:

* grid: an equilateral triangular lattice of edge ``edge_km`` laid out in the
  tangent plane at the box centre and mapped onto the sphere by the inverse
  gnomonic projection.  Cells are the lattice triangles, ``clon/clat`` are the
  spherical circumcentres, ICON vertices are the lattice nodes and
  ``cells_of_vertex`` is int32 ``(6, nvert)`` with ``-2`` fill at the border
  (the layout horizon_svf.pyx:46-48 documents);
* terrain: a Gaussian random field synthesised spectrally on a regular grid at
  half the edge length (amplitude ~ k^(-beta/2) for a power spectrum k^-beta),
  scaled to a target standard deviation and sampled bilinearly at the cell
  circumcentres.

All arrays are numpy; nothing here touches a GPU.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

RAD_EARTH = 6371229.0  # ICON/COSMO earth radius [m] (mo_lradtopo_horayzon.cpp:584)


@dataclass
class Domain:
    """One ICON-like domain in the units the reference call expects."""

    clon: np.ndarray  # f64 [ncell] rad
    clat: np.ndarray  # f64 [ncell] rad
    hsurf: np.ndarray  # f64 [ncell] m
    vlon: np.ndarray  # f64 [nvert] rad
    vlat: np.ndarray  # f64 [nvert] rad
    cells_of_vertex: np.ndarray  # i32 [6, nvert], -2 fill
    meta: dict = field(default_factory=dict)
    # planar (tangent-plane) coordinates of the cell centres [m]; used by the
    # terrain recipes and by tests that need a closed-form terrain
    cx: np.ndarray | None = None
    cy: np.ndarray | None = None

    @property
    def ncell(self) -> int:
        return int(self.clon.size)

    @property
    def nvert(self) -> int:
        return int(self.vlon.size)

    def args(self):
        """Positional array arguments of horizon_svf_comp_py."""
        return (self.clon, self.clat, self.hsurf, self.vlon, self.vlat,
                self.cells_of_vertex)


# ---------------------------------------------------------------------------
# Lattice and projection
# ---------------------------------------------------------------------------

def _tangent_basis(lon0: float, lat0: float):
    up = np.array([math.cos(lat0) * math.cos(lon0),
                   math.cos(lat0) * math.sin(lon0), math.sin(lat0)])
    east = np.array([-math.sin(lon0), math.cos(lon0), 0.0])
    north = np.array([-math.sin(lat0) * math.cos(lon0),
                      -math.sin(lat0) * math.sin(lon0), math.cos(lat0)])
    return up, east, north


def _plane_to_unit(x, y, lon0, lat0):
    """Inverse gnomonic projection: tangent-plane (x, y) [m] -> unit vectors."""
    up, east, north = _tangent_basis(lon0, lat0)
    p = (RAD_EARTH * up[None, :] + x[:, None] * east[None, :]
         + y[:, None] * north[None, :])
    return p / np.linalg.norm(p, axis=1, keepdims=True)


def _unit_to_lonlat(u):
    lon = np.arctan2(u[:, 1], u[:, 0])
    lat = np.arcsin(np.clip(u[:, 2], -1.0, 1.0))
    return lon, lat


def triangular_lattice(width_m: float, height_m: float, edge_m: float,
                       x0: float = 0.0, y0: float = 0.0):
    """Equilateral lattice nodes and triangles covering a width x height box
    centred on (x0, y0).  Returns (xv, yv, tri[ntri, 3]) with node numbering
    row-major."""
    dy = edge_m * math.sqrt(3.0) / 2.0
    nx = int(math.floor(width_m / edge_m + 1e-9)) + 1
    ny = int(math.floor(height_m / dy + 1e-9)) + 1
    if nx < 2 or ny < 2:
        raise ValueError("box smaller than one lattice cell")
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    xv = (ii + 0.5 * (jj % 2)) * edge_m
    yv = jj * dy
    xv = xv - 0.5 * ((nx - 1) * edge_m + 0.5 * edge_m) + x0
    yv = yv - 0.5 * (ny - 1) * dy + y0
    node = lambda i, j: j * nx + i  # noqa: E731
    i = np.arange(nx - 1)
    tris = []
    for j in range(ny - 1):
        if j % 2 == 0:
            a = np.stack([node(i, j), node(i + 1, j), node(i, j + 1)], 1)
            b = np.stack([node(i, j + 1), node(i + 1, j + 1), node(i + 1, j)], 1)
        else:
            a = np.stack([node(i, j), node(i + 1, j), node(i + 1, j + 1)], 1)
            b = np.stack([node(i, j + 1), node(i + 1, j + 1), node(i, j)], 1)
        # interleave so that neighbouring cells have neighbouring indices
        t = np.empty((2 * (nx - 1), 3), dtype=np.int64)
        t[0::2] = a
        t[1::2] = b
        tris.append(t)
    return xv.ravel(), yv.ravel(), np.concatenate(tris, 0)


def _cells_of_vertex(tri: np.ndarray, nvert: int, rng) -> np.ndarray:
    """(6, nvert) int32 adjacency with -2 fill; slot order shuffled per vertex
    (the reference sorts the adjacent cells by angle itself,
    mo_lradtopo_horayzon.cpp:246-261, so the slot order carries no meaning)."""
    ncell = tri.shape[0]
    v = tri.ravel()
    c = np.repeat(np.arange(ncell, dtype=np.int64), 3)
    order = np.argsort(v, kind="stable")
    v, c = v[order], c[order]
    start = np.searchsorted(v, np.arange(nvert))
    slot = np.arange(v.size) - start[v]
    if slot.max() > 5:
        raise ValueError("vertex with more than 6 adjacent cells")
    cov = np.full((6, nvert), -2, dtype=np.int32)
    if rng is not None:
        # random permutation of the 6 slots per vertex
        perm = np.argsort(rng.random((nvert, 6)), axis=1)
        slot = perm[v, slot]
    cov[slot, v] = c.astype(np.int32)
    return cov


def lattice_domain(width_km: float, height_km: float, edge_km: float,
                   lon0_deg: float = 8.5, lat0_deg: float = 46.5,
                   seed: int = 0, shuffle: bool = True) -> Domain:
    """ICON-like regional grid without terrain (hsurf = 0)."""
    lon0, lat0 = math.radians(lon0_deg), math.radians(lat0_deg)
    xv, yv, tri = triangular_lattice(width_km * 1e3, height_km * 1e3,
                                     edge_km * 1e3)
    rng = np.random.default_rng(seed) if shuffle else None
    uv = _plane_to_unit(xv, yv, lon0, lat0)
    vlon, vlat = _unit_to_lonlat(uv)
    # spherical circumcentre of each triangle: normal of the plane through
    # its three corners, oriented outwards
    a, b, c = uv[tri[:, 0]], uv[tri[:, 1]], uv[tri[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n *= np.sign(np.sum(n * a, axis=1))[:, None]
    clon, clat = _unit_to_lonlat(n)
    cov = _cells_of_vertex(tri, xv.size, rng)
    cx = (xv[tri[:, 0]] + xv[tri[:, 1]] + xv[tri[:, 2]]) / 3.0
    cy = (yv[tri[:, 0]] + yv[tri[:, 1]] + yv[tri[:, 2]]) / 3.0
    meta = dict(width_km=width_km, height_km=height_km, edge_km=edge_km,
                lon0_deg=lon0_deg, lat0_deg=lat0_deg)
    return Domain(clon=clon, clat=clat, hsurf=np.zeros_like(clon), vlon=vlon,
                  vlat=vlat, cells_of_vertex=cov, meta=meta, cx=cx, cy=cy)


# ---------------------------------------------------------------------------
# Terrain
# ---------------------------------------------------------------------------

def gaussian_random_field(nx: int, ny: int, spacing: float, beta: float,
                          std: float, rng) -> np.ndarray:
    """Spectral synthesis of a zero-mean field with power ~ k^-beta."""
    noise = rng.standard_normal((ny, nx))
    f = np.fft.rfft2(noise)
    ky = np.fft.fftfreq(ny, d=spacing)[:, None]
    kx = np.fft.rfftfreq(nx, d=spacing)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.zeros_like(k)
    amp[k > 0] = k[k > 0] ** (-beta / 2.0)
    field_ = np.fft.irfft2(f * amp, s=(ny, nx))
    field_ -= field_.mean()
    field_ *= std / field_.std()
    return field_


def _bilinear(grid: np.ndarray, x0: float, y0: float, spacing: float, x, y):
    ny, nx = grid.shape
    fx = np.clip((x - x0) / spacing, 0.0, nx - 1.000001)
    fy = np.clip((y - y0) / spacing, 0.0, ny - 1.000001)
    i = np.floor(fx).astype(np.int64)
    j = np.floor(fy).astype(np.int64)
    tx, ty = fx - i, fy - j
    return ((1 - tx) * (1 - ty) * grid[j, i] + tx * (1 - ty) * grid[j, i + 1]
            + (1 - tx) * ty * grid[j + 1, i] + tx * ty * grid[j + 1, i + 1])


def _field_on_box(width_m, height_m, spacing, beta, std, rng):
    nx = int(math.ceil(width_m / spacing)) + 8
    ny = int(math.ceil(height_m / spacing)) + 8
    g = gaussian_random_field(nx, ny, spacing, beta, std, rng)
    x0 = -0.5 * (nx - 1) * spacing
    y0 = -0.5 * (ny - 1) * spacing
    return g, x0, y0


def terrain_grf(dom: Domain, beta: float, std: float, seed: int,
                clip_zero: bool = True) -> np.ndarray:
    """Seeded Gaussian random field sampled at the cell centres."""
    rng = np.random.default_rng(seed)
    w = dom.meta["width_km"] * 1e3
    h = dom.meta["height_km"] * 1e3
    sp = dom.meta["edge_km"] * 1e3 / 2.0
    g, x0, y0 = _field_on_box(w, h, sp, beta, std, rng)
    z = _bilinear(g, x0, y0, sp, dom.cx, dom.cy)
    if clip_zero:
        z = np.maximum(z, 0.0)
    return z


def add_ridges(dom: Domain, z: np.ndarray, n: int, height: float,
               width_m: float, seed: int) -> np.ndarray:
    """Add n straight Gaussian ridges (FWHM = width_m) with seeded position
    and orientation."""
    rng = np.random.default_rng(seed)
    w = dom.meta["width_km"] * 1e3
    h = dom.meta["height_km"] * 1e3
    sig = width_m / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    out = z.copy()
    for _ in range(n):
        px = rng.uniform(-0.35, 0.35) * w
        py = rng.uniform(-0.35, 0.35) * h
        ang = rng.uniform(0.0, math.pi)
        d = (dom.cx - px) * math.sin(ang) - (dom.cy - py) * math.cos(ang)
        out += height * np.exp(-0.5 * (d / sig) ** 2)
    return out


def add_cliffs(dom: Domain, z: np.ndarray, n: int, step: float,
               seed: int) -> np.ndarray:
    """Add n straight cliff lines, each raising one side by ``step`` metres."""
    rng = np.random.default_rng(seed)
    w = dom.meta["width_km"] * 1e3
    h = dom.meta["height_km"] * 1e3
    out = z.copy()
    for _ in range(n):
        px = rng.uniform(-0.4, 0.4) * w
        py = rng.uniform(-0.4, 0.4) * h
        ang = rng.uniform(0.0, 2.0 * math.pi)
        d = (dom.cx - px) * math.sin(ang) - (dom.cy - py) * math.cos(ang)
        out += step * (d > 0.0)
    return out


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

def config1(seed: int = 101, size_km: float = 100.0,
            edge_km: float = 2.0) -> Domain:
    """Plain-method CPU case: 100 x 100 km at 2 km edge (~5.8k cells), field
    k^-3, std 500 m, clipped at 0."""
    dom = lattice_domain(size_km, size_km, edge_km, seed=seed)
    dom.hsurf = terrain_grf(dom, beta=3.0, std=500.0, seed=seed)
    dom.meta.update(name="c1", seed=seed)
    return dom


def config2(seed: int = 202, width_km: float = 600.0, height_km: float = 400.0,
            edge_km: float = 1.0) -> Domain:
    """Plain GPU case: 600 x 400 km at 1 km edge (~554k cells), alpine-like
    field (k^-3, std 1200 m) plus 3 Gaussian ridges of 2500 m and 20 km
    width, rescaled to a maximum of 4500 m."""
    dom = lattice_domain(width_km, height_km, edge_km, seed=seed)
    z = terrain_grf(dom, beta=3.0, std=1200.0, seed=seed, clip_zero=False)
    z = add_ridges(dom, z, 3, 2500.0, 20e3, seed=seed + 1)
    z = z - np.percentile(z, 15.0)
    z = np.maximum(z, 0.0)
    z *= 4500.0 / z.max()
    dom.hsurf = z
    dom.meta.update(name="c2", seed=seed)
    return dom


def tilted_plane(size_km: float, edge_km: float, slope_east: float,
                 slope_north: float, seed: int = 7) -> Domain:
    """Grid whose mesh points lie exactly (to FP64) on the tilted plane
    z = slope_east * x + slope_north * y of the tangent frame at the box
    centre.  hsurf is solved per point so that the ECEF point (R + h) u lies
    on that plane; the horizon then has a closed form (see the tests)."""
    dom = lattice_domain(size_km, size_km, edge_km, seed=seed)
    lon0 = math.radians(dom.meta["lon0_deg"])
    lat0 = math.radians(dom.meta["lat0_deg"])
    up, east, north = _tangent_basis(lon0, lat0)
    u = np.stack([np.cos(dom.clat) * np.cos(dom.clon),
                  np.cos(dom.clat) * np.sin(dom.clon), np.sin(dom.clat)], 1)
    # point p = (R + h) u ; plane through R*up with normal m
    m = up - slope_east * east - slope_north * north
    # (p - R up) . m = 0  ->  (R + h) (u . m) = R (up . m)
    rh = RAD_EARTH * np.dot(up, m) / (u @ m)
    dom.hsurf = rh - RAD_EARTH
    dom.meta.update(name="tilted_plane", slope_east=slope_east,
                    slope_north=slope_north, plane_normal=m.tolist())
    return dom


def _subset_lattice(xv, yv, tri, keep):
    """Triangles tri[keep] with their vertices renumbered."""
    t = tri[keep]
    used = np.unique(t)
    remap = np.full(xv.size, -1, np.int64)
    remap[used] = np.arange(used.size)
    return xv[used], yv[used], remap[t]


def config3(seed: int = 303, core_km: float = 60.0, core_edge_km: float = 0.1,
            ring_km: float = 40.0, ring_edge_km: float = 0.5) -> Domain:
    """Fine-mesh stress: a core_km box at core_edge_km (~831k cells at the
    defaults) inside an outer ring_km-wide ring at ring_edge_km (~148k cells)
    that supplies the distant horizon.  Core and ring share no vertices: the
    seam vertices have fewer than six cells (-2 fill), as where two ICON
    grids meet.  Field: k^-2.5, std 900 m, plus 300 m cliff steps, shifted
    to a minimum of 0."""
    lon0, lat0 = math.radians(8.5), math.radians(46.5)
    half = core_km * 500.0
    xc, yc, tc = triangular_lattice(core_km * 1e3, core_km * 1e3, core_edge_km * 1e3)
    outer = core_km + 2 * ring_km
    xr, yr, tr = triangular_lattice(outer * 1e3, outer * 1e3, ring_edge_km * 1e3)
    cxr = (xr[tr[:, 0]] + xr[tr[:, 1]] + xr[tr[:, 2]]) / 3.0
    cyr = (yr[tr[:, 0]] + yr[tr[:, 1]] + yr[tr[:, 2]]) / 3.0
    gap = ring_edge_km * 1e3
    keep = (np.abs(cxr) > half + gap) | (np.abs(cyr) > half + gap)
    xr, yr, tr = _subset_lattice(xr, yr, tr, keep)
    xv = np.concatenate([xc, xr])
    yv = np.concatenate([yc, yr])
    tri = np.concatenate([tc, tr + xc.size])
    rng = np.random.default_rng(seed)
    uv = _plane_to_unit(xv, yv, lon0, lat0)
    vlon, vlat = _unit_to_lonlat(uv)
    a, b, c = uv[tri[:, 0]], uv[tri[:, 1]], uv[tri[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n *= np.sign(np.sum(n * a, axis=1))[:, None]
    clon, clat = _unit_to_lonlat(n)
    cov = _cells_of_vertex(tri, xv.size, rng)
    cx = (xv[tri[:, 0]] + xv[tri[:, 1]] + xv[tri[:, 2]]) / 3.0
    cy = (yv[tri[:, 0]] + yv[tri[:, 1]] + yv[tri[:, 2]]) / 3.0
    meta = dict(width_km=outer, height_km=outer, edge_km=ring_edge_km, core_km=core_km,
                core_edge_km=core_edge_km, lon0_deg=8.5, lat0_deg=46.5, name="c3", seed=seed,
                ncore=int(tc.shape[0]))
    dom = Domain(clon=clon, clat=clat, hsurf=np.zeros_like(clon), vlon=vlon, vlat=vlat,
                 cells_of_vertex=cov, meta=meta, cx=cx, cy=cy)
    frng = np.random.default_rng(seed + 1)
    sp = core_edge_km * 1e3 * 2.0
    g, gx0, gy0 = _field_on_box(outer * 1e3, outer * 1e3, sp, 2.5, 900.0, frng)
    z = _bilinear(g, gx0, gy0, sp, cx, cy)
    z = add_cliffs(dom, z, 6, 300.0, seed + 2)
    dom.hsurf = z - z.min()
    return dom


def config4(n_domains: int = 64, seed: int = 404, cells_min: float = 5e3,
            cells_max: float = 4e5):
    """Batched uneven domains: cell counts log-uniform in [cells_min,
    cells_max], edge lengths in {2, 1, 0.5} km, nhori in {24, 48}, each
    domain from its own seed with the config-1 field recipe.  Returns a list
    of (Domain, nhori)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_domains):
        n = float(np.exp(rng.uniform(np.log(cells_min), np.log(cells_max))))
        edge = float(rng.choice([2.0, 1.0, 0.5]))
        nhori = int(rng.choice([24, 48]))
        side = math.sqrt(n * math.sqrt(3.0) / 4.0 * edge * edge)  # km, square box
        s = seed * 1000 + k
        dom = lattice_domain(side, side, edge, lon0_deg=8.5 + 0.01 * k, lat0_deg=46.5, seed=s)
        dom.hsurf = terrain_grf(dom, beta=3.0, std=500.0, seed=s)
        dom.meta.update(name=f"c4_{k}", seed=s, nhori=nhori)
        out.append((dom, nhori))
    return out


CONFIGS = {"c1": config1, "c2": config2}
