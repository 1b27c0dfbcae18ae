"""Seeded synthetic workloads in the fm3dt_parallel file contract.

Everything the solver reads is produced here from a seed: the B-spline
vertex model (``.vtx``), teleseismic plane-wave boundary times
(``itimes.out`` and, for the side-face mode, the four side faces), local
point sources, and surface receivers.  Geometry conventions follow the
reference reader (fm3dt_parallel.f90:6103-6126):

* heights are km above the datum (positive up); radial index 1 is the top
  and grows downward, ``h = rgor - (ir-1)*dnr``;
* colatitude ``theta = (90 - lat)*pi/180`` grows southward with the x index;
* longitude ``phi = lon*pi/180`` grows eastward with the z index;
* ``pi`` is the reference's default-real literal (fm3dt_parallel.f90:61),
  i.e. float32(3.1415926535898) widened to double.

Teleseismic boundary times use a 1-D velocity model and a constant ray
parameter P (s/rad) from a virtual epicentre: ``T = T0 + P*(Delta - D0) +
tau(r)`` with ``tau(r) = int_{r_bot}^{r} sqrt(1/v^2 - P^2/r^2) dr``.  Because
``Delta`` is a great-circle distance, this is an exact solution of the
spherical eikonal equation in the 1-D model, which makes it an analytic
pin for the solver as well as an input generator.
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
from typing import List, Optional

import numpy as np

PI_REF = float(np.float32(3.1415926535898))  # fm3dt_parallel.f90:61 (default-real literal)
EARTH_KM = 6371.0


@dataclasses.dataclass
class GridDesc:
    """Vertex grid + propagation grid, exactly the six header lines of a .vtx file."""
    nvr: int
    nvx: int
    nvz: int
    gor: float
    goxd: float
    gozd: float
    dvr: float
    dvxd: float
    dvzd: float
    nnr: int
    nnx: int
    nnz: int
    rgor: float
    rgoxd: float
    rgozd: float
    dnr: float
    dnxd: float
    dnzd: float
    earth: float = EARTH_KM

    # derived quantities in the reference's own arithmetic (fm3dt_parallel.f90:6110-6117)
    @property
    def dnx(self) -> float:
        return self.dnxd * PI_REF / 180.0

    @property
    def dnz(self) -> float:
        return self.dnzd * PI_REF / 180.0

    @property
    def rgox(self) -> float:
        return (90.0 - self.rgoxd) * PI_REF / 180.0

    @property
    def rgoz(self) -> float:
        return self.rgozd * PI_REF / 180.0

    @property
    def dvx(self) -> float:
        return self.dvxd * PI_REF / 180.0

    @property
    def dvz(self) -> float:
        return self.dvzd * PI_REF / 180.0

    @property
    def gox(self) -> float:
        return (90.0 - self.goxd) * PI_REF / 180.0

    @property
    def goz(self) -> float:
        return self.gozd * PI_REF / 180.0

    @property
    def nnodes(self) -> int:
        return self.nnr * self.nnx * self.nnz

    @property
    def nvert(self) -> int:
        return (self.nvr + 2) * (self.nvx + 2) * (self.nvz + 2)


def v1d(depth_km):
    """1-D P model: 5.8 km/s at the datum, +3.1 km/s per 400 km (linear)."""
    return 5.8 + 3.1 * np.asarray(depth_km, dtype=np.float64) / 400.0


def make_grid(nn=(61, 61, 61), depth_km=400.0, span_deg=(10.0, 10.0),
              centre=(-35.0, 145.0), nv=(12, 12, 12), margin_cells=3.0,
              earth=EARTH_KM) -> GridDesc:
    """Propagation grid of ``nn`` nodes over [0, depth_km] x span, and a vertex
    grid of ``nv`` interior vertices that covers it with ``margin_cells``
    propagation cells to spare on every side (so every node, halo included,
    has a full 4-tap B-spline support)."""
    nnr, nnx, nnz = nn
    dnr = depth_km / (nnr - 1)
    dnxd = span_deg[0] / (nnx - 1)
    dnzd = span_deg[1] / (nnz - 1)
    rgor = 0.0
    rgoxd = centre[0] + span_deg[0] / 2.0
    rgozd = centre[1] - span_deg[1] / 2.0
    nvr, nvx, nvz = nv
    gor = rgor + margin_cells * dnr
    vbot = rgor - depth_km - margin_cells * dnr
    dvr = (gor - vbot) / (nvr - 1)
    goxd = rgoxd + margin_cells * dnxd
    vsouth = rgoxd - span_deg[0] - margin_cells * dnxd
    dvxd = (goxd - vsouth) / (nvx - 1)
    gozd = rgozd - margin_cells * dnzd
    veast = rgozd + span_deg[1] + margin_cells * dnzd
    dvzd = (veast - gozd) / (nvz - 1)
    return GridDesc(nvr, nvx, nvz, gor, goxd, gozd, dvr, dvxd, dvzd,
                    nnr, nnx, nnz, rgor, rgoxd, rgozd, dnr, dnxd, dnzd, earth)


def _smooth121(a: np.ndarray) -> np.ndarray:
    for ax in range(a.ndim):
        p = np.concatenate([np.take(a, [0], axis=ax), a, np.take(a, [-1], axis=ax)], axis=ax)
        n = a.shape[ax]
        a = 0.25 * np.take(p, range(0, n), axis=ax) + 0.5 * np.take(p, range(1, n + 1), axis=ax) \
            + 0.25 * np.take(p, range(2, n + 2), axis=ax)
    return a


def make_velv(g: GridDesc, amp: float = 0.05, seed: int = 1) -> np.ndarray:
    """Vertex velocities, shape (nvz+2, nvx+2, nvr+2) C-order == r fastest,
    the .vtx record order.  1-D gradient times (1 + delta), delta a smoothed
    Gaussian field scaled to max|delta| = amp."""
    rng = np.random.default_rng(seed)
    k = np.arange(g.nvr + 2)
    hv = g.gor - (k - 1) * g.dvr
    base = v1d(-hv)[None, None, :]
    d = rng.standard_normal((g.nvz + 2, g.nvx + 2, g.nvr + 2))
    d = _smooth121(d)
    if amp > 0:
        d = d * (amp / np.max(np.abs(d)))
    else:
        d = d * 0.0
    return np.ascontiguousarray(base * (1.0 + d))


def make_velv_slab(g: GridDesc, seed: int = 4, slab_amp: float = 0.10, slab_km: float = 60.0,
                   dip_deg: float = 45.0, nblobs: int = 10, blob_amp: float = 0.05,
                   blob_km=(30.0, 60.0)) -> np.ndarray:
    """Sharp-contrast vertex model (same layout as make_velv): the 1-D
    gradient times (1 + slab + blobs).  The slab is a layer slab_km thick
    through the grid centre at mid-depth, striking along phi and dipping
    dip_deg along theta, +slab_amp inside with a 10 km cosine taper; the
    blobs are Gaussians of radius U(blob_km) at uniform positions, each
    +-blob_amp (random sign)."""
    rng = np.random.default_rng(seed)
    k = np.arange(g.nvr + 2)
    depth = -(g.gor - (k - 1) * g.dvr)
    lat = g.goxd - (np.arange(g.nvx + 2) - 1) * g.dvxd
    lon = g.gozd + (np.arange(g.nvz + 2) - 1) * g.dvzd
    km = 111.19
    lat0, lon0 = float(np.mean(lat)), float(np.mean(lon))
    X = ((lat - lat0) * km)[None, :, None]
    Y = ((lon - lon0) * km * np.cos(np.deg2rad(lat0)))[:, None, None]
    D = depth[None, None, :]
    dmid = 0.5 * (depth[1] + depth[-2])
    dip = np.deg2rad(dip_deg)
    dist = np.abs(np.sin(dip) * X + np.cos(dip) * (D - dmid))
    half, taper = 0.5 * slab_km, 10.0
    slab = np.where(dist <= half - taper, 1.0,
                    np.where(dist >= half, 0.0, 0.5 * (1.0 + np.cos(np.pi * (dist - half + taper) / taper))))
    d = slab_amp * np.broadcast_to(slab, (Y.shape[0], X.shape[1], D.shape[2]))
    span = [(X.min(), X.max()), (Y.min(), Y.max()), (depth[1], depth[-2])]
    for _ in range(nblobs):
        c = [rng.uniform(lo, hi) for lo, hi in span]
        r = rng.uniform(*blob_km)
        a = blob_amp * rng.choice([-1.0, 1.0])
        d = d + a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (D - c[2]) ** 2) / r ** 2)
    base = v1d(depth)[None, None, :]
    return np.ascontiguousarray(base * (1.0 + d))


def node_coords(g: GridDesc):
    """(h[ir], theta[ix], phi[iz]) of propagation nodes, reference arithmetic."""
    h = np.array([g.rgor - (i - 1) * g.dnr for i in range(1, g.nnr + 1)])
    th = np.array([g.rgox + (i - 1) * g.dnx for i in range(1, g.nnx + 1)])
    ph = np.array([g.rgoz + (i - 1) * g.dnz for i in range(1, g.nnz + 1)])
    return h, th, ph


@dataclasses.dataclass
class TeleSource:
    baz_deg: float
    inc_deg: float
    t0: float            # boundary time at distance D0 (s)
    d0: float            # epicentral distance of the grid centre (rad)
    p: float = 0.0       # ray parameter (s/rad), filled by tele_sources
    th_e: float = 0.0    # virtual epicentre colatitude (rad)
    ph_e: float = 0.0    # virtual epicentre longitude (rad)


def _destination(th0, ph0, az, dist):
    """Point at angular distance ``dist`` along azimuth ``az`` (from north,
    clockwise) from (colat th0, lon ph0); all radians."""
    lat0 = np.pi / 2 - th0
    lat1 = np.arcsin(np.sin(lat0) * np.cos(dist) + np.cos(lat0) * np.sin(dist) * np.cos(az))
    dlon = np.arctan2(np.sin(az) * np.sin(dist) * np.cos(lat0),
                      np.cos(dist) - np.sin(lat0) * np.sin(lat1))
    return np.pi / 2 - lat1, ph0 + dlon


def great_circle(th1, ph1, th2, ph2):
    """Angular distance (rad), haversine form (well conditioned at small range)."""
    lat1 = np.pi / 2 - th1
    lat2 = np.pi / 2 - th2
    s = np.sin((lat2 - lat1) / 2) ** 2 + np.cos(lat1) * np.cos(lat2) * np.sin((ph2 - ph1) / 2) ** 2
    return 2.0 * np.arcsin(np.sqrt(s))


def tele_sources(g: GridDesc, n: int, seed: int, inc_range=(15.0, 35.0),
                 d0_deg: float = 60.0) -> List[TeleSource]:
    rng = np.random.default_rng(seed)
    h, th, ph = node_coords(g)
    thc = 0.5 * (th[0] + th[-1])
    phc = 0.5 * (ph[0] + ph[-1])
    r_bot = g.earth + h[-1]
    v_bot = float(v1d(-h[-1]))
    out = []
    for _ in range(n):
        baz = float(rng.uniform(0.0, 360.0))
        inc = float(rng.uniform(*inc_range))
        t0 = float(rng.uniform(420.0, 600.0))
        s = TeleSource(baz, inc, t0, np.deg2rad(d0_deg))
        s.p = r_bot * np.sin(np.deg2rad(inc)) / v_bot
        s.th_e, s.ph_e = _destination(thc, phc, np.deg2rad(baz), s.d0)
        out.append(s)
    return out


def _tau(g: GridDesc, p: float, h_nodes: np.ndarray) -> np.ndarray:
    """tau(h) = int_{h_bot}^{h} sqrt(1/v^2 - p^2/r^2) dh' with 8-point
    Gauss-Legendre per propagation cell (the integrand is smooth)."""
    xg, wg = np.polynomial.legendre.leggauss(8)
    h_bot = h_nodes[-1]
    out = np.zeros_like(h_nodes)
    # integrate from the bottom upward, cell by cell
    hs = h_nodes[::-1]
    acc = 0.0
    res = [0.0]
    for a, b in zip(hs[:-1], hs[1:]):
        hm = 0.5 * (a + b)
        hr = 0.5 * (b - a)
        hh = hm + hr * xg
        r = g.earth + hh
        q = np.sqrt(np.maximum(1.0 / v1d(-hh) ** 2 - (p / r) ** 2, 0.0))
        acc += hr * np.sum(wg * q)
        res.append(acc)
    out[:] = np.array(res)[::-1]
    assert abs(h_nodes[-1] - h_bot) == 0.0
    return out


def tele_field_1d(g: GridDesc, s: TeleSource) -> np.ndarray:
    """Exact 1-D-model plane-wave field on every node, shape (nnr, nnx, nnz)."""
    h, th, ph = node_coords(g)
    tau = _tau(g, s.p, h)
    T, P = np.meshgrid(th, ph, indexing="ij")
    dist = great_circle(T, P, s.th_e, s.ph_e)
    horiz = s.t0 + s.p * (dist - s.d0)
    return tau[:, None, None] + horiz[None, :, :]


def tele_bottom_face(g: GridDesc, s: TeleSource) -> np.ndarray:
    """Bottom-face boundary times, shape (nnx, nnz) == itimes.out record
    (phi fastest; fm3dt_parallel.f90:6208-6219, 6908-6913)."""
    h, th, ph = node_coords(g)
    T, P = np.meshgrid(th, ph, indexing="ij")
    dist = great_circle(T, P, s.th_e, s.ph_e)
    return np.ascontiguousarray(s.t0 + s.p * (dist - s.d0))


def tele_boundary_nodes(g: GridDesc, s: TeleSource, sides: bool = True):
    """Fixed-node set for a plane wave injected on the bottom face and, with
    `sides`, on the side faces the wave enters through (in theta and in phi,
    the face where the 1-D times are earlier): (node index, time) with node
    index (ir*nnx + ix)*nnz + iz (0-based, ir = 0 at the top) and 1-D-model
    (tau-p) times.  The faces it leaves through stay free: fixing 1-D times
    there would contradict the 3-D field that reaches them."""
    T = tele_field_1d(g, s)
    m = np.zeros(T.shape, bool)
    m[-1] = True
    if sides:
        bot = T[-1]
        m[:, 0 if bot[0, :].mean() <= bot[-1, :].mean() else -1, :] = True
        m[:, :, 0 if bot[:, 0].mean() <= bot[:, -1].mean() else -1] = True
    idx = np.flatnonzero(m.ravel())
    return idx.astype(np.int64), T.ravel()[idx].astype(np.float64)


def tele_faces(g: GridDesc, s: TeleSource):
    """The fixed set of tele_boundary_nodes as face arrays for
    fmm3d_solve_tele_sides: (bottom (nnx, nnz), theta face (nnr, nnz), phi
    face (nnr, nnx), side (theta, phi) with 0 the low-index face and 1 the
    high-index one), 1-D-model times."""
    T = tele_field_1d(g, s)
    bot = T[-1]
    xs = 0 if bot[0, :].mean() <= bot[-1, :].mean() else 1
    zs = 0 if bot[:, 0].mean() <= bot[:, -1].mean() else 1
    return (bot.copy(), T[:, -1 if xs else 0, :].copy(), T[:, :, -1 if zs else 0].copy(),
            np.array([xs, zs], np.int32))


def point_source_rad(g: GridDesc, src: "LocalSource"):
    """(height km, colatitude rad, longitude rad) of a local source."""
    return np.array([-src.depth_km, (90.0 - src.lat) * PI_REF / 180.0, src.lon * PI_REF / 180.0])


@dataclasses.dataclass
class LocalSource:
    depth_km: float
    lat: float
    lon: float
    radius_nodes: int = 10
    refine: int = 5


def local_sources(g: GridDesc, n: int, seed: int, depth_range=(5.0, 80.0),
                  radius_range=(5, 20)) -> List[LocalSource]:
    rng = np.random.default_rng(seed)
    south = g.rgoxd - (g.nnx - 1) * g.dnxd
    east = g.rgozd + (g.nnz - 1) * g.dnzd
    out = []
    for _ in range(n):
        d = float(rng.uniform(*depth_range))
        lat = float(rng.uniform(south + 0.2 * (g.rgoxd - south), g.rgoxd - 0.2 * (g.rgoxd - south)))
        lon = float(rng.uniform(g.rgozd + 0.2 * (east - g.rgozd), east - 0.2 * (east - g.rgozd)))
        r = int(rng.integers(radius_range[0], radius_range[1] + 1))
        out.append(LocalSource(d, lat, lon, r, 5))
    return out


def receivers(g: GridDesc, n: int, seed: int, max_per_quadrant: Optional[int] = None) -> np.ndarray:
    """(n, 3) array of (height km, lat deg, lon deg) at the datum, uniform
    over the grid interior (one cell kept clear of every edge).  With
    ``max_per_quadrant`` the draw is repeated until no 2x2 surface quadrant
    holds more than that many receivers (the reference sizes its per-rank
    receiver table nrc/2, fm3dt_parallel.f90:6335)."""
    rng = np.random.default_rng(seed)
    south = g.rgoxd - (g.nnx - 1) * g.dnxd
    east = g.rgozd + (g.nnz - 1) * g.dnzd
    for _ in range(1000):
        lat = rng.uniform(south + g.dnxd, g.rgoxd - g.dnxd, n)
        lon = rng.uniform(g.rgozd + g.dnzd, east - g.dnzd, n)
        if max_per_quadrant is None:
            break
        # reference split: low ranks take the remainder (fm3dt_parallel.f90:6127-6150)
        xs = (g.nnx + 1) // 2
        zs = (g.nnz + 1) // 2
        ix = np.floor(((90.0 - lat) * PI_REF / 180.0 - g.rgox) / g.dnx).astype(int) + 1
        iz = np.floor((lon * PI_REF / 180.0 - g.rgoz) / g.dnz).astype(int) + 1
        q = (ix > xs).astype(int) * 2 + (iz > zs).astype(int)
        if np.bincount(q, minlength=4).max() <= max_per_quadrant:
            break
    return np.stack([np.zeros(n), lat, lon], axis=1)


def receivers_rad(rcv_deg: np.ndarray) -> np.ndarray:
    """(height km, lat deg, lon deg) -> (height km, colat rad, lon rad), the
    reference's conversion (fm3dt_parallel.f90:6310-6311)."""
    r = np.asarray(rcv_deg, dtype=np.float64)
    return np.stack([r[:, 0], (90.0 - r[:, 1]) * PI_REF / 180.0, r[:, 2] * PI_REF / 180.0], axis=1)


# ---------------------------------------------------------------- file contract

def write_vtx(path: str, g: GridDesc, velv: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write(f"{g.nvr} {g.nvx} {g.nvz}\n")
        f.write(f"{g.gor!r} {g.goxd!r} {g.gozd!r}\n")
        f.write(f"{g.dvr!r} {g.dvxd!r} {g.dvzd!r}\n")
        f.write(f"{g.nnr} {g.nnx} {g.nnz}\n")
        f.write(f"{g.rgor!r} {g.rgoxd!r} {g.rgozd!r}\n")
        f.write(f"{g.dnr!r} {g.dnxd!r} {g.dnzd!r}\n")
        flat = np.asarray(velv, dtype=np.float64).reshape(-1)
        f.write("\n".join(repr(float(v)) for v in flat))
        f.write("\n")


IN_TEMPLATE = """ccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
c INPUT PARAMETERS
cccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
sources.dat            c: File containing source information
itimes.out             c: File containing initial traveltimes
receivers.dat          c: File containing receiver positions
gridc.vtx              c: File containing velocity grid information
{earth:.1f}                 c: Earth radius in km
{fom}                      c: Use first-order(0) or mixed-order(1) scheme
1                      c: Limit traveltime field to receivers (0=no,1=yes)
0.4     0.4            c: (lat,lon) cushion around receiver array (degrees)
ccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
c OUTPUT FILES
ccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccccc
1                      c: find source-receiver traveltimes (0=no,1=yes)
rtravel.out            c: Name of file containing source-receiver traveltimes
{cfd}                      c: Calculate partial (Frechet) derivatives (0=no,1=yes)
frechet.out            c: Name of file containing Frechet derivatives
"""


def write_case(d: str, g: GridDesc, velv: np.ndarray, srcs: List[TeleSource],
               rcv: np.ndarray, fom: int = 1, cfd: int = 1) -> dict:
    """Write the five input files of one forward run into directory ``d``;
    returns the bottom faces array (nsrc, nnx, nnz) and file hashes."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "fm3dt_parallel.in"), "w") as f:
        f.write(IN_TEMPLATE.format(earth=g.earth, fom=fom, cfd=cfd))
    write_vtx(os.path.join(d, "gridc.vtx"), g, velv)
    with open(os.path.join(d, "sources.dat"), "w") as f:
        f.write("1\n")
        f.write(f"{len(srcs)}\n")
        for s in srcs:
            lat = 90.0 - np.rad2deg(s.th_e)
            lon = np.rad2deg(s.ph_e)
            f.write(f"  {lat:.6f}  {lon:.6f}  0.0\n  P\n")
    with open(os.path.join(d, "receivers.dat"), "w") as f:
        f.write("1\n")
        f.write(f"{len(rcv)}\n")
        for h, la, lo in rcv:
            f.write(f"{float(h)!r} {float(la)!r} {float(lo)!r}\n")
    faces = np.stack([tele_bottom_face(g, s) for s in srcs]) if srcs else np.zeros((0, g.nnx, g.nnz))
    faces.astype("<f8").tofile(os.path.join(d, "itimes.out"))
    hashes = {}
    for fn in ("gridc.vtx", "itimes.out", "receivers.dat", "sources.dat"):
        with open(os.path.join(d, fn), "rb") as f:
            hashes[fn] = hashlib.sha256(f.read()).hexdigest()
    return {"faces": faces, "sha256": hashes}


# ----------------------------------------------------------- named workloads

def tele_case(nn=(61, 61, 61), depth_km=400.0, span_deg=(10.0, 10.0), nv=(12, 12, 12),
              amp=0.05, nsrc=6, nrc=20, seed=1):
    """One teleseismic workload: grid, vertex model, sources, receivers."""
    g = make_grid(nn, depth_km, span_deg, nv=nv)
    velv = make_velv(g, amp, seed)
    srcs = tele_sources(g, nsrc, seed)
    rcv = receivers(g, nrc, seed, max_per_quadrant=nrc // 2)
    return g, velv, srcs, rcv


def config_321(nsrc=64, nrc=500, seed=4):
    """321^3 nodes, 0-600 km, 15x15 deg, 33^3 vertices, a +10 % dipping slab
    (60 km thick, 45 deg) and ten +-5 % Gaussian blobs (radius 30-60 km)."""
    g = make_grid((321, 321, 321), 600.0, (15.0, 15.0), nv=(33, 33, 33))
    velv = make_velv_slab(g, seed)
    srcs = tele_sources(g, nsrc, seed)
    rcv = receivers(g, nrc, seed)
    return g, velv, srcs, rcv


def config_201(nsrc=50, nrc=300, seed=2):
    """201^3 nodes, 0-600 km, 15x15 deg, 25^3 vertices, +-3 %."""
    return tele_case((201, 201, 201), 600.0, (15.0, 15.0), (25, 25, 25), 0.03, nsrc, nrc, seed)
