"""Builders of the five workloads the path is measured and checked on.

  cantilever()      steel bar 1.0 x 0.1 x 0.1 m, 40x4x4 Kuhn cells (3840 tets, 1025 nodes),
                    x=0 clamped, 1 kN tip load on x=1 ramped over 1 ms, elastic, 5000 steps
  plate()           flat plate 10 x 4 x 0.02 m, 350x144x4 cells (1,209,600 tets), nodes
                    jittered +-5% of the spacing, edges clamped, 200 kPa ramped over 5 ms
  stiffened_panel() 6 x 3 m, 12 mm plate + 4 flat bars 200 x 12 mm (~812k tets), J2 with
                    sigma_y 355 MPa and H = 1 GPa, r = 0.5 m sphere at 5 m/s into mid-panel
  batch_cases()     64 plates with tet counts log-uniform in [1e4, 3e5], pressure 50-400 kPa,
                    yield 235-460 MPa
  served_case()     a stiffened panel scaled to a target tet count with indenter speed 2-10 m/s
  hex_plate()       solid-shell plating 30 x 25 m x 12 mm, 1200 x 1000 x 1 hexes
                    (1,200,000 elements), edges clamped, 50 kPa on the top face ramped over
                    5 ms, elastic
  hex_zone()        a zone patch of solid-shell plating, 1.2 x 0.6 m x 12 mm, edges clamped,
                    a rigid sphere punched into its middle (the elastic zone::Solver solve)

Every builder takes a `scale` (or explicit sizes) so the same geometry can be run at
a size the FP64 oracle finishes in seconds.
"""
import math

import numpy as np

from . import CONTACT, PIN_ALL, Case, HexCase, Mesh

STEEL = dict(E=210e9, nu=0.3, rho=7850.0)
TIP_NODE = (2 * 5 + 2) * 41 + 40  # node (i, j, k) = (40, 2, 2) of the 40x4x4 cantilever


def cantilever(nx=40, ny=4, nz=4, total_load=1000.0, ramp=1e-3, **kw):
    m = Mesh.box(1.0, 0.1, 0.1, nx, ny, nz)
    m.flag_plane(0, 0.0, PIN_ALL)
    m.traction(0, 1.0, (0.0, 0.0, -total_load / (0.1 * 0.1)))
    opts = dict(STEEL, load_ramp=ramp, dt_safety=0.9, sample_every=100)
    opts.update(kw)
    return Case.from_mesh(m, **opts)


def plate(lx=10.0, ly=4.0, lz=0.02, nx=350, ny=144, nz=4, pressure=200e3, ramp=5e-3,
          jitter=0.05, seed=20260101, plastic=False, yield_=355e6, **kw):
    m = Mesh.box(lx, ly, lz, nx, ny, nz)
    if jitter:
        m.jitter(jitter, seed)
    for axis, v in ((0, 0.0), (0, lx), (1, 0.0), (1, ly)):
        m.flag_plane(axis, v, PIN_ALL)
    m.traction(2, lz, (0.0, 0.0, -pressure))
    opts = dict(STEEL, load_ramp=ramp, dt_safety=0.9, sample_every=100, plastic=plastic,
                yield_=yield_, hardening=1e9)
    opts.update(kw)
    return Case.from_mesh(m, **opts)


def plate_scaled(scale):
    """The large plate's shape at 1/scale of its cells per in-plane axis (same physics)."""
    nx, ny = max(2, round(350 / scale)), max(2, round(144 / scale))
    return plate(nx=nx, ny=ny)


PANEL = dict(lx=6.0, ly=3.0, t_plate=0.012, n_stiff=4, web_h=0.2, web_t=0.012)


def stiffened_panel(nx=462, ny_bay=45, nz_web=16, speed=5.0, radius=0.5, yield_=355e6,
                    hardening=1e9, geometry=None, **kw):
    g = dict(PANEL)
    if geometry:
        g.update(geometry)
    m = Mesh.panel(g["lx"], g["ly"], g["t_plate"], g["n_stiff"], g["web_h"], g["web_t"], nx,
                   ny_bay, nz_web)
    for axis, v in ((0, 0.0), (0, g["lx"]), (1, 0.0), (1, g["ly"])):
        m.flag_plane(axis, v, PIN_ALL)
    cx, cy = 0.5 * g["lx"], 0.5 * g["ly"]
    m.flag_disc(2, 0.0, cx, cy, radius, CONTACT)
    indenter = dict(radius=radius, center=(cx, cy, -radius), velocity=(0.0, 0.0, speed),
                    kappa=0.1)
    opts = dict(STEEL, plastic=True, yield_=yield_, hardening=hardening, indenter=indenter,
                dt_safety=0.9, sample_every=100)
    opts.update(kw)
    return Case.from_mesh(m, **opts)


def panel_tets(nx, ny_bay, nz_web, n_stiff=4):
    ny = (n_stiff + 1) * ny_bay + n_stiff
    return 6 * (nx * ny + nx * n_stiff * nz_web)


def panel_cells_for(target_tets):
    """(nx, ny_bay, nz_web) of a 6 x 3 m stiffened panel with about target_tets tets."""
    best = None
    lo, hi = 0.005, 1.0
    for _ in range(60):
        h = math.sqrt(lo * hi)
        nx, ny_bay, nz = max(2, round(6.0 / h)), max(1, round(0.5904 / h)), max(1, round(0.2 / h))
        t = panel_tets(nx, ny_bay, nz)
        if best is None or abs(t - target_tets) < abs(best[3] - target_tets):
            best = (nx, ny_bay, nz, t)
        if t > target_tets:
            lo = h
        else:
            hi = h
    return best[:3]


def served_case(target_tets, speed, **kw):
    nx, ny_bay, nz = panel_cells_for(target_tets)
    return stiffened_panel(nx=nx, ny_bay=ny_bay, nz_web=nz, speed=speed, **kw)


def batch_cases(n=64, seed=7, lo=1e4, hi=3e5, **kw):
    rng = np.random.default_rng(seed)
    h = 10.0 / 350
    cases = []
    for _ in range(n):
        t = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        ny = max(2, round(math.sqrt(t / 24 / 2.5)))
        nx = max(2, round(2.5 * ny))
        p = rng.uniform(50e3, 400e3)
        sy = rng.uniform(235e6, 460e6)
        cases.append(plate(lx=nx * h, ly=ny * h, nx=nx, ny=ny, pressure=p, ramp=2.5e-4,
                           jitter=0.0, plastic=True, yield_=sy, **kw))
    return cases


def _clamp_edges(X, lx, ly, tol=1e-9):
    flags = np.zeros(len(X), np.uint32)
    edge = (X[:, 0] < tol) | (X[:, 0] > lx - tol) | (X[:, 1] < tol) | (X[:, 1] > ly - tol)
    flags[edge] = PIN_ALL
    return flags


def hex_plate(lx=30.0, ly=25.0, thickness=0.012, nx=1200, ny=1000, nz=1, pressure=50e3,
              ramp=5e-3, origin=(0.0, 0.0, 0.0), **kw):
    """Solid-shell plate (makePlateMesh) with clamped edges under a ramped pressure.
    `origin` shifts the whole plate (the rest coordinates the solve works from)."""
    opts = dict(STEEL, load_ramp=ramp, dt_safety=0.9, sample_every=100)
    opts.update(kw)
    c = HexCase.plate(lx, ly, thickness, nx, ny, nz, pressure=pressure, **opts)
    c.flags = _clamp_edges(c.X, lx, ly)
    c.X = np.ascontiguousarray(c.X + np.asarray(origin, np.float64))
    return c


def hex_plate_scaled(scale):
    """The hex plate's shape at 1/scale of its elements per in-plane axis."""
    return hex_plate(nx=max(2, round(1200 / scale)), ny=max(2, round(1000 / scale)))


def hex_zone(lx=1.2, ly=0.6, thickness=0.012, nx=48, ny=24, nz=1, radius=0.15, speed=2.0,
             origin=(0.0, 0.0, 0.0), **kw):
    """A zone patch: clamped edges, a sphere of `radius` moving at `speed` into the middle
    of the bottom face, contact on every free node (zone.cpp's indenter, elastic)."""
    c = HexCase.plate(lx, ly, thickness, nx, ny, nz)
    c.flags = _clamp_edges(c.X, lx, ly)
    c.flags[(c.flags & PIN_ALL) == 0] |= CONTACT
    o = np.asarray(origin, np.float64)
    c.X = np.ascontiguousarray(c.X + o)
    centre = o + np.array([0.5 * lx, 0.5 * ly, -0.5 * thickness - radius])
    c.indenter = dict(radius=radius, center=tuple(centre), velocity=(0.0, 0.0, speed), kappa=0.1)
    c.E, c.nu, c.rho = STEEL["E"], STEEL["nu"], STEEL["rho"]
    c.dt_safety, c.sample_every = 0.9, 100
    for k, v in kw.items():
        setattr(c, k, v)
    return c

