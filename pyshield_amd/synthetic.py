"""Seeded synthetic soundings for the microphysics workloads.

The generator draws a physically consistent state for a block of columns:

* interfaces ``p_{k+1/2} = ptop + (ps - ptop) (k/nz)^1.5`` (k = 0 is the model
  top), hydrostatic ``delp = p_{k+1/2} - p_{k-1/2}``, ``ps = 1000 hPa``;
* ``T(z) = max(300 - 6.5 K/km z, 200 K) + N(0, 1 K)`` (plus a per-block offset),
  with z from hypsometric integration upward, iterated twice;
* ``delz = -(Rd Tv / g) ln(p_{k+1/2}/p_{k-1/2}) < 0``;
* ``qv = RH qsat_w(T, rho)`` with ``RH ~ U(0.3, 1.05)`` per cell;
* 1-2 cloud layers per column with base ~ U(0, cloud_top - 1 km) and thickness
  ~ U(1, 5) km; inside them ql, qi, qr, qs, qg ~ LogNormal(ln median, sigma)
  independently, zero elsewhere;
* ``w ~ N(0, 0.5 m/s)``, ``ua, va ~ N(0, 10 m/s)``, ``qa = 0``,
  ``land ~ Bernoulli(0.3)``, ``area = 1e10 m^2``.

The seed of block ``tile`` of workload ``cfg`` is ``1000 cfg + tile``.
Everything is float64, shaped ``[ni, nj, nz]`` (2D: ``[ni, nj]``), k contiguous.
"""

from dataclasses import dataclass

import numpy as np

from . import constants as C


@dataclass
class Workload:
    name: str
    ni: int
    nj: int
    nz: int
    ntiles: int
    timestep: float
    ptop: float = 300.0
    tile_offsets: bool = False
    cloud_top: float = 12000.0
    median: float = 1.0e-4
    qg_median: float = 1.0e-4
    calls: int = 1

    @property
    def ncol(self):
        return self.ni * self.nj * self.ntiles


WORKLOADS = {
    # single 12x12 tile x 79 levels, dt 150 s (the reference's CPU-sized case)
    "tile12": Workload("tile12", 12, 12, 79, 1, 150.0),
    # cubed-sphere C96: 6 x 96 x 96 columns x 79 levels, dt 225 s, tile offsets
    "c96": Workload("c96", 96, 96, 79, 6, 225.0, tile_offsets=True),
    # production shape C384: 6 x 384 x 384 columns x 79 levels, dt 150 s
    "c384": Workload("c384", 384, 384, 79, 6, 150.0, calls=10),
    # C48: 6 x 48 x 48 x 79 (adjoint workload)
    "c48": Workload("c48", 48, 48, 79, 6, 150.0),
    # deep columns: one 768 x 768 tile x 127 levels, ptop 1 hPa, dt 90 s
    "deep768": Workload("deep768", 768, 768, 127, 1, 90.0, ptop=100.0,
                        cloud_top=12000.0, qg_median=5.0e-4),
}
WORKLOAD_IDS = {"tile12": 1, "c96": 2, "c384": 3, "c48": 4, "deep768": 5}


def qsat_water(t, den):
    """Saturation mixing ratio over water (closed form of
    microphysics_funcs.py:53-64), used only to build synthetic humidity."""
    return (C.E00 * np.exp((C.DC_VAP * np.log(t / C.TICE)
                            + C.LV0 * (t - C.TICE) / (t * C.TICE)) / C.RVGAS)
            ) / (C.RVGAS * t * den)


def make_state(ni, nj, nz, seed, *, ptop=300.0, t_offset=0.0, cloud_top=12000.0,
               median=1.0e-4, qg_median=None, ps=1.0e5):
    """One block of ni x nj columns.  Returns a dict of float64 arrays."""
    rng = np.random.Generator(np.random.PCG64(seed))
    qg_median = median if qg_median is None else qg_median
    shape = (ni, nj, nz)
    kk = np.arange(nz + 1, dtype=np.float64) / nz
    pint = ptop + (ps - ptop) * kk ** 1.5                      # [nz+1], top -> surface
    delp1 = np.diff(pint)                                        # [nz]
    logp = np.log(pint[1:] / pint[:-1])                          # ln(p_bot/p_top) > 0
    pmid = 0.5 * (pint[1:] + pint[:-1])

    noise = rng.standard_normal(shape)
    # hypsometric heights at layer mid-points, iterated twice
    t = np.full(shape, 288.0)
    for _ in range(2):
        dzl = C.RDGAS * t / C.GRAV * logp                        # layer thickness > 0
        zint_below = np.cumsum(dzl[..., ::-1], axis=-1)[..., ::-1] - dzl   # bottom of layer
        zmid = zint_below + 0.5 * dzl
        t = np.maximum(300.0 - 6.5e-3 * zmid, 200.0) + noise + t_offset
    den = pmid / (C.RDGAS * t)
    rh = rng.uniform(0.3, 1.05, shape)
    qv = rh * qsat_water(t, den)
    tv = t * (1.0 + C.ZVIR * qv)
    delz = -(C.RDGAS * tv / C.GRAV) * logp

    # cloud layers
    nlay = rng.integers(1, 3, size=(ni, nj))
    cond = np.zeros(shape, dtype=bool)
    for layer in range(2):
        base = rng.uniform(0.0, max(cloud_top - 1000.0, 0.0), size=(ni, nj))
        thick = rng.uniform(1000.0, 5000.0, size=(ni, nj))
        active = (nlay > layer)[..., None]
        inside = (zmid >= base[..., None]) & (zmid <= (base + thick)[..., None])
        cond |= active & inside
    out = {}
    for name, med in (("ql", median), ("qi", median), ("qr", median), ("qs", median),
                      ("qg", qg_median)):
        v = rng.lognormal(np.log(med), 1.0, shape)
        out[name] = np.where(cond, v, 0.0)
    out.update(
        pt=t,
        qv=qv,
        qa=np.zeros(shape),
        ua=rng.normal(0.0, 10.0, shape),
        va=rng.normal(0.0, 10.0, shape),
        delp=np.broadcast_to(delp1, shape).copy(),
        delz=delz,
        w=rng.normal(0.0, 0.5, shape),
        land=(rng.uniform(0.0, 1.0, (ni, nj)) < 0.3).astype(np.float64),
        area=np.full((ni, nj), 1.0e10),
    )
    return out


def make_workload(name, *, tiles=None):
    """The whole workload as one block [ntiles*ni, nj, nz] (tiles stacked on i)."""
    w = WORKLOADS[name]
    tiles = range(w.ntiles) if tiles is None else tiles
    blocks = []
    for tile in tiles:
        off = (-10.0 + 4.0 * tile) if w.tile_offsets else 0.0
        blocks.append(make_state(w.ni, w.nj, w.nz, 1000 * WORKLOAD_IDS[name] + tile,
                                 ptop=w.ptop, t_offset=off, cloud_top=w.cloud_top,
                                 median=w.median, qg_median=w.qg_median))
    if len(blocks) == 1:
        return blocks[0]
    return {k: np.concatenate([b[k] for b in blocks], axis=0) for k in blocks[0]}


STATE_3D = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz", "w")
STATE_2D = ("land", "area")
