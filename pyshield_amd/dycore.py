"""``ApplyPhysicsToDycore`` for one tile on one rank, on the HIP kernels.

Mirrors pySHiELD/update/fv_update_phys.py:77-195: the constructor takes the
namelist and the driver grid data (vlon*, vlat*, es1_*, ew2_*, edge_vect_*,
grid_type) and ``__call__(state, u_dt, v_dt, t_dt, dt)`` applies the physics
tendencies in place through ``shield_apply_physics_to_dycore``: moist_cv,
update_pressure_and_surface_winds and AGrid2DGridPhysics.  Two steps of the
reference's call come from the caller, as callables:

* the one-point halo update of u_dt and v_dt (fv_update_phys.py:174-188):
  ``halo_update(u_dt, v_dt)`` on the halo'd tendencies (the halo of a
  cubed-sphere tile comes from its neighbours; a doubly periodic domain on
  one rank is filled here when no callable is given);
* CubedToLatLon (NDSL, not part of pySHiELD), which recomputes ua, va from
  the updated D-grid winds: ``cubed_to_latlon(u, v, ua, va)``.

Fields are float64 torch tensors on the GPU in the shapes of
``shield_dycore_fields`` (include/shield_mp.h); there is no CPU path.
"""

import ctypes

import torch

from . import _lib

RDGAS = 287.05
CP_AIR = 1004.6
KAPPA = RDGAS / CP_AIR
GRID_NAMES = ("vlon1", "vlon2", "vlon3", "vlat1", "vlat2", "vlat3", "es1_1", "es1_2", "es1_3", "ew2_1", "ew2_2",
              "ew2_3", "edge_vect_w", "edge_vect_e", "edge_vect_s", "edge_vect_n")


def _ptr(t, name, shape):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda:
        raise TypeError(f"{name}: expected a float64 GPU tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def apply_physics_to_dycore(fields: dict, grid: dict, dt: float, npx: int, npy: int, grid_type: int = 0,
                            con_cp: float = CP_AIR, kappa: float = KAPPA, stream=None):
    """One shield_apply_physics_to_dycore call; ``fields`` holds the state,
    u_dt, v_dt, t_dt (see include/shield_mp.h for names and shapes)."""
    ni, nj, nk = fields["pt"].shape
    shapes = {n: (ni, nj, nk) for n in ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel", "pt", "t_dt",
                                        "delp", "ua", "va")}
    shapes.update({n: (ni, nj, nk + 1) for n in ("pe", "peln", "pk")})
    shapes.update({n: (ni, nj) for n in ("ps", "u_srf", "v_srf")})
    shapes.update(u=(ni, nj + 1, nk), v=(ni + 1, nj, nk), u_dt=(ni + 2, nj + 2, nk), v_dt=(ni + 2, nj + 2, nk))
    gshapes = {n: (ni + 2, nj + 2) for n in GRID_NAMES[:6]}
    gshapes.update({n: (ni, nj + 1) for n in ("es1_1", "es1_2", "es1_3")})
    gshapes.update({n: (ni + 1, nj) for n in ("ew2_1", "ew2_2", "ew2_3")})
    gshapes.update(edge_vect_w=(nj,), edge_vect_e=(nj,), edge_vect_s=(ni,), edge_vect_n=(ni,))
    f = _lib.ShieldDycoreFields()
    for n, shp in shapes.items():
        setattr(f, n, _ptr(fields[n], n, shp))
    if grid_type <= 3:
        for n, shp in gshapes.items():
            setattr(f, n, _ptr(grid[n], n, shp))
    f.dt, f.con_cp, f.kappa = float(dt), float(con_cp), float(kappa)
    f.grid_type, f.npx, f.npy = int(grid_type), int(npx), int(npy)
    shape = _lib.ShieldMpShape(ni, nj, nk, nj * nk, nk, nj, 1)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib().shield_apply_physics_to_dycore(ctypes.byref(shape), ctypes.byref(f), ctypes.c_void_p(s)),
               "shield_apply_physics_to_dycore")


def _unwrap(x):
    """A torch tensor from a tensor or a Quantity-like object with .data."""
    return x.data if hasattr(x, "data") and not isinstance(x, torch.Tensor) else x


def periodic_halo_update(u_dt, v_dt):
    """Fill the one-point halo of [ni + 2, nj + 2, nk] tendencies of a doubly
    periodic domain held by one rank (the edges; the corners are never read)."""
    for x in (u_dt, v_dt):
        x[0, 1:-1] = x[-2, 1:-1]
        x[-1, 1:-1] = x[1, 1:-1]
        x[1:-1, 0] = x[1:-1, -2]
        x[1:-1, -1] = x[1:-1, 1]


class ApplyPhysicsToDycore:
    """fv_update_phys.py:77-195 for one tile on one rank.

    ``namelist`` needs npx, npy; ``grid_info`` the grid fields of GRID_NAMES
    and grid_type.  ``halo_update(u_dt, v_dt)`` fills the one-point halo of
    the [ni + 2, nj + 2, nk] tendencies (the reference's u_dt / v_dt halo
    updaters); without one, a doubly periodic domain (grid_type > 3) is filled
    periodically and a cubed-sphere tile must be handed halo'd tendencies
    whose halo is already filled.  ``cubed_to_latlon(u, v, ua, va)`` is NDSL's
    CubedToLatLon, run after the wind update when given.

    ``__call__(state, u_dt, v_dt, t_dt, dt)`` takes the tendencies either with
    their halo ([ni + 2, nj + 2, nk], used in place) or as the compute domain
    ([ni, nj, nk]: copied into a halo'd buffer kept between calls, and zeroed
    afterwards as the reference's set_winds_zero zeroes them when
    grid_type <= 3)."""

    def __init__(self, namelist, grid_info, halo_update=None, cubed_to_latlon=None):
        self.namelist = namelist
        self.grid_type = int(getattr(grid_info, "grid_type", 0))
        self.grid = {n: _unwrap(getattr(grid_info, n)) for n in GRID_NAMES} if self.grid_type <= 3 else {}
        self.halo_update = halo_update
        self.cubed_to_latlon = cubed_to_latlon
        # the surface winds, kept as the reference keeps them, and the halo'd
        # tendency buffers (allocated at the first call, when the tile's
        # shape is known)
        self.u_srf = self.v_srf = None
        self._hu = self._hv = None

    def _halo_buffers(self, ni, nj, nk, device):
        if self._hu is None or tuple(self._hu.shape) != (ni + 2, nj + 2, nk) or self._hu.device != device:
            self._hu = torch.zeros((ni + 2, nj + 2, nk), dtype=torch.float64, device=device)
            self._hv = torch.zeros_like(self._hu)
        return self._hu, self._hv

    def __call__(self, state, u_dt, v_dt, t_dt, dt: float):
        fields = {n: _unwrap(getattr(state, n)) for n in ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel",
                                                          "pt", "delp", "pe", "peln", "pk", "ps", "ua", "va", "u",
                                                          "v")}
        u_dt, v_dt, t_dt = _unwrap(u_dt), _unwrap(v_dt), _unwrap(t_dt)
        ni, nj, nk = fields["pt"].shape
        dev = fields["pt"].device
        if tuple(u_dt.shape) == (ni + 2, nj + 2, nk) and tuple(v_dt.shape) == (ni + 2, nj + 2, nk):
            hu, hv, compact = u_dt, v_dt, False
        elif tuple(u_dt.shape) == (ni, nj, nk) and tuple(v_dt.shape) == (ni, nj, nk):
            if self.halo_update is None and self.grid_type <= 3:
                raise ValueError("compute-domain tendencies on a cubed-sphere tile need halo_update")
            hu, hv = self._halo_buffers(ni, nj, nk, dev)
            hu[1:-1, 1:-1].copy_(u_dt)
            hv[1:-1, 1:-1].copy_(v_dt)
            compact = True
        else:
            raise ValueError(f"u_dt / v_dt: expected {(ni + 2, nj + 2, nk)} or {(ni, nj, nk)}")
        if self.halo_update is not None:
            self.halo_update(hu, hv)
        elif compact:
            periodic_halo_update(hu, hv)
        if self.u_srf is None or tuple(self.u_srf.shape) != (ni, nj) or self.u_srf.device != dev:
            self.u_srf = torch.zeros((ni, nj), dtype=torch.float64, device=dev)
            self.v_srf = torch.zeros_like(self.u_srf)
        fields.update(u_dt=hu, v_dt=hv, t_dt=t_dt, u_srf=self.u_srf, v_srf=self.v_srf)
        apply_physics_to_dycore(fields, self.grid, dt, self.namelist.npx, self.namelist.npy, self.grid_type)
        if compact and self.grid_type <= 3:
            u_dt.zero_()
            v_dt.zero_()
        if self.cubed_to_latlon is not None:
            self.cubed_to_latlon(fields["u"], fields["v"], fields["ua"], fields["va"])
