"""pySHiELD's Physics driver (pySHiELD/stencils/physics.py:196-366) on the GPU:
the column-local prologue (atmos_phys_driver_statein, get_prs_fv3,
get_phi_fv3 and prepare_microphysics) in one kernel, the microphysics, and
the forward-Euler state update.

Fields are float64 GPU tensors (or objects holding one in ``.data``) with k
contiguous and k = 0 at the model top: layer fields [X, Y, nz], interface
fields (phii, prsi) [X, Y, nz + 1].  The reference keeps nz + 1 levels for
every field; the extra level of a layer field is padding no result reads.
"""

import ctypes

import torch

from . import _lib
from ._config import PhysicsConfig
from .microphysics import Microphysics, _domain, _tensor
from .update import update_physics_state_with_tendencies

SPECIES = ("qvapor", "qliquid", "qrain", "qice", "qsnow", "qgraupel", "qo3mr")


def _geom(t, what, shape):
    if t.dtype != torch.float64 or not t.is_cuda:
        raise TypeError(f"{what}: expected a float64 GPU tensor")
    if tuple(t.shape) != tuple(shape) or t.stride(2) != 1:
        raise ValueError(f"{what}: expected shape {tuple(shape)} with contiguous k")


def physics_prologue(delz, delp, q: dict, qsgs_tke, pt, phii, prsi, prsik, dm, delprsi, del_gz, phil,
                     ptop: float, omga=None, dz=None, wmp=None, tend=None, stream=None):
    """atmos_phys_driver_statein, get_prs_fv3, get_phi_fv3 and (with omga)
    prepare_microphysics on the GPU, in place: ``q`` maps SPECIES to tensors;
    ``delp`` comes back as the layer-mean pressure, as in the reference;
    ``tend`` lists the ten tendencies prepare_microphysics zeroes (udt, vdt,
    pt_dt, qv_dt, ql_dt, qr_dt, qi_dt, qs_dt, qg_dt, qa_dt)."""
    T = _tensor
    delp = T(delp)
    nx, ny, nz = delp.shape
    layer = [T(delz), delp, *[T(q[n]) for n in SPECIES], T(qsgs_tke), T(pt), T(dm), T(delprsi), T(del_gz), T(phil)]
    if omga is not None:
        layer += [T(omga), T(dz), T(wmp), *[T(x) for x in tend]]
    for x in layer:
        _geom(x, "layer field", (nx, ny, nz))
        if x.stride()[:2] != delp.stride()[:2]:
            raise ValueError("layer fields must share their i/j strides")
    iface = [T(phii), T(prsi), T(prsik)]
    for x in iface:
        _geom(x, "interface field", (nx, ny, nz + 1))
        if x.stride()[:2] != iface[0].stride()[:2]:
            raise ValueError("interface fields must share their i/j strides")
    f = _lib.ShieldPrologueFields()
    f.delz, f.delp = T(delz).data_ptr(), delp.data_ptr()
    for i, n in enumerate(SPECIES):
        f.q[i] = T(q[n]).data_ptr()
    f.qsgs_tke, f.pt = T(qsgs_tke).data_ptr(), T(pt).data_ptr()
    f.phii, f.prsi, f.prsik = (x.data_ptr() for x in iface)
    f.dm, f.delprsi, f.del_gz, f.phil = (T(x).data_ptr() for x in (dm, delprsi, del_gz, phil))
    if omga is not None:
        f.omga, f.dz, f.wmp = (T(x).data_ptr() for x in (omga, dz, wmp))
        for i, x in enumerate(tend):
            f.tend[i] = T(x).data_ptr()
    f.si_i, f.sj_i = iface[0].stride(0), iface[0].stride(1)
    f.ptop = float(ptop)
    shape = _lib.ShieldMpShape(nx, ny, nz, delp.stride(0), delp.stride(1), 0, 0)
    if stream is None:
        stream = torch.cuda.current_stream(delp.device).cuda_stream
    _lib.check(_lib.lib().shield_physics_prologue(ctypes.byref(shape), ctypes.byref(f), ctypes.c_void_p(stream)),
               "shield_physics_prologue")


class Physics:
    """Physics(stencil_factory, quantity_factory, grid_data, namelist) and
    __call__(physics_state, timestep) of physics.py:196-366.  ``grid_data``
    provides ``ptop`` and ``area``; ``physics_state`` the reference's
    PhysicsState fields (phii, prsi, delz, delp, the seven tracers, qsgs_tke,
    qcld, pt, ua, va, omga, delprsi, dz, wmp, phil, the physics_updated_*
    fields) and, with GFS microphysics, ``microphysics`` (a MicrophysicsState
    over the same fields).  ``precision`` (extension) selects the
    microphysics precision."""

    def __init__(self, stencil_factory, quantity_factory, grid_data, namelist: PhysicsConfig,
                 precision: str = "fp64"):
        schemes = [getattr(s, "value", s) for s in namelist.schemes]
        self._ptop = float(grid_data.ptop)
        nx, ny, nz = _domain(stencil_factory)[:3]
        device = getattr(quantity_factory, "device", None) or _tensor(grid_data.area).device
        z = lambda k: torch.zeros((nx, ny, k), dtype=torch.float64, device=device)  # noqa: E731
        self._prsik = z(nz + 1)
        self._dm3d = z(nz)
        self._del_gz = z(nz)
        self._gfs_microphysics = "GFS_microphysics" in schemes
        if self._gfs_microphysics:
            self._microphysics = Microphysics(stencil_factory, quantity_factory, grid_data, namelist,
                                              precision=precision)

    @property
    def microphysics(self):
        return self._microphysics if self._gfs_microphysics else None

    def __call__(self, physics_state, timestep: float):
        s = physics_state
        q = {n: getattr(s, n) for n in SPECIES}
        extra = {}
        if self._gfs_microphysics:
            m = s.microphysics
            extra = dict(omga=s.omga, dz=s.dz, wmp=s.wmp,
                         tend=[m.udt, m.vdt, m.pt_dt, m.qv_dt, m.ql_dt, m.qr_dt, m.qi_dt, m.qs_dt, m.qg_dt,
                               m.qa_dt])
        physics_prologue(s.delz, s.delp, q, s.qsgs_tke, s.pt, s.phii, s.prsi, self._prsik, self._dm3d, s.delprsi,
                         self._del_gz, s.phil, self._ptop, **extra)
        if not self._gfs_microphysics:
            return
        m = s.microphysics
        self._microphysics(m, timestep)
        T = _tensor
        update_physics_state_with_tendencies(
            T(s.qvapor), T(s.qliquid), T(s.qrain), T(s.qice), T(s.qsnow), T(s.qgraupel), T(s.qcld), T(s.pt),
            T(s.ua), T(s.va), T(m.qv_dt), T(m.ql_dt), T(m.qr_dt), T(m.qi_dt), T(m.qs_dt), T(m.qg_dt),
            T(m.qa_dt), T(m.pt_dt), T(m.udt), T(m.vdt), T(s.physics_updated_specific_humidity),
            T(s.physics_updated_qliquid), T(s.physics_updated_qrain), T(s.physics_updated_qice),
            T(s.physics_updated_qsnow), T(s.physics_updated_qgraupel), T(s.physics_updated_cloud_fraction),
            T(s.physics_updated_pt), T(s.physics_updated_ua), T(s.physics_updated_va), timestep)
