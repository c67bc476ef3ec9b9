"""Drop-in ``Microphysics`` / ``MicrophysicsState`` on the MI355X kernels.

Mirrors pySHiELD/stencils/microphysics.py:1556-2263: same class names,
constructor and ``__call__(state, timestep)`` signatures, same in-place
semantics (the ten tendencies are accumulated with ``+=``; ``qice``/``qsnow``
and their tendencies are rewritten when ``de_ice``; ``wmp`` when
``do_sedi_w``), same persistent cloud-fraction accumulator ``qaz`` carried
across calls.  Fields are torch tensors on the GPU (float64, [X, Y, Z] with k
contiguous) or objects exposing such a tensor as ``.data`` (a Quantity-like
wrapper); halo offsets are taken from a view of the compute domain.

The whole call runs as one fused HIP kernel through the C ABI
``shield_mp_forward``; there is no CPU path.
"""

import ctypes

import torch

from . import _lib
from ._config import PhysicsConfig


def _tensor(x):
    """The torch tensor behind a field (a tensor or an object with .data)."""
    t = x if isinstance(x, torch.Tensor) else getattr(x, "data", None)
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"expected a torch tensor or an object with a tensor .data, got {type(x)}")
    return t


def _check3(name, t, shape):
    if t.dtype != torch.float64:
        raise TypeError(f"{name}: expected float64, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError(f"{name}: expected a GPU tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.dim() == 3 and t.shape[2] > 1 and t.stride(2) != 1:
        raise ValueError(f"{name}: the k dimension must be contiguous (stride 1)")


def _same_strides(ts, what):
    s0 = ts[0].stride()
    for t in ts[1:]:
        if t.stride() != s0:
            raise ValueError(f"all {what} fields must share the same strides "
                             f"({s0} vs {t.stride()})")
    return s0


class MicrophysicsState:
    """pt, qvapor, qliquid, qrain, qice, qsnow, qgraupel, qcld, ua, va, delp,
    delz, omga: same as the physics state; ``tendency`` is the template the ten
    tendency fields are copied from; delprsi is the layer pressure thickness,
    wmp the vertical velocity and land the land mask
    (microphysics.py:1556-1620)."""

    def __init__(self, pt, qvapor, qliquid, qrain, qice, qsnow, qgraupel, qcld, ua, va,
                 delp, delz, omga, delprsi, wmp, dz, tendency, land):
        self.pt = pt
        self.qvapor = qvapor
        self.qliquid = qliquid
        self.qrain = qrain
        self.qice = qice
        self.qsnow = qsnow
        self.qgraupel = qgraupel
        self.qcld = qcld
        self.ua = ua
        self.va = va
        self.delp = delp
        self.delz = delz
        self.omga = omga
        tend = _tensor(tendency)
        self.qv_dt = tend.clone()
        self.ql_dt = tend.clone()
        self.qr_dt = tend.clone()
        self.qi_dt = tend.clone()
        self.qs_dt = tend.clone()
        self.qg_dt = tend.clone()
        self.qa_dt = tend.clone()
        self.udt = tend.clone()
        self.vdt = tend.clone()
        self.pt_dt = tend.clone()
        self.delprsi = delprsi
        self.wmp = wmp
        self.dz = dz
        self.land = land


class GridIndexing:
    """The part of NDSL's GridIndexing the microphysics reads: the compute
    domain (nx, ny, nz)."""

    def __init__(self, domain):
        self.domain = tuple(int(d) for d in domain)


class StencilFactory:
    """Minimal stand-in for NDSL's StencilFactory: carries the grid indexing
    and the torch device the fields live on."""

    def __init__(self, domain, device="cuda"):
        self.grid_indexing = GridIndexing(domain)
        self.device = torch.device(device)


class QuantityFactory:
    """Allocates zero-initialised float64 device fields (NDSL QuantityFactory.zeros)."""

    def __init__(self, domain, device="cuda"):
        self.domain = tuple(int(d) for d in domain)
        self.device = torch.device(device)

    def zeros(self, dims=None, units="unknown", dtype=torch.float64):
        shape = self.domain if dims is None or len(dims) == 3 else self.domain[:2]
        return torch.zeros(shape, dtype=torch.float64, device=self.device)


class GridData:
    """Grid metrics the microphysics reads: the cell area (microphysics.py:1659)."""

    def __init__(self, area):
        self.area = area


def _domain(stencil_factory):
    if hasattr(stencil_factory, "grid_indexing"):
        return tuple(stencil_factory.grid_indexing.domain)
    return tuple(stencil_factory)


class Microphysics:
    """GFDL single-moment cloud microphysics (microphysics.py:1623-2263).

    ``precision`` (extension): ``"fp64"`` runs every operation in FP64, the
    reference's arithmetic; ``"mixed"`` runs process rates in FP32 with FP64
    state, transfers, sedimentation/flux sums and precipitation.
    """

    def __init__(self, stencil_factory, quantity_factory, grid_data, namelist: PhysicsConfig,
                 precision: str = "fp64"):
        if precision not in _lib.MODES:
            raise ValueError(f"precision must be one of {sorted(_lib.MODES)}")
        self.namelist = namelist
        self._mode = _lib.MODES[precision]
        self.precision = precision
        nx, ny, nz = _domain(stencil_factory)[:3]
        self._domain = (nx, ny, nz)
        self._area = grid_data.area
        device = getattr(quantity_factory, "device", None) or _tensor(grid_data.area).device
        # persistent device state: the cloud-fraction accumulator (never reset
        # between calls, microphysics.py:1674) and the surface precipitation
        # of the last call in mm/day (kept private by the reference,
        # microphysics.py:1664-1667; exposed here as 2D fields)
        self._qaz = torch.zeros((nx, ny, nz), dtype=torch.float64, device=device)
        self.rain = torch.zeros((nx, ny), dtype=torch.float64, device=device)
        self.snow = torch.zeros_like(self.rain)
        self.ice = torch.zeros_like(self.rain)
        self.graupel = torch.zeros_like(self.rain)
        self._params = _lib.setup(namelist.mp_namelist(), float(namelist.dt_atmos or 1.0))
        self._timestep = float(self._params.timestep)

    # microphysics.py:1959-1962
    def _update_timestep_if_needed(self, timestep: float):
        if timestep != self._timestep:
            _lib.check(_lib.lib().shield_mp_update_timestep(ctypes.byref(self._params), float(timestep)),
                       "shield_mp_update_timestep")
            self._timestep = float(timestep)

    @property
    def params(self):
        return self._params

    @property
    def qaz(self):
        return self._qaz

    def __call__(self, state: MicrophysicsState, timestep: float):
        self._update_timestep_if_needed(timestep)
        nx, ny, nz = self._domain
        fields = {
            "pt": state.pt, "qv": state.qvapor, "ql": state.qliquid, "qr": state.qrain,
            "qi": state.qice, "qs": state.qsnow, "qg": state.qgraupel, "qa": state.qcld,
            "ua": state.ua, "va": state.va, "delp": state.delprsi, "delz": state.delz,
            "w": state.wmp,
            "pt_dt": state.pt_dt, "qv_dt": state.qv_dt, "ql_dt": state.ql_dt, "qr_dt": state.qr_dt,
            "qi_dt": state.qi_dt, "qs_dt": state.qs_dt, "qg_dt": state.qg_dt, "qa_dt": state.qa_dt,
            "udt": state.udt, "vdt": state.vdt,
            "qaz": self._qaz,
            "land": state.land, "area": self._area,
            "rain": self.rain, "snow": self.snow, "ice": self.ice, "graupel": self.graupel,
        }
        forward({k: _tensor(v) for k, v in fields.items()}, self._params, self._mode)


FIELDS_3D = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz", "w",
             "pt_dt", "qv_dt", "ql_dt", "qr_dt", "qi_dt", "qs_dt", "qg_dt", "qa_dt", "udt", "vdt",
             "qaz")
FIELDS_2D = ("land", "area", "rain", "snow", "ice", "graupel")


def _mp_args(fields: dict):
    """(shape, shield_mp_fields) of a field dict; checks dtypes, shapes, strides."""
    pt = fields["pt"]
    ni, nj, nk = pt.shape
    t3 = [fields[n] for n in FIELDS_3D if fields.get(n) is not None]
    t2 = [fields[n] for n in FIELDS_2D if fields.get(n) is not None]
    for n in FIELDS_3D:
        if fields.get(n) is not None:
            _check3(n, fields[n], (ni, nj, nk))
    for n in FIELDS_2D:
        if fields.get(n) is not None:
            _check3(n, fields[n], (ni, nj))
    si, sj, _ = _same_strides(t3, "3D")
    si2, sj2 = _same_strides(t2, "2D")
    shape = _lib.ShieldMpShape(ni, nj, nk, si, sj, si2, sj2)
    f = _lib.ShieldMpFields()
    for n in _lib.FIELD_NAMES:
        t = fields.get(n)
        setattr(f, n, t.data_ptr() if t is not None else None)
    return shape, f


def forward(fields: dict, params, mode=_lib.SHIELD_MP_FP64, stream=None):
    """One microphysics call on tensors: ``fields`` maps the names of
    shield_mp_fields to float64 GPU tensors ([ni, nj, nk], k contiguous; 2D
    [ni, nj]).  Launches on ``stream`` (default: torch's current stream)."""
    if isinstance(mode, str):
        mode = _lib.MODES[mode]
    shape, f = _mp_args(fields)
    if stream is None:
        stream = torch.cuda.current_stream(fields["pt"].device).cuda_stream
    rc = _lib.lib().shield_mp_forward(ctypes.byref(shape), ctypes.byref(params), ctypes.byref(f),
                                      int(mode), ctypes.c_void_p(stream))
    _lib.check(rc, "shield_mp_forward")


def forward_update(fields: dict, update: dict, params, dt: float, rdt: float, q_min: float = 1.0e-9,
                   mode=_lib.SHIELD_MP_FP64, stream=None):
    """The microphysics call and, in the same kernel, the state update that
    Physics.__call__ and UpdateAtmosphereState.__call__ run after it
    (physics.py:331-366, update_atmos_state.py:301-349 without
    ApplyPhysicsToDycore): ``physics_updated_*`` by forward Euler over ``dt``,
    fill_gfs_delp of the updated vapour (threshold ``q_min``) with the dycore
    delp, and the GFS -> FV3 tracer and delp conversion with the dycore
    tendencies accumulated over ``rdt``.

    ``update`` holds: "updated" (dict of the ten physics_updated tensors keyed
    by pyshield_amd._lib.UPDATED_NAMES), "u_dt", "v_dt", "pt_dt", "tracers"
    (dict keyed by TRACER_NAMES), "delp" (all [ni, nj, nk] with the strides of
    the microphysics fields) and "prsi" ([ni, nj, nk + 1], k contiguous)."""
    if isinstance(mode, str):
        mode = _lib.MODES[mode]
    shape, f = _mp_args(fields)
    ni, nj, nk = fields["pt"].shape
    ref = fields["pt"]
    layer = [update["updated"][n] for n in _lib.UPDATED_NAMES]
    layer += [update["u_dt"], update["v_dt"], update["pt_dt"], update["delp"]]
    layer += [update["tracers"][n] for n in _lib.TRACER_NAMES]
    for t in layer:
        _check3("update field", t, (ni, nj, nk))
        if t.stride()[:2] != ref.stride()[:2]:
            raise ValueError("state-update fields must have the strides of the microphysics fields")
    prsi = update["prsi"]
    _check3("prsi", prsi, (ni, nj, nk + 1))
    u = _lib.ShieldUpdateFields()
    for i, n in enumerate(_lib.UPDATED_NAMES):
        u.updated[i] = update["updated"][n].data_ptr()
    for i, n in enumerate(_lib.TRACER_NAMES):
        u.tracers[i] = update["tracers"][n].data_ptr()
    u.u_dt, u.v_dt, u.pt_dt = (update[k].data_ptr() for k in ("u_dt", "v_dt", "pt_dt"))
    u.delp, u.prsi = update["delp"].data_ptr(), prsi.data_ptr()
    u.si_i, u.sj_i = prsi.stride(0), prsi.stride(1)
    u.dt, u.rdt, u.q_min = float(dt), float(rdt), float(q_min)
    if stream is None:
        stream = torch.cuda.current_stream(ref.device).cuda_stream
    rc = _lib.lib().shield_mp_forward_update(ctypes.byref(shape), ctypes.byref(params), ctypes.byref(f),
                                             ctypes.byref(u), int(mode), ctypes.c_void_p(stream))
    _lib.check(rc, "shield_mp_forward_update")


# ---------------------------------------------------------------- adjoint
AD_STATE_NAMES = ("pt", "qv", "ql", "qr", "qi", "qs", "qg", "qa", "ua", "va", "delp", "delz")
AD_OUT_NAMES = _lib.ADJ_OUT_NAMES


def ad_params(params):
    """The six differentiable process parameters of a shield_mp_params, in the
    order of shield_mp_adjoint_fields.params: fac_rc (autoconversion
    threshold qc0 = fac_rc ccn), vr_fac, vs_fac, vg_fac, vi_fac (fall-speed
    coefficients), qi_gen (ice generation)."""
    n = params.nml
    return [params.fac_rc, n.vr_fac, n.vs_fac, n.vg_fac, n.vi_fac, n.qi_gen]


def with_ad_params(params, values):
    """A copy of ``params`` whose six process parameters are ``values``."""
    p = _lib.ShieldMpParams()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(params), ctypes.sizeof(p))
    v = [float(x) for x in values]
    p.fac_rc = v[0]
    p.nml.vr_fac, p.nml.vs_fac, p.nml.vg_fac, p.nml.vi_fac, p.nml.qi_gen = v[1:]
    return p


def adjoint(fields: dict, cotangents: dict, grads: dict, params, workspace=None, waves=2048, stream=None):
    """Reverse-mode adjoint of one FP64 forward call (shield_mp_adjoint).
    ``fields``: the forward's inputs (as for forward(); output entries are not
    used).  ``cotangents``: cotangents of the outputs keyed by AD_OUT_NAMES
    (the 2D precipitation entries may be missing) and, with do_sedi_w, "w"
    (optional, the vertical velocity the call rewrites).  ``grads``: tensors
    that receive the input cotangents, keyed by AD_STATE_NAMES and "qaz"
    (optional), "w" (optional, do_sedi_w) and "params" (6 float64 values).  All 3D and 2D tensors share
    the strides of fields["pt"] and fields["land"].  Returns the workspace
    (reusable for the next call)."""
    shape, f = _mp_args(fields)
    ref3, ref2 = fields["pt"], fields["land"]
    a = _lib.ShieldMpAdjointFields()
    for n in AD_OUT_NAMES:
        t = cotangents.get(n)
        if t is not None:
            ref = ref2 if n in FIELDS_2D else ref3
            _check3("cotangent " + n, t, tuple(ref.shape))
            if t.stride() != ref.stride():
                raise ValueError(f"cotangent {n} must have the strides of the forward's fields")
        setattr(a, "d_" + n, t.data_ptr() if t is not None else None)
    for n in AD_STATE_NAMES + ("qaz",):
        t = grads.get(n)
        if t is not None:
            _check3("gradient " + n, t, tuple(ref3.shape))
            if t.stride() != ref3.stride():
                raise ValueError(f"gradient {n} must have the strides of the forward's fields")
        setattr(a, "g_" + n, t.data_ptr() if t is not None else None)
    # do_sedi_w: the cotangent of w after the call (None: zero) and the
    # tensor that receives that of w before it (None: not written)
    for key, attr, src in (("w", "d_w", cotangents), ("w", "g_w", grads)):
        t = src.get(key)
        if t is not None:
            _check3(f"{attr[0]}_w", t, tuple(ref3.shape))
            if t.stride() != ref3.stride():
                raise ValueError("w's cotangents must have the strides of the forward's fields")
        setattr(a, attr, t.data_ptr() if t is not None else None)
    gp = grads["params"]
    if gp.dtype != torch.float64 or gp.numel() != 6 or not gp.is_contiguous() or gp.device != ref3.device:
        raise ValueError("grads['params'] must be a contiguous float64 tensor of 6 values on the fields' device")
    a.params = gp.data_ptr()
    L = _lib.lib()
    nbytes = int(L.shield_mp_adjoint_workspace_bytes(ctypes.byref(shape), ctypes.byref(params), int(waves)))
    if nbytes == 0:
        raise _lib.ShieldError("shield_mp_adjoint: unsupported shape (1 <= ni * nj, 2 <= nk <= 128)")
    if workspace is None or workspace.numel() * 8 < nbytes or workspace.device != ref3.device:
        workspace = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=ref3.device)
    if stream is None:
        stream = torch.cuda.current_stream(ref3.device).cuda_stream
    rc = L.shield_mp_adjoint(ctypes.byref(shape), ctypes.byref(params), ctypes.byref(f), ctypes.byref(a),
                             ctypes.c_void_p(workspace.data_ptr()), ctypes.c_size_t(workspace.numel() * 8),
                             ctypes.c_void_p(stream))
    _lib.check(rc, "shield_mp_adjoint")
    return workspace


def _strided_like(ref):
    """An uninitialised float64 tensor with ref's shape and strides (halo'd
    Quantity views keep their layout, which forward() and adjoint() require
    of every field)."""
    return torch.empty_strided(tuple(ref.shape), tuple(ref.stride()), dtype=torch.float64, device=ref.device)


# the adjoint's workspace, kept between backward calls (per device)
_AD_WORKSPACE = {}


class MicrophysicsFunction(torch.autograd.Function):
    """One microphysics call as a differentiable function of the twelve state
    inputs (AD_STATE_NAMES), the incoming qaz and the six process parameters
    (a float64 tensor in the order of ad_params).  Returns the ten tendencies
    and the four precipitation rates (AD_OUT_NAMES).  The forward runs the
    HIP forward in ``mode``; the backward runs the HIP adjoint, which
    recomputes the FP64 forward of every column and follows the branch it
    takes at each point.  Tendencies are those of this call (overwritten,
    not accumulated).  Halo'd views are accepted: outputs and gradients take
    the state's strides.  The adjoint's workspace is kept for the next
    backward on the same device."""

    @staticmethod
    def forward(ctx, params, mode, adp, land, area, qaz, *state):
        p = with_ad_params(params, adp.detach().cpu().tolist())
        p.accumulate = 0
        ref = state[0]
        f = {n: t.detach() for n, t in zip(AD_STATE_NAMES, state)}
        f.update(land=land.detach(), area=area.detach())
        f["qaz"] = _strided_like(ref).copy_(qaz.detach())
        f["w"] = _strided_like(ref).zero_()
        out = {n: _strided_like(land if n in FIELDS_2D else ref) for n in AD_OUT_NAMES}
        f.update(out)
        forward(f, p, mode)
        ctx.p = p
        ctx.save_for_backward(land, area, qaz, *state)
        return tuple(out[n] for n in AD_OUT_NAMES)

    @staticmethod
    def backward(ctx, *gout):
        land, area, qaz, *state = ctx.saved_tensors
        ref = state[0]
        f = {n: t.detach() for n, t in zip(AD_STATE_NAMES, state)}
        f.update(land=land, area=area, qaz=qaz)
        # the forward's w (do_sedi_w), zero: w reaches none of the outputs
        f["w"] = _strided_like(ref).zero_()
        cot = {}
        for n, g in zip(AD_OUT_NAMES, gout):
            like = land if n in FIELDS_2D else ref
            cot[n] = _strided_like(like).zero_() if g is None else _strided_like(like).copy_(g)
        grads = {n: _strided_like(ref) for n in AD_STATE_NAMES + ("qaz",)}
        grads["params"] = torch.zeros(6, dtype=torch.float64, device=ref.device)
        key = ref.device
        _AD_WORKSPACE[key] = adjoint(f, cot, grads, ctx.p, workspace=_AD_WORKSPACE.get(key))
        gadp = grads["params"].cpu()
        return (None, None, gadp, None, None, grads["qaz"]) + tuple(grads[n] for n in AD_STATE_NAMES)


def microphysics_ad(state: dict, params, adp=None, mode="fp64"):
    """Differentiable microphysics call on a state dict (AD_STATE_NAMES plus
    land, area and qaz, all [ni, nj, nk] / [ni, nj] float64 CUDA tensors with
    k contiguous).  ``adp``: the six process parameters as a CPU float64
    tensor (default: those of ``params``).  Returns a dict of AD_OUT_NAMES."""
    if adp is None:
        adp = torch.tensor(ad_params(params), dtype=torch.float64)
    outs = MicrophysicsFunction.apply(params, mode, adp, state["land"], state["area"], state["qaz"],
                                      *[state[n] for n in AD_STATE_NAMES])
    return dict(zip(AD_OUT_NAMES, outs))
