"""CPU oracle for the tsw_rtm RTM path (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module, and only as the checker or the timed CPU baseline.  The
das_amd product never imports it.

Wraps ``oracle/liboracle.so`` (rtm_oracle.cpp, the C++ restatement of
tsw_rtm/c_engine/wave2d.c and of the time loops in tsw_rtm/wave_solver.py) and
restates, with numpy, the multi-shot post-processing of
tsw_rtm/rtm_engine.py:106-171.  ``ref_lib()`` loads the reference engine
compiled from its own source by oracle/Makefile (``oracle/_ref``), when built.

Pinning of each function is done in tests/test_oracle.py against the golden
vectors in tests/golden/ (made from wave2d.c compiled from source and from the
reference's numpy/scipy expressions) and against invariants.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libwave2d_ref.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_L = ctypes.c_longlong

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        for sfx, T in (("f32", _F), ("f64", _D)):
            getattr(L, f"ora_acoustic_step_{sfx}").argtypes = [_P] * 5 + [_I, _I, T, T, T, _I]
            getattr(L, f"ora_elastic_step_{sfx}").argtypes = [_P] * 10 + [_I, _I, T, T, T, _I]
            getattr(L, f"ora_correlate_{sfx}").argtypes = [_P, _P, _P, _I, _I, T]
            getattr(L, f"ora_forward_{sfx}").argtypes = (
                [_P, _P, _I, _I, _D, _D, _D, _I, _P, _I, _I, _I, _P, _I, _I, _I, _P, _P])
            getattr(L, f"ora_backward_{sfx}").argtypes = (
                [_P, _P, _I, _I, _D, _D, _D, _I, _P, _I, _P, _I, _I, _P, _L, _P, _I])
        L.ora_build_sponge_f32.argtypes = [_P, _I, _I, _I, _F, _I]
        L.ora_build_sponge_f64.argtypes = [_P, _I, _I, _I, _D, _I]
        L.ora_migrate_shot_f32.argtypes = (
            [_P, _P, _P, _I, _I, _D, _D, _D, _I, _P, _I, _I, _P, _I, _P, _I, _I, _I, _P, _P, _P, _P])
        L.ora_migrate_shot_f32.restype = _I
        L.ora_migrate_shot_f64.argtypes = (
            [_P, _P, _P, _I, _I, _D, _D, _D, _I, _P, _I, _I, _P, _I, _P, _I, _I, _I, _P, _P])
        L.ora_migrate_shot_f64.restype = _I
        L.ora_laplacian_f32.argtypes = [_P, _P, _I, _I, _D, _D]
        for sfx in ("f32", "f64"):
            getattr(L, f"ora_elastic_forward_{sfx}").argtypes = (
                [_P] * 5 + [_I, _I, _D, _D, _D, _I, _P, _I, _I, _I, _P, _I, _I, _I, _P, _P, _P])
            getattr(L, f"ora_elastic_backward_{sfx}").argtypes = (
                [_P] * 5 + [_I, _I, _D, _D, _D, _I, _P, _I, _P, _I, _I, _P, _L, _P, _I])
        for sfx in ("f32", "f64"):
            f = getattr(L, f"ora_misfit_gradient_{sfx}")
            f.argtypes = ([_P, _P, _I, _I, _D, _D, _D, _I, _P, _I, _I, _I, _P, _I, _I]
                          + [_P] * 5)
            f.restype = _D
        _lib = L
    return _lib


def ref_lib():
    """wave2d.c compiled from source (oracle/Makefile), or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = ctypes.CDLL(REF_PATH)
        R.acoustic_2d_step.argtypes = [_P] * 5 + [_I, _I, _F, _F, _F, _I]
        R.elastic_2d_step.argtypes = [_P] * 10 + [_I, _I, _F, _F, _F, _I]
        R.correlate_imaging_condition.argtypes = [_P, _P, _P, _I, _I, _F]
        R.build_cerjan_sponge.argtypes = [_P, _I, _I, _I, _F, _I]
        _ref = R
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _sfx(dt):
    return "f32" if np.dtype(dt) == np.float32 else "f64"


# ---------------------------------------------------------------- kernels
def acoustic_step(p_curr, p_prev, p_next, vel, sponge, dx, dz, dt, free_surface=True):
    """In place on ``p_next`` (wave2d.c:19-82).  dtype from p_next."""
    T = p_next.dtype
    nz, nx = p_next.shape
    # converted copies are held in `args` until the call returns
    args = [_c(p_curr, T), _c(p_prev, T), None, _c(vel, T), None if sponge is None else _c(sponge, T)]
    getattr(lib(), f"ora_acoustic_step_{_sfx(T)}")(
        _ptr(args[0]), _ptr(args[1]), _ptr(p_next), _ptr(args[3]), _ptr(args[4]),
        nx, nz, dx, dz, dt, int(bool(free_surface)))
    return p_next


def elastic_step(fields, lam, mu, bx, bz, sponge, dx, dz, dt, free_surface=True):
    """In place on the 5 arrays of ``fields`` = (vx, vz, sxx, szz, sxz)."""
    T = fields[0].dtype
    nz, nx = fields[0].shape
    par = [_c(a, T) for a in (lam, mu, bx, bz)]
    sp = None if sponge is None else _c(sponge, T)
    getattr(lib(), f"ora_elastic_step_{_sfx(T)}")(
        *[_ptr(f) for f in fields], *[_ptr(p) for p in par], _ptr(sp),
        nx, nz, dx, dz, dt, int(bool(free_surface)))
    return fields


def correlate(src, rec, image, w):
    T = image.dtype
    nz, nx = image.shape
    a, b = _c(src, T), _c(rec, T)
    getattr(lib(), f"ora_correlate_{_sfx(T)}")(_ptr(a), _ptr(b), _ptr(image), nx, nz, w)
    return image


def build_sponge(nx, nz, n_sponge=30, alpha=0.0015, free_surface=True, dtype=np.float32):
    out = np.ones((nz, nx), dtype=dtype)
    getattr(lib(), f"ora_build_sponge_{_sfx(dtype)}")(
        _ptr(out), nx, nz, n_sponge, alpha, int(bool(free_surface)))
    return out


def laplacian(image, dx, dz):
    img = _c(image, np.float32)
    out = np.empty_like(img)
    lib().ora_laplacian_f32(_ptr(img), _ptr(out), img.shape[1], img.shape[0], dx, dz)
    return out


# ---------------------------------------------------------------- loops
@dataclass
class Grid:
    nx: int
    nz: int
    dx: float
    dz: float
    dt: float
    n_sponge: int = 30
    sponge_alpha: float = 0.0015
    free_surface: bool = True

    def sponge(self, dtype=np.float32):
        return build_sponge(self.nx, self.nz, self.n_sponge, self.sponge_alpha,
                            self.free_surface, dtype)


def forward(g: Grid, vel, src, six, siz, rix, riz=2, k=1, save=False, dtype=np.float32):
    """wave_solver.py:128-186 -> (rec (nt,nrec), snaps or None)."""
    T = np.dtype(dtype)
    src = _c(src, T)
    nt = len(src)
    rix = _c(rix, np.int32)
    rec = np.zeros((nt, len(rix)), dtype=T)
    snaps = np.zeros(((nt - 1) // k + 1, g.nz, g.nx), dtype=T) if save else None
    sp = g.sponge(T)
    vel = _c(vel, T)
    getattr(lib(), f"ora_forward_{_sfx(T)}")(
        _ptr(vel), _ptr(sp), g.nx, g.nz, g.dx, g.dz, g.dt, int(g.free_surface),
        _ptr(src), nt, int(six), int(siz), _ptr(rix), len(rix), int(riz), int(k),
        _ptr(rec), _ptr(snaps))
    return rec, snaps


def backward(g: Grid, vel, data, rix, riz, snaps, image, k=1):
    """wave_solver.py:188-242; accumulates into ``image`` in place."""
    T = image.dtype
    data = _c(data, T)
    nt, nrec = data.shape
    rix = _c(rix, np.int32)
    snaps = _c(snaps, T)
    sp = g.sponge(T)
    vel = _c(vel, T)
    getattr(lib(), f"ora_backward_{_sfx(T)}")(
        _ptr(vel), _ptr(sp), g.nx, g.nz, g.dx, g.dz, g.dt, int(g.free_surface),
        _ptr(data), nt, _ptr(rix), nrec, int(riz), _ptr(snaps), len(snaps), _ptr(image), int(k))
    return image


def migrate_shot(g: Grid, vel_fwd, vel_bwd, src, six, siz, data, rix, riz=2, k=2,
                 dtype=np.float32, use_reference_kernels=False):
    """rtm_engine.py:49-104 -> (image, illumination).

    ``use_reference_kernels`` drives wave2d.c compiled from source (FP32)."""
    T = np.dtype(dtype)
    src = _c(src, T)
    data = _c(data, T)
    nt = data.shape[0]
    rix = _c(rix, np.int32)
    img = np.zeros((g.nz, g.nx), dtype=T)
    ill = np.zeros((g.nz, g.nx), dtype=T)
    sp = g.sponge(T)
    vel_fwd, vel_bwd = _c(vel_fwd, T), _c(vel_bwd, T)
    common = [_ptr(vel_fwd), _ptr(vel_bwd), _ptr(sp), g.nx, g.nz, g.dx, g.dz, g.dt,
              int(g.free_surface), _ptr(src), int(six), int(siz), _ptr(data), nt, _ptr(rix),
              len(rix), int(riz), int(k), _ptr(img), _ptr(ill)]
    if T == np.float32:
        es = ec = None
        if use_reference_kernels:
            R = ref_lib()
            if R is None:
                raise FileNotFoundError(REF_PATH)
            es = ctypes.cast(R.acoustic_2d_step, ctypes.c_void_p).value
            ec = ctypes.cast(R.correlate_imaging_condition, ctypes.c_void_p).value
        rc = lib().ora_migrate_shot_f32(*common, es, ec)
    else:
        rc = lib().ora_migrate_shot_f64(*common)
    if rc != 0:
        raise MemoryError("oracle snapshot allocation failed")
    return img, ill


def postprocess(g: Grid, stacked, illum, mute_sponge=True, sponge=None):
    """rtm_engine.py:145-171 -> (raw, laplacian, normalized)."""
    stacked = np.array(stacked, dtype=np.float32, copy=True)
    illum = np.array(illum, dtype=np.float32, copy=True)
    sp = g.sponge(np.float32) if sponge is None else sponge
    mask = sp < 0.99
    if mute_sponge:
        stacked[mask] = 0.0
        illum[mask] = 0.0
    mx = np.max(illum)
    eps = 1e-4 * mx if mx > 0 else 1e-12
    normalized = stacked / (np.sqrt(illum) + eps)
    lap = laplacian(stacked, g.dx, g.dz)
    if mute_sponge:
        lap[mask] = 0.0
    return stacked, lap, normalized


def migrate_multiple(g: Grid, vel_fwd, vel_bwd, sources, six_list, siz_list, gathers, rix, riz=2,
                     k=2, mute_sponge=True, dtype=np.float32):
    """rtm_engine.py:106-171 (stack in FP32, shot order)."""
    stacked = np.zeros((g.nz, g.nx), dtype=np.float32)
    illum = np.zeros((g.nz, g.nx), dtype=np.float32)
    for src, six, siz, d in zip(sources, six_list, siz_list, gathers):
        im, il = migrate_shot(g, vel_fwd, vel_bwd, src, six, siz, d, rix, riz, k, dtype)
        stacked += im.astype(np.float32)
        illum += il.astype(np.float32)
    return postprocess(g, stacked, illum, mute_sponge)


def ricker(f0, dt, t_length):
    """wave_solver.py:74-81 closed form."""
    nt = int(round(t_length / dt))
    t = np.arange(nt) * dt
    tau = t - 1.0 / f0
    a = (np.pi * f0 * tau) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32), t


# ---------------------------------------------------------------- elastic
def elastic_parameters(vp, vs, rho):
    """(lambda, mu, buoyancy_x, buoyancy_z) from Vp, Vs, rho (cell-centred)."""
    vp = np.asarray(vp, np.float64)
    vs = np.asarray(vs, np.float64)
    rho = np.asarray(rho, np.float64)
    mu = rho * vs * vs
    lam = rho * vp * vp - 2.0 * mu
    b = 1.0 / rho
    return tuple(np.ascontiguousarray(a, np.float32) for a in (lam, mu, b, b))


def elastic_forward(g: Grid, model, src, six, siz, rix, riz=2, k=1, save=False, illum=False,
                    dtype=np.float32):
    """Elastic forward loop -> (rec (nt, nrec) strain rate, vz snaps or None,
    sum of squared vz snapshots or None)."""
    T = np.dtype(dtype)
    par = [_c(a, T) for a in model]
    src = _c(src, T)
    nt = len(src)
    rix = _c(rix, np.int32)
    rec = np.zeros((nt, len(rix)), dtype=T)
    ns = (nt - 1) // k + 1
    snaps = np.zeros((ns, g.nz, g.nx), dtype=T) if save else None
    ill = np.zeros((g.nz, g.nx), dtype=T) if illum else None
    sp = g.sponge(T)
    getattr(lib(), f"ora_elastic_forward_{_sfx(T)}")(
        *[_ptr(a) for a in par], _ptr(sp), g.nx, g.nz, g.dx, g.dz, g.dt, int(g.free_surface),
        _ptr(src), nt, int(six), int(siz), _ptr(rix), len(rix), int(riz), int(k), _ptr(rec),
        _ptr(snaps), _ptr(ill))
    return rec, snaps, ill


def elastic_backward(g: Grid, model, data, rix, riz, snaps, image, k=1):
    T = image.dtype
    par = [_c(a, T) for a in model]
    data = _c(data, T)
    nt, nrec = data.shape
    rix = _c(rix, np.int32)
    snaps = _c(snaps, T)
    sp = g.sponge(T)
    getattr(lib(), f"ora_elastic_backward_{_sfx(T)}")(
        *[_ptr(a) for a in par], _ptr(sp), g.nx, g.nz, g.dx, g.dz, g.dt, int(g.free_surface),
        _ptr(data), nt, _ptr(rix), nrec, int(riz), _ptr(snaps), len(snaps), _ptr(image), int(k))
    return image


def elastic_migrate_shot(g: Grid, model_fwd, model_bwd, src, six, siz, data, rix, riz=2, k=2,
                         dtype=np.float32):
    """Elastic RTM of one shot -> (image, illumination = sum vz_snap^2 * dt*k)."""
    T = np.dtype(dtype)
    _, snaps, ill = elastic_forward(g, model_fwd, src, six, siz, rix, riz, k, save=True,
                                    illum=True, dtype=T)
    ill = ill * T.type(g.dt * k)
    img = np.zeros((g.nz, g.nx), dtype=T)
    elastic_backward(g, model_bwd, data, rix, riz, snaps, img, k)
    return img, ill


def misfit_gradient(g: Grid, vel, src, six, siz, rix, riz, d_obs, dtype=np.float64, res=None):
    """Misfit J = 1/2 |R P - d_obs|^2 of forward() and dJ/dvel by the discrete
    adjoint with every level kept (rtm_oracle.cpp, misfit_gradient).  `res`,
    when given, replaces the residual (adjoint tests).  Returns
    (J, grad (nz,nx), pred (nt,nrec), probe (nt) = dJ/dU^m at the source)."""
    T = np.dtype(dtype)
    src = _c(src, T)
    nt = len(src)
    rix = _c(rix, np.int32)
    vel = _c(vel, T)
    d_obs = _c(d_obs, T)
    res = None if res is None else _c(res, T)
    sp = g.sponge(T)
    grad = np.zeros((g.nz, g.nx), T)
    pred = np.zeros((nt, len(rix)), T)
    probe = np.zeros(nt, T)
    loss = getattr(lib(), f"ora_misfit_gradient_{_sfx(T)}")(
        _ptr(vel), _ptr(sp), g.nx, g.nz, g.dx, g.dz, g.dt, int(g.free_surface), _ptr(src), nt,
        int(six), int(siz), _ptr(rix), len(rix), int(riz), _ptr(d_obs), _ptr(res), _ptr(grad),
        _ptr(pred), _ptr(probe))
    return loss, grad, pred, probe
