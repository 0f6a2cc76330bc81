"""Seeded synthetic workloads for the RTM path (host-side input generation).

Models and data follow the documented benchmark configurations: a two-layer
1500/2500 m/s model cut by a 60-degree dipping fault with 50 m throw and a
smoothed +-3% perturbation, a migration model smoothed over 100 m, a Ricker
source and a surface receiver line.  Gaussian smoothing uses
scipy.ndimage.gaussian_filter with sigma in metres / dx; randomness uses
numpy.random.default_rng(seed).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np
from scipy import ndimage, signal

from .wave_solver import ricker_wavelet


def faulted_two_layer(nz, nx, dx, seed=0, x_offset_m=0.0, v_top=1500.0, v_bot=2500.0,
                      depth_m=500.0, fault_x_m=800.0, dip_deg=60.0, throw_m=50.0,
                      perturb=0.03, corr_m=25.0):
    """Two layers cut by a dipping normal fault (hanging wall to the right of
    the fault trace), times a smoothed multiplicative perturbation."""
    z = (np.arange(nz) * dx)[:, None]
    x = (np.arange(nx) * dx + x_offset_m)[None, :]
    # fault plane: x_f(z) = fault_x + z / tan(dip)
    xf = fault_x_m + z / np.tan(np.deg2rad(dip_deg))
    interface = np.where(x > xf, depth_m, depth_m - throw_m)
    v = np.where(z >= interface, v_bot, v_top).astype(np.float64)
    rng = np.random.default_rng(seed)
    noise = ndimage.gaussian_filter(rng.standard_normal((nz, nx)), corr_m / dx)
    noise /= max(noise.std(), 1e-12)
    v *= 1.0 + np.clip(0.01 * noise, -perturb, perturb)
    return v.astype(np.float32)


def smooth(v, sigma_m, dx):
    return ndimage.gaussian_filter(np.asarray(v, np.float64), sigma_m / dx).astype(np.float32)


def bandpass(traces, dt, f_lo=0.8, f_hi=4.5, order=4):
    """Zero-phase Butterworth band-pass along time (axis 0)."""
    sos = signal.butter(order, [f_lo, f_hi], btype="band", fs=1.0 / dt, output="sos")
    return signal.sosfiltfilt(sos, np.asarray(traces, np.float64), axis=0).astype(np.float32)


@dataclass
class Shot:
    """One virtual shot on its own (possibly cropped) grid."""
    nx: int
    nz: int
    nt: int
    vel_fwd: np.ndarray
    vel_bwd: np.ndarray
    src: np.ndarray
    six: int
    siz: int
    data: np.ndarray            # (nt, nrec)
    rix: np.ndarray
    riz: int
    x0: int = 0                 # window offset in the global image / resident model (columns)
    z0: int = 0                 # window offset (rows)

    @property
    def cells(self):
        return self.nx * self.nz


@dataclass
class ElasticShot:
    """One elastic virtual shot: models are (lambda, mu, buoyancy_x, buoyancy_z)."""
    nx: int
    nz: int
    nt: int
    model_fwd: tuple
    model_bwd: tuple
    src: np.ndarray
    six: int
    siz: int
    data: np.ndarray            # (nt, nrec) vertical strain rate
    rix: np.ndarray
    riz: int
    x0: int = 0
    kind: str = "elastic"
    z0: int = 0

    @property
    def cells(self):
        return self.nx * self.nz


@dataclass
class PlainAcousticShot:
    """Plain acoustic RTM, one virtual shot: 200 x 400 at 5 m, dt 0.5 ms, nt 4000."""
    nz: int = 200
    nx: int = 400
    dx: float = 5.0
    dt: float = 5e-4
    nt: int = 4000
    n_sponge: int = 30
    k: int = 2
    f0: float = 2.5
    seed: int = 0
    true_vel: np.ndarray = field(default=None, repr=False)
    mig_vel: np.ndarray = field(default=None, repr=False)
    src: np.ndarray = field(default=None, repr=False)
    rix: np.ndarray = field(default=None, repr=False)
    six: int = 200
    siz: int = 2
    riz: int = 2

    def __post_init__(self):
        self.true_vel = faulted_two_layer(self.nz, self.nx, self.dx, self.seed)
        self.mig_vel = smooth(self.true_vel, 100.0, self.dx)
        self.src, _ = ricker_wavelet(self.f0, self.dt, self.nt * self.dt)
        self.rix = np.arange(200, dtype=np.int32) * 2


def make_plain_acoustic_data(cfg: PlainAcousticShot, solver):
    """Receiver data: forward-modelled in the true model, band-passed 0.8-4.5 Hz."""
    rec, _ = solver.forward_propagate_acoustic(cfg.true_vel, cfg.src, cfg.six, cfg.siz, cfg.rix,
                                               cfg.riz)
    return bandpass(rec, cfg.dt)


def layered_faulted_vp(nz, nx, dx, seed=0, v_min=1800.0, v_max=3500.0, n_layers=5,
                       fault_x_m=None, dip_deg=60.0, throw_m=50.0, perturb=0.03, corr_m=25.0):
    """Faulted layered Vp model (v_min..v_max with depth) with a smoothed
    seeded +-3% perturbation."""
    z = (np.arange(nz) * dx)[:, None]
    x = (np.arange(nx) * dx)[None, :]
    depth = nz * dx
    fault_x_m = 0.5 * nx * dx if fault_x_m is None else fault_x_m
    xf = fault_x_m + z / np.tan(np.deg2rad(dip_deg))
    shift = np.where(x > xf, 0.0, -throw_m)
    layer = np.clip(np.floor((z - shift) / (depth / n_layers)), 0, n_layers - 1)
    v = v_min + (v_max - v_min) * layer / (n_layers - 1)
    rng = np.random.default_rng(seed)
    noise = ndimage.gaussian_filter(rng.standard_normal((nz, nx)), corr_m / dx)
    noise /= max(noise.std(), 1e-12)
    return (v * (1.0 + np.clip(0.01 * noise, -perturb, perturb))).astype(np.float32)


def elastic_from_vp(vp):
    """Vs = Vp/1.8, Gardner density rho = 310 Vp^0.25 (SI), then
    (lambda, mu, buoyancy_x, buoyancy_z)."""
    from .wave_solver import elastic_parameters

    vp = np.asarray(vp, np.float64)
    return elastic_parameters(vp, vp / 1.8, 310.0 * vp ** 0.25)


@dataclass
class PlainElasticShot:
    """Plain 2D elastic P-SV RTM, one shot: 400 x 1600 at 5 m, dt 0.4 ms,
    nt 8000; vertical-force Ricker 2 Hz; 800 vertical strain-rate receivers
    at 10 m spacing."""
    nz: int = 400
    nx: int = 1600
    dx: float = 5.0
    dt: float = 4e-4
    nt: int = 8000
    n_sponge: int = 30
    k: int = 2
    f0: float = 2.0
    n_rec: int = 800
    seed: int = 1
    siz: int = 2
    riz: int = 2

    def __post_init__(self):
        vp = layered_faulted_vp(self.nz, self.nx, self.dx, self.seed)
        self.model_true = elastic_from_vp(vp)
        self.model_mig = elastic_from_vp(smooth(vp, 100.0, self.dx))
        self.src, _ = ricker_wavelet(self.f0, self.dt, self.nt * self.dt)
        self.six = self.nx // 2
        self.rix = (np.arange(self.n_rec) * 2).astype(np.int32)

    @property
    def cell_updates(self):
        return 2 * self.nt * self.nx * self.nz


def correlogram_gather(rng, nt, dt, offsets_m, v_app=1800.0, f0=2.5, f_lo=0.8, f_hi=4.5):
    """Whitened, 1-bit-normalised synthetic cross-correlograms (nt, nrec):
    band-limited (f_lo..f_hi) noise correlated against a Ricker arrival delayed
    by offset / v_app, then sign() and spectral whitening inside the band.
    Built in the frequency domain (noise power spectrum times the arrival's
    spectrum) with multithreaded single-precision FFTs."""
    from scipy import fft as sfft

    nrec = len(offsets_m)
    nf = int(sfft.next_fast_len(2 * nt, real=True))
    freqs = np.fft.rfftfreq(nf, dt).astype(np.float32)
    band = (freqs >= f_lo) & (freqs <= f_hi)
    nb = int(band.sum())
    # |N(f)|^2 of band-limited Gaussian noise: a chi-square(2) draw per bin
    power = np.zeros((len(freqs), nrec), np.float32)
    power[band] = rng.standard_exponential((nb, nrec), dtype=np.float32)
    # Ricker spectrum, delayed by t0 = 1/f0 + |offset|/v_app
    fb = freqs[band][:, None]
    amp = (fb * fb / np.float32(f0 ** 3)) * np.exp(-(fb * fb) / np.float32(f0 * f0))
    delay = (1.0 / f0 + np.abs(np.asarray(offsets_m, np.float32)) / v_app)[None, :]
    spec = np.zeros((len(freqs), nrec), np.complex64)
    spec[band] = amp * power[band] * np.exp(np.complex64(-2j * np.pi) * fb * delay)
    xc = sfft.irfft(spec, n=nf, axis=0, workers=-1)[:nt]
    xc = np.sign(xc).astype(np.float32)
    S = sfft.rfft(xc, n=nf, axis=0, workers=-1)
    mag = np.abs(S)
    S = np.where(band[:, None], S / (mag + np.float32(1e-3) * mag.max()), 0).astype(np.complex64)
    g = sfft.irfft(S, n=nf, axis=0, workers=-1)[:nt]
    g /= max(float(np.abs(g).max()), 1e-30)
    return (np.float32(1e-3) * g).astype(np.float32)


@dataclass
class BatchedWindows:
    """Batched acoustic RTM: 64 virtual shots on cropped windows nz = 300,
    nx ~ U{600..1600}, nt ~ U{3000..6000}, dx 5 m, dt 0.5 ms, k 2."""
    n_shots: int = 64
    nz: int = 300
    nx_range: tuple = (600, 1600)
    nt_range: tuple = (3000, 6000)
    dx: float = 5.0
    dt: float = 5e-4
    k: int = 2
    n_sponge: int = 30
    f0: float = 2.5
    seed: int = 2
    global_nx: int = 0
    shots: List[Shot] = field(default_factory=list, repr=False)

    def build(self, with_data=True, scale=1.0):
        """Creates the shots.  ``scale`` < 1 shrinks nt (for quick tests)."""
        rng = np.random.default_rng(self.seed)
        nxs = rng.integers(self.nx_range[0], self.nx_range[1] + 1, self.n_shots)
        nts = rng.integers(self.nt_range[0], self.nt_range[1] + 1, self.n_shots)
        nts = np.maximum((nts * scale).astype(int), 2)
        self.global_nx = int(nxs.max()) + 2000
        gvel = faulted_two_layer(self.nz, self.global_nx, self.dx, self.seed,
                                 fault_x_m=0.5 * self.global_nx * self.dx)
        gmig = smooth(gvel, 100.0, self.dx)
        self.shots = []
        for i in range(self.n_shots):
            nx, nt = int(nxs[i]), int(nts[i])
            x0 = int(rng.integers(0, self.global_nx - nx + 1))
            six = int(rng.integers(nx // 4, 3 * nx // 4))
            rix = (np.arange(nx // 2) * 2).astype(np.int32)
            src, _ = ricker_wavelet(self.f0, self.dt, nt * self.dt)
            data = None
            if with_data:
                data = correlogram_gather(rng, nt, self.dt, (rix - six) * self.dx, f0=self.f0)
            vw = np.ascontiguousarray(gmig[:, x0:x0 + nx])
            self.shots.append(Shot(nx=nx, nz=self.nz, nt=nt, vel_fwd=vw, vel_bwd=vw, src=src,
                                   six=six, siz=2, data=data, rix=rix, riz=2, x0=x0))
        return self

    def cell_updates(self):
        return sum(2 * s.nt * s.cells for s in self.shots)


@dataclass
class GradientShots:
    """Misfit gradient: 300 x 800 at 5 m, dt 0.5 ms, nt 5000, 16 surface shots
    (Ricker 2.5 Hz) over 400 receivers every 10 m; observed data modelled in
    the two-layer faulted model, gradient taken at its 100 m smoothing."""
    nz: int = 300
    nx: int = 800
    dx: float = 5.0
    dt: float = 5e-4
    nt: int = 5000
    n_shots: int = 16
    n_sponge: int = 30
    f0: float = 2.5
    seed: int = 0
    true_vel: np.ndarray = field(default=None, repr=False)
    start_vel: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        self.true_vel = faulted_two_layer(self.nz, self.nx, self.dx, self.seed)
        self.start_vel = smooth(self.true_vel, 100.0, self.dx)

    def source(self):
        return ricker_wavelet(self.f0, self.dt, self.nt * self.dt)[0]

    def geometry(self):
        """[(six, siz, rix, riz)] of the shots."""
        rix = (np.arange(self.nx // 2) * 2).astype(np.int32)
        sx = np.linspace(self.n_sponge + 10, self.nx - self.n_sponge - 11, self.n_shots)
        return [(int(round(x)), 2, rix, 2) for x in sx]

    def cells(self):
        return self.nx * self.nz


@dataclass
class RequestSpec:
    """A served request before its arrays are built (kind, window, length, seed)."""
    kind: str
    nz: int
    nx: int
    nt: int
    x0: int
    six: int
    seed: int

    @property
    def cells(self):
        return self.nx * self.nz


class ServedStream:
    """Served RTM: a seeded stream of virtual-shot requests, nz in
    {256, 384, 512} x nx in {512, 1024, 2048}, nt ~ U{2000..8000}, 80% acoustic
    and 20% elastic, all cropped from one shared faulted model, each with its
    own seeded correlogram gather (receivers every 2 cells on row 2)."""

    def __init__(self, n_requests=2000, seed=4, dx=5.0, dt=4e-4, k=2, n_sponge=30,
                 elastic_fraction=0.2, f0=2.5, bank_traces=2048):
        self.dx, self.dt, self.k, self.n_sponge, self.f0 = dx, dt, k, n_sponge, f0
        rng = np.random.default_rng(seed)
        self.global_nx = 2048 + 1024
        self.nz_max = 512
        vp = layered_faulted_vp(self.nz_max, self.global_nx, dx, seed)
        self.vel = smooth(vp, 100.0, dx)           # acoustic migration model
        self.ela = elastic_from_vp(self.vel)       # elastic migration model
        self.specs = []
        for i in range(n_requests):
            kind = "elastic" if rng.random() < elastic_fraction else "acoustic"
            nz = int(rng.choice([256, 384, 512]))
            nx = int(rng.choice([512, 1024, 2048]))
            nt = int(rng.integers(2000, 8001))
            x0 = int(rng.integers(0, self.global_nx - nx + 1))
            six = int(rng.integers(nx // 4, 3 * nx // 4))
            self.specs.append(RequestSpec(kind, nz, nx, nt, x0, six, int(rng.integers(1 << 30))))
        # bank of whitened 1-bit correlogram traces, 8000 samples
        brng = np.random.default_rng(seed + 1)
        offs = (brng.random(bank_traces) - 0.5) * 2 * 2048 * dx
        self.bank = correlogram_gather(brng, 8000, dt, offs, f0=f0).T.copy()  # (traces, nt)

    def make(self, spec: RequestSpec, resident: bool = False):
        """The request's arrays.  With ``resident`` the models are left out
        (None): the server migrates in the window of its resident models
        (``RTMServer.set_models(self.vel, self.vel, self.ela, self.ela)``)."""
        rng = np.random.default_rng(spec.seed)
        rix = (np.arange(spec.nx // 2) * 2).astype(np.int32)
        pick = rng.choice(self.bank.shape[0], len(rix), replace=len(rix) > self.bank.shape[0])
        data = np.ascontiguousarray(self.bank[pick, :spec.nt].T)
        src, _ = ricker_wavelet(self.f0, self.dt, spec.nt * self.dt)
        win = (slice(0, spec.nz), slice(spec.x0, spec.x0 + spec.nx))
        if resident:
            if spec.kind == "elastic":
                return ElasticShot(nx=spec.nx, nz=spec.nz, nt=spec.nt, model_fwd=None,
                                   model_bwd=None, src=src, six=spec.six, siz=2, data=data,
                                   rix=rix, riz=2, x0=spec.x0)
            return Shot(nx=spec.nx, nz=spec.nz, nt=spec.nt, vel_fwd=None, vel_bwd=None, src=src,
                        six=spec.six, siz=2, data=data, rix=rix, riz=2, x0=spec.x0)
        if spec.kind == "elastic":
            m = tuple(np.ascontiguousarray(a[win]) for a in self.ela)
            return ElasticShot(nx=spec.nx, nz=spec.nz, nt=spec.nt, model_fwd=m, model_bwd=m,
                               src=src, six=spec.six, siz=2, data=data, rix=rix, riz=2,
                               x0=spec.x0)
        v = np.ascontiguousarray(self.vel[win])
        return Shot(nx=spec.nx, nz=spec.nz, nt=spec.nt, vel_fwd=v, vel_bwd=v, src=src,
                    six=spec.six, siz=2, data=data, rix=rix, riz=2, x0=spec.x0)
