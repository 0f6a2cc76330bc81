"""Misfit gradient by the hand-written adjoint (das_amd.fwi).

The forward model is pinned to the existing path (FP32 data bit-identical to
forward_propagate_acoustic) and to the oracle (FP64).  The adjoint has no
counterpart in the reference; it is pinned by what the mathematics fixes:
the dot-product identity <F s, r> = <s, F^T r> in FP64, central finite
differences of J along seeded smooth directions in FP64, linearity over shots,
and the bitwise equality of the checkpointed gradient with the one that keeps
every level."""
import numpy as np
import pytest
import torch
from scipy import ndimage

pytestmark = pytest.mark.gpu

DX, DT, NSP = 5.0, 4e-4, 10
NZ, NX, NT = 48, 80, 200


def _models():
    """Smooth background and a true model with a 15% step at 60 m depth
    (reflections return within the record)."""
    rng = np.random.default_rng(7)
    v0 = 2000.0 + 150.0 * ndimage.gaussian_filter(rng.standard_normal((NZ, NX)), 3.0)
    v_true = v0 * np.where(np.arange(NZ)[:, None] >= 12, 1.15, 1.0)
    return v_true.astype(np.float32), v0.astype(np.float32)


def _shots(v_obs=None, n=1, dtype=np.float64, fs=True, seed=0):
    """Shots on the surface with receivers on every other column (columns 0
    and 1 are never-updated edge cells); d_obs modelled in v_obs (zero data
    when v_obs is None)."""
    from das_amd.fwi import MisfitShot
    from das_amd.wave_solver import ricker_wavelet

    src, _ = ricker_wavelet(30.0, DT, NT * DT)
    rix = np.arange(0, NX, 2).astype(np.int32)
    out = []
    for i in range(n):
        six = 20 + 23 * i
        out.append(MisfitShot(src=src.astype(dtype), six=six, siz=2, rix=rix, riz=2,
                              d_obs=np.zeros((NT, len(rix)), dtype)))
    if v_obs is not None:
        op = _op(out, dtype, fs=fs)
        op.forward_loss(_dev(v_obs, dtype))
        for i, s in enumerate(out):
            s.d_obs = op.predicted(i)
        op.close()
    return out


def _op(shots, dtype=np.float64, ckpt=0, fs=True, nt=None):
    from das_amd.fwi import WaveformMisfit

    tdt = torch.float64 if dtype == np.float64 else torch.float32
    return WaveformMisfit(NX, NZ, DX, DX, DT, shots, n_sponge=NSP, free_surface_top=fs,
                          dtype=tdt, ckpt_interval=ckpt)


def _dev(a, dtype=np.float64):
    return torch.tensor(np.asarray(a, dtype), device="cuda")


def test_fp32_prediction_equals_forward_propagate():
    from das_amd.wave_solver import WaveSolver2D

    v_true, v0 = _models()
    shots = _shots(None, n=2, dtype=np.float32)
    op = _op(shots, np.float32)
    loss = op.forward_loss(_dev(v_true, np.float32))
    solver = WaveSolver2D(NX, NZ, DX, DX, DT, n_sponge=NSP)
    total = 0.0
    for i, s in enumerate(shots):
        rec, _ = solver.forward_propagate_acoustic(v_true, s.src, s.six, s.siz, s.rix, s.riz)
        got = op.predicted(i)
        np.testing.assert_array_equal(got, rec)
        total += 0.5 * float(np.sum(rec.astype(np.float64) ** 2))
    assert loss == pytest.approx(total, rel=1e-12)


@pytest.mark.parametrize("fs", [True, False])
def test_fp64_prediction_matches_oracle(fs):
    from oracle import oracle as O

    v_true, _ = _models()
    shots = _shots(None, n=1, fs=fs)
    op = _op(shots, fs=fs)
    op.forward_loss(_dev(v_true))
    g = O.Grid(nx=NX, nz=NZ, dx=DX, dz=DX, dt=DT, n_sponge=NSP,
               sponge_alpha=float(np.float32(0.0015)), free_surface=fs)
    s = shots[0]
    rec, _ = O.forward(g, v_true.astype(np.float64), s.src, s.six, s.siz, s.rix, s.riz,
                       dtype=np.float64)
    got = op.predicted(0)
    err = np.linalg.norm(got - rec) / np.linalg.norm(rec)
    assert err < 1e-13, err


@pytest.mark.parametrize("fs,six,siz", [(True, 30, 2), (True, 1, 2), (False, 40, 3),
                                        (False, 0, 4), (True, 77, 45)])
def test_dot_product_fp64(fs, six, siz):
    """<F s, r> = <s, F^T r>: F maps the source time function to the data,
    F^T r is read from the adjoint field at the source cell; the adjoint
    field also equals the oracle's."""
    from das_amd.fwi import MisfitShot
    from oracle import oracle as O

    _, v0 = _models()
    rng = np.random.default_rng(11 + six)
    s = rng.standard_normal(NT)
    rix = np.arange(0, NX, 3).astype(np.int32)
    riz = 2 if fs else 3
    shot = MisfitShot(src=s, six=six, siz=siz, rix=rix, riz=riz,
                      d_obs=np.zeros((NT, len(rix))))
    op = _op([shot], fs=fs)
    v = _dev(v0)
    op.forward_loss(v)
    fs_ = op.predicted(0)
    r = rng.standard_normal(fs_.shape)
    op.set_residual(0, r)
    op.gradient(v)
    lam = op.adjoint_at_source(0)
    inj = DT * DT
    lhs = float(np.sum(fs_ * r))
    rhs = inj * float(np.sum(s * lam))
    assert lhs != 0.0
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)
    g = O.Grid(nx=NX, nz=NZ, dx=DX, dz=DX, dt=DT, n_sponge=NSP,
               sponge_alpha=float(np.float32(0.0015)), free_surface=fs)
    _, _, _, probe = O.misfit_gradient(g, v0.astype(np.float64), s, six, siz, rix, riz,
                                       np.zeros_like(r), res=r)
    assert np.linalg.norm(lam - probe) <= 1e-12 * np.linalg.norm(probe)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-4)])
def test_gradient_matches_oracle(dtype, tol):
    from oracle import oracle as O

    v_true, v0 = _models()
    shots = _shots(v_true, n=2, dtype=dtype)
    op = _op(shots, dtype)
    v = _dev(v0, dtype)
    loss = op.forward_loss(v)
    got = op.gradient(v).cpu().numpy()
    g = O.Grid(nx=NX, nz=NZ, dx=DX, dz=DX, dt=DT, n_sponge=NSP,
               sponge_alpha=float(np.float32(0.0015)), free_surface=True)
    want, J = np.zeros((NZ, NX), dtype), 0.0
    for s in shots:
        j, gr, _, _ = O.misfit_gradient(g, v0.astype(dtype), s.src, s.six, s.siz, s.rix, s.riz,
                                        s.d_obs, dtype=dtype)
        want += gr
        J += j
    assert loss == pytest.approx(J, rel=max(tol, 1e-12))
    got, want = got.astype(np.float64), want.astype(np.float64)  # no FP32 underflow in norms
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print("gradient vs oracle (%s): rel L2 %.3e" % (np.dtype(dtype).name, err))
    assert err < tol, err


def _directions(n, seed=0, sigma_m=20.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        phi = ndimage.gaussian_filter(rng.standard_normal((NZ, NX)), sigma_m / DX)
        out.append(phi / np.abs(phi).max())
    return out


def test_gradient_matches_central_fd_fp64():
    """Central differences of J along 20 seeded smooth perturbations
    dv = eps * v * phi (max|phi| = 1) against <g, dv>, eps = 1e-3, 1e-4, 1e-5
    relative: the error of an exact gradient falls as eps^2."""
    v_true, v0 = _models()
    shots = _shots(v_true, n=2)
    op = _op(shots)
    v = _dev(v0)
    op.forward_loss(v)
    g = op.gradient(v).cpu().numpy()
    v064 = v0.astype(np.float64)
    table = []
    for phi in _directions(20):
        row = []
        for eps in (1e-3, 1e-4, 1e-5):
            dv = eps * v064 * phi
            fd = (op.forward_loss(_dev(v064 + dv)) - op.forward_loss(_dev(v064 - dv))) / 2.0
            row.append(abs(fd - float(np.sum(g * dv))) / abs(fd))
        table.append(row)
    t = np.array(table)
    print("FD relative error, eps = 1e-3 / 1e-4 / 1e-5: median %s, max %s"
          % (np.median(t, 0), t.max(0)))
    assert t[:, 2].max() < 1e-4, t
    assert np.all((t[:, 2] < 1e-8) | (t[:, 1] / t[:, 2] > 30)), t


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_checkpointed_gradient_equals_stored(dtype):
    v_true, v0 = _models()
    shots = _shots(v_true, n=2, dtype=dtype)
    grads = []
    for ck in (-1, 0, 7, 1, NT + 5):
        op = _op(shots, dtype, ckpt=ck)
        grads.append(op.gradient(_dev(v0, dtype)).cpu().numpy())
        op.close()
    assert np.abs(grads[0]).max() > 0
    for g in grads[1:]:
        np.testing.assert_array_equal(g, grads[0])


def test_shots_add_up_fp64():
    """Shots of equal length advance in one launch per step, shots of another
    length in a group of their own; the misfit and gradient are the sums of
    the single-shot ones."""
    from das_amd.fwi import MisfitShot

    v_true, v0 = _models()
    shots = _shots(v_true, n=3)
    s1 = shots[1]
    shots[1] = MisfitShot(src=s1.src[:NT - 50], six=s1.six, siz=s1.siz, rix=s1.rix, riz=s1.riz,
                          d_obs=s1.d_obs[:NT - 50])
    v = _dev(v0)
    op = _op(shots)
    loss = op.forward_loss(v)
    both = op.gradient(v).cpu().numpy()
    single, lsum = 0.0, 0.0
    for s in shots:
        o = _op([s])
        lsum += o.forward_loss(v)
        single = single + o.gradient(v).cpu().numpy()
    assert loss == pytest.approx(lsum, rel=1e-13)
    err = np.linalg.norm(both - single) / np.linalg.norm(single)
    assert err < 1e-13, err


def test_fp32_gradient_close_to_fp64():
    v_true, v0 = _models()
    g64 = _op(_shots(v_true, n=2)).gradient(_dev(v0)).cpu().numpy()
    g32 = _op(_shots(v_true, n=2, dtype=np.float32), np.float32).gradient(
        _dev(v0, np.float32)).cpu().numpy()
    err = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print("FP32 gradient vs FP64: rel L2 %.3e" % err)
    assert err < 1e-3, err


def test_autograd_function():
    v_true, v0 = _models()
    op = _op(_shots(v_true, n=1))
    v = _dev(v0).requires_grad_(True)
    loss = op(v)
    (3.0 * loss).backward()
    g = op.gradient(v.detach())
    torch.testing.assert_close(v.grad, 3.0 * g, rtol=0, atol=0)
    # a forward at another velocity between forward and backward: the
    # backward recomputes at the saved velocity
    w = _dev(v0 * 1.01).requires_grad_(True)
    lw = op(w)
    op.forward_loss(_dev(v0))
    lw.backward()
    torch.testing.assert_close(w.grad, op.gradient(w.detach()), rtol=0, atol=0)
    with pytest.raises(ValueError):
        op(_dev(v0, np.float32))


@pytest.mark.parametrize("config", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_blocked_forward_equals_per_step(config):
    """Temporally blocked FP32 forward and recompute: data, misfit and
    gradient bit-identical to one launch per step, with checkpoints and with
    every level kept, for shots of two lengths."""
    from das_amd.fwi import MisfitShot

    v_true, v0 = _models()
    shots = _shots(v_true, n=3, dtype=np.float32)
    s1 = shots[1]
    shots[1] = MisfitShot(src=s1.src[:NT - 37], six=s1.six, siz=s1.siz, rix=s1.rix, riz=s1.riz,
                          d_obs=s1.d_obs[:NT - 37])
    v = _dev(v0, np.float32)
    for ck in (0, -1, 7):
        ref = _op(shots, np.float32, ckpt=ck)
        ref.set_blocking(0)
        lr = ref.forward_loss(v)
        gr = ref.gradient(v).cpu().numpy()
        op = _op(shots, np.float32, ckpt=ck)
        op.set_blocking(config)
        lb = op.forward_loss(v)
        for i in range(3):
            np.testing.assert_array_equal(op.predicted(i), ref.predicted(i))
        assert lb == lr
        np.testing.assert_array_equal(op.gradient(v).cpu().numpy(), gr)
        op.close()
        ref.close()
    with pytest.raises(ValueError):
        _op(shots[:1], np.float64).set_blocking(1)


def test_full_size_checkpointed_equals_stored():
    """The 300 x 800, nt 5000 gradient workload (4 of its 16 shots): sqrt(nt)
    checkpoints with recompute give the gradient of keeping every level, bit
    for bit, in FP32 with the blocked forward."""
    from das_amd.fwi import MisfitShot, WaveformMisfit
    from das_amd.workloads import GradientShots

    cfg = GradientShots()
    src = cfg.source()
    geo = cfg.geometry()[::4]
    shots = [MisfitShot(src=src, six=a, siz=b, rix=r, riz=z,
                        d_obs=np.zeros((cfg.nt, len(r)), np.float32)) for a, b, r, z in geo]
    op = WaveformMisfit(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, shots, n_sponge=cfg.n_sponge)
    op.forward_loss(_dev(cfg.true_vel, np.float32))
    obs = [op.predicted(i) for i in range(len(shots))]
    op.close()
    for s, d in zip(shots, obs):
        s.d_obs = d
    v = _dev(cfg.start_vel, np.float32)
    grads = []
    for ck in (0, -1):
        op = WaveformMisfit(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, shots,
                            n_sponge=cfg.n_sponge, ckpt_interval=ck)
        grads.append(op.gradient(v).cpu().numpy())
        op.close()
    assert np.abs(grads[0]).max() > 0
    np.testing.assert_array_equal(grads[0], grads[1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_two_step_adjoint_equals_single_steps(dtype):
    """Two adjoint steps per launch give the single-step gradient and adjoint
    field bit for bit, with segments of odd and even length."""
    v_true, v0 = _models()
    shots = _shots(v_true, n=2, dtype=dtype)
    v = _dev(v0, dtype)
    for ck in (7, 0, -1):
        out = []
        for steps in (1, 2):
            op = _op(shots, dtype, ckpt=ck)
            if dtype == np.float32:
                op.set_adjoint_blocking(0)  # the per-step kernels
            op.set_adjoint_steps(steps)
            g = op.gradient(v).cpu().numpy()
            out.append((g, op.adjoint_at_source(0), op.adjoint_at_source(1)))
            op.close()
        for a, b in zip(*out):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("fs", [True, False])
@pytest.mark.parametrize("config", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_blocked_adjoint_equals_per_step(config, fs):
    """Temporally blocked FP32 adjoint: gradient and adjoint at the sources
    bit-identical to one launch per step, with checkpoints of several lengths
    and with every level kept, for shots of two lengths whose sources and
    receivers sit on the free-surface rows and in never-updated edge cells."""
    from das_amd.fwi import MisfitShot
    from das_amd.wave_solver import ricker_wavelet

    v_true, v0 = _models()
    src, _ = ricker_wavelet(30.0, DT, NT * DT)
    src = src.astype(np.float32)
    geo = [(20, 2, np.arange(0, NX, 2), 2), (1, 1, np.arange(NX), 0),
           (NX - 2, NZ - 3, np.arange(3, NX - 1, 5), 1), (40, 5, np.array([7, 7, 30]), 3)]
    shots = [MisfitShot(src=src if i != 1 else src[:NT - 29], six=a, siz=b,
                        rix=r.astype(np.int32), riz=z,
                        d_obs=np.zeros((NT if i != 1 else NT - 29, len(r)), np.float32))
             for i, (a, b, r, z) in enumerate(geo)]
    op = _op(shots, np.float32, fs=fs)
    op.forward_loss(_dev(v_true, np.float32))
    for i, sh in enumerate(shots):
        sh.d_obs = op.predicted(i)
    op.close()
    v = _dev(v0, np.float32)
    for ck in (0, -1, 7):
        out = []
        for blk in (0, config):
            op = _op(shots, np.float32, ckpt=ck, fs=fs)
            op.set_adjoint_blocking(blk)
            g = op.gradient(v).cpu().numpy()
            out.append([g] + [op.adjoint_at_source(i) for i in range(len(shots))])
            op.close()
        assert np.abs(out[0][0]).max() > 0
        for a, b in zip(*out):
            np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError):
        _op(shots[:1], np.float64).set_adjoint_blocking(1)


@pytest.mark.parametrize("groups", [2, 3, 8])
def test_launch_groups_equal_one_group(monkeypatch, groups):
    """The blocked passes split a group's shots into launch groups on their
    own streams (TSW_FWI_GROUPS): loss, predicted data, gradient and adjoint
    at the sources equal one group's bit for bit, with checkpoints and with
    every level kept, for five shots of one length (sets of uneven size) and
    one of another."""
    from das_amd.fwi import MisfitShot
    from das_amd.wave_solver import ricker_wavelet

    v_true, v0 = _models()
    src, _ = ricker_wavelet(30.0, DT, NT * DT)
    src = src.astype(np.float32)
    geo = [(20, 2, np.arange(0, NX, 2), 2), (1, 1, np.arange(NX), 0), (NX - 2, NZ - 3, np.arange(3, NX - 1, 5), 1),
           (40, 5, np.array([7, 7, 30]), 3), (33, 0, np.arange(5, NX, 7), 2), (12, 4, np.arange(NX - 9, NX), 1)]
    shots = [MisfitShot(src=src if i != 5 else src[:NT - 17], six=a, siz=b, rix=r.astype(np.int32), riz=z,
                        d_obs=np.zeros((NT if i != 5 else NT - 17, len(r)), np.float32))
             for i, (a, b, r, z) in enumerate(geo)]
    op = _op(shots, np.float32)
    op.forward_loss(_dev(v_true, np.float32))
    for i, sh in enumerate(shots):
        sh.d_obs = op.predicted(i)
    op.close()
    v = _dev(v0, np.float32)
    for ck in (0, 7):
        out = []
        for g in (1, groups):
            monkeypatch.setenv("TSW_FWI_GROUPS", str(g))
            op = _op(shots, np.float32, ckpt=ck)
            loss = op.forward_loss(v)
            grad = op.gradient(v).cpu().numpy()
            out.append([np.float64(loss), grad] + [op.predicted(i) for i in range(len(shots))]
                       + [op.adjoint_at_source(i) for i in range(len(shots))])
            op.close()
        assert np.abs(out[0][1]).max() > 0
        for a, b in zip(*out):
            np.testing.assert_array_equal(a, b)
