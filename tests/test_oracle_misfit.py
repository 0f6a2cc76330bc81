"""The oracle's misfit gradient (discrete adjoint, every level kept) pinned
by what the mathematics fixes: the dot-product identity <F s, r> = <s, F^T r>
for the source-to-data map, central finite differences of J along smooth
directions and along single cells, and the forward data equal to the oracle's
forward loop.  Small grids: runs in seconds on the CPU."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O

NZ, NX, NT, DX, DT = 32, 48, 200, 5.0, 4e-4


def _setup(fs=True, seed=0):
    rng = np.random.default_rng(seed)
    g = O.Grid(nx=NX, nz=NZ, dx=DX, dz=DX, dt=DT, n_sponge=6, free_surface=fs)
    v0 = 2000.0 + 150.0 * ndimage.gaussian_filter(rng.standard_normal((NZ, NX)), 3.0)
    v_true = v0 * np.where(np.arange(NZ)[:, None] >= 12, 1.15, 1.0)
    src = O.ricker(30.0, DT, NT * DT)[0].astype(np.float64)
    rix = np.arange(0, NX, 2).astype(np.int32)
    riz = 2 if fs else 3
    return g, v0, v_true, src, rix, riz, rng


@pytest.mark.parametrize("fs,six,siz", [(True, 20, 2), (True, 1, 2), (True, 45, 25),
                                        (False, 24, 3), (False, 0, 4), (False, 30, 0)])
def test_dot_product_identity(fs, six, siz):
    g, v0, _, _, rix, riz, rng = _setup(fs, seed=six)
    s = rng.standard_normal(NT)
    zero = np.zeros((NT, len(rix)))
    _, _, pred, _ = O.misfit_gradient(g, v0, s, six, siz, rix, riz, zero)
    r = rng.standard_normal(pred.shape)
    _, _, _, probe = O.misfit_gradient(g, v0, s, six, siz, rix, riz, zero, res=r)
    lhs = float(np.sum(pred * r))
    rhs = DT * DT * float(np.sum(s * probe))
    assert lhs != 0.0
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_forward_data_equals_forward_loop():
    g, v0, _, src, rix, riz, _ = _setup()
    J, _, pred, _ = O.misfit_gradient(g, v0, src, 20, 2, rix, riz, np.zeros((NT, len(rix))))
    rec, _ = O.forward(g, v0, src, 20, 2, rix, riz, dtype=np.float64)
    np.testing.assert_array_equal(pred, rec)
    assert J == pytest.approx(0.5 * np.sum(rec ** 2), rel=1e-13)


@pytest.mark.parametrize("fs", [True, False])
def test_gradient_matches_central_differences(fs):
    g, v0, v_true, src, rix, riz, rng = _setup(fs, seed=3)
    six, siz = 24, riz
    zero = np.zeros((NT, len(rix)))
    _, _, d_obs, _ = O.misfit_gradient(g, v_true, src, six, siz, rix, riz, zero)
    J, grad, _, _ = O.misfit_gradient(g, v0, src, six, siz, rix, riz, d_obs)
    assert J > 0
    def fd_err(dv):
        jp = O.misfit_gradient(g, v0 + dv, src, six, siz, rix, riz, d_obs)[0]
        jm = O.misfit_gradient(g, v0 - dv, src, six, siz, rix, riz, d_obs)[0]
        fd = (jp - jm) / 2.0
        return abs(fd - float(np.sum(grad * dv))) / abs(fd)

    # Taylor test: the central-difference error of an exact gradient falls as
    # eps^2 (a wrong gradient leaves an eps-independent floor)
    for _ in range(5):
        phi = ndimage.gaussian_filter(rng.standard_normal((NZ, NX)), 4.0)
        phi /= np.abs(phi).max()
        e4, e5 = fd_err(1e-4 * v0 * phi), fd_err(1e-5 * v0 * phi)
        assert e5 < 1e-4, (e4, e5)
        assert e5 < 1e-8 or e4 / e5 > 30, (e4, e5)
    # single cells: interior, the largest gradient, next to the surface, sponge
    big = np.unravel_index(np.argmax(np.abs(grad)), grad.shape)
    for iz, ix in ((12, 24), big, (2, 10), (8, 4)):
        dv = np.zeros_like(v0)
        dv[iz, ix] = 1e-5 * v0[iz, ix]
        assert fd_err(dv) < 1e-4, (iz, ix)
    # cells the step never reads through C carry no gradient
    assert np.all(grad[:2, :] == 0) and np.all(grad[:, :2] == 0)
    assert np.all(grad[-2:, :] == 0) and np.all(grad[:, -2:] == 0)


def test_fp32_gradient_close_to_fp64():
    g, v0, v_true, src, rix, riz, _ = _setup(True, seed=4)
    zero = np.zeros((NT, len(rix)))
    _, _, d_obs, _ = O.misfit_gradient(g, v_true, src, 24, 2, rix, riz, zero)
    _, g64, _, _ = O.misfit_gradient(g, v0, src, 24, 2, rix, riz, d_obs)
    _, g32, _, _ = O.misfit_gradient(g, v0.astype(np.float32), src, 24, 2, rix, riz, d_obs,
                                     dtype=np.float32)
    err = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    assert err < 1e-3, err
