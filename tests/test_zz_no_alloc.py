"""The allocation-free correlation entry point (tsw_xcorr_device_noalloc),
called through the C ABI: equal bit for bit to tsw_xcorr_device on two calls
in a row, the second smaller, with and without accumulation; the exact FP64
lag sum on sampled pairs; device free memory unchanged across a call; sizes
beyond its fixed scratch refused with an error."""
import ctypes

import numpy as np
import pytest

from das_amd import _native as N

pytestmark = pytest.mark.gpu


def _call(fn, data, src, rec, nlag, out, acc):
    import torch

    st = torch.cuda.current_stream().cuda_stream
    N.check(fn(ctypes.c_void_p(data.data_ptr()), data.shape[0], data.shape[1], N.ptr(src),
               len(src), N.ptr(rec), len(rec), nlag, acc, ctypes.c_void_p(out.data_ptr()),
               ctypes.c_void_p(st)))


def _free_bytes():
    f, t = N.LL(), N.LL()
    N.check(N.lib().tsw_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value


def test_noalloc_equals_allocating_entry_point():
    import torch

    L = N.lib()
    N.check(L.tsw_init(0))
    g = torch.Generator().manual_seed(5)
    # (nt, nch, nsrc, nrec, nlag): the second call is smaller than the first
    for nt, nch, nsrc, nrec, nlag in ((4000, 96, 12, 40, 150), (1500, 40, 5, 17, 60)):
        print(f"case nt={nt} nch={nch} nsrc={nsrc} nrec={nrec} nlag={nlag}", flush=True)
        need = L.tsw_xcorr_noalloc_scratch_bytes(nt, nsrc, nrec, nlag)
        assert 0 < need <= L.tsw_xcorr_noalloc_scratch_max()
        data = torch.randn(nt, nch, generator=g).cuda()
        src = ((np.arange(nsrc) * 7) % nch).astype(np.int32)
        rec = ((np.arange(nrec) * 2 + 1) % nch).astype(np.int32)
        nl = 2 * nlag + 1
        for acc in (0, 1):
            base = torch.randn(nsrc, nl, nrec, generator=g).cuda()
            a, b = base.clone(), base.clone()
            torch.cuda.synchronize()
            _call(L.tsw_xcorr_device, data, src, rec, nlag, a, acc)
            f0 = _free_bytes()
            _call(L.tsw_xcorr_device_noalloc, data, src, rec, nlag, b, acc)
            assert _free_bytes() == f0
            torch.cuda.synchronize()
            assert torch.equal(a, b), f"acc={acc}"
            if acc == 0:
                # exact lag sums of sampled pairs (FP64), rounded once to FP32
                x = data.cpu().numpy().astype(np.float64)
                got = b.cpu().numpy()
                for s_i, r_i in ((0, 0), (nsrc - 1, nrec - 1), (nsrc // 2, 3)):
                    sv, rv = x[:, src[s_i]], x[:, rec[r_i]]
                    full = np.correlate(rv, sv, mode="full")  # lag -(nt-1)..nt-1
                    ref = full[nt - 1 - nlag:nt + nlag].astype(np.float32)
                    np.testing.assert_allclose(got[s_i, :, r_i], ref, rtol=2e-6, atol=1e-6)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    import torch

    L = N.lib()
    N.check(L.tsw_init(0))
    nt, nch, nsrc, nrec, nlag = 30000, 1600, 64, 800, 175
    need = L.tsw_xcorr_noalloc_scratch_bytes(nt, nsrc, nrec, nlag)
    assert need > L.tsw_xcorr_noalloc_scratch_max()
    data = torch.zeros(8, dtype=torch.float32, device="cuda")  # never read: refused first
    out = torch.zeros(8, dtype=torch.float32, device="cuda")
    src = np.arange(nsrc, dtype=np.int32)
    rec = np.arange(nrec, dtype=np.int32) + 100
    with pytest.raises(ValueError, match="scratch"):
        _call(L.tsw_xcorr_device_noalloc, _Shape(data, nt, nch), src, rec, nlag, out, 0)


class _Shape:
    """A device pointer presented with a (nt, nch) shape, for a call that is
    refused before the data are read."""

    def __init__(self, t, nt, nch):
        self._t, self.shape = t, (nt, nch)

    def data_ptr(self):
        return self._t.data_ptr()
