"""hft_analysis_noalloc: the history diagnostics with partial sums in the
library's fixed device array (no allocation), called through the C ABI and
checked bit for bit against hft_analysis on two calls in a row, the second
on a smaller batch; oversize batches are refused."""
import ctypes

import numpy as np
import pytest
import torch

from hipft_amd import workloads as W
from hipft_amd.core import FluxTransport, HftError, lib

pytestmark = pytest.mark.gpu


def _model(ntm, npm1, nr, seed):
    t, p = W.clustered_theta(ntm, 2.0), W.uniform_phi(npm1)
    g = FluxTransport(t, p, nr, mode="independent")
    g.set_state(np.stack([W.bipole_map(t, p, 12, seed + r) for r in range(nr)]), layout="pt")
    g.set_diffusion(np.linspace(200.0, 800.0, nr))
    g.flux_transport_step(1.0, strang=True, subcycles=0)
    return g


def _noalloc(g, out, f_ptr=None, ld=0):
    out.fill_(float("nan"))
    torch.cuda.synchronize()
    rc = lib().hft_analysis_noalloc(g._h, f_ptr, ctypes.c_int64(ld), ctypes.c_void_p(out.data_ptr()))
    assert rc == 0, lib().hft_last_error().decode()
    g.synchronize()
    return out.cpu().numpy()


def test_noalloc_analysis_matches_hft_analysis(gpu_available):
    big = _model(91, 181, 4, 10)
    small = _model(33, 65, 2, 20)
    out_big = torch.empty((big.nr, 14), dtype=torch.float64, device="cuda")
    out_small = torch.empty((small.nr, 14), dtype=torch.float64, device="cuda")
    ld = small.npm - 2 + 5
    buf = torch.zeros((small.nr, small.ntm, ld), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    # first call: the larger batch, the context's own state
    h_big = _noalloc(big, out_big)
    # second call in a row: a smaller batch, from a caller's device array with
    # padded rows; the scratch array still holds the first call's partials
    small.get_state_device(buf.data_ptr(), ld)
    small.synchronize()
    h_small = _noalloc(small, out_small, ctypes.c_void_p(buf.data_ptr()), ld)
    # and the first batch again
    h_big2 = _noalloc(big, out_big)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert np.array_equal(h_big, big.analysis())
    assert np.array_equal(h_small, small.analysis())
    assert np.array_equal(h_big2, h_big)
    assert free1 == free0  # nothing allocated on the device
    big.close()
    small.close()


def test_noalloc_analysis_refuses_oversize_batches(gpu_available):
    # 512 rows -> 64 row blocks x 16 partials; 1025 realizations exceed 2^20
    t, p = W.clustered_theta(512, 2.0), W.uniform_phi(8)
    g = FluxTransport(t, p, 1025, mode="independent")
    out = torch.zeros((1025, 14), dtype=torch.float64, device="cuda")
    with pytest.raises(HftError, match="exceed"):
        g.analysis_noalloc(ctypes.c_void_p(out.data_ptr()))
    g.close()
