"""Source terms (hipft_amd/sources.py, oracle orc_rfe_frame): interpolation
of a source map onto the mesh (interp2d, src/hipft.f90:7992-8063), the
polar-method normal generator (get_gaussian, 8616-8670) and random flux
emergence frames (generate_rfe 8809-8921 with balance_flux 8752-8801).

The random stream is not the Fortran compiler's RANDOM_NUMBER, so the RNG is
checked in distribution only; a frame from a given random map is checked
against closed forms (bilinear exactness, flux balance, pole averages, the
unsigned-flux budget the model scales imply)."""
import math

import numpy as np
import pytest

from hipft_amd import workloads as W
from hipft_amd.sources import GaussianStream, interp2d
from oracle.oracle import Oracle

RSUN_CM2 = 4.84416e21  # src/hipft.f90:119


def test_interp2d_identity_and_bilinear():
    xi = np.linspace(0.0, math.pi, 23)
    yi = np.sort(np.random.default_rng(1).uniform(0.0, 2 * math.pi, 31))
    fi = np.random.default_rng(2).standard_normal((23, 31))
    # same mesh: the data itself (a = 0 or 1 at the nodes)
    assert np.array_equal(interp2d(xi, yi, fi, xi, yi), fi)
    # a bilinear function is reproduced inside the mesh, zero outside
    a, b, c, d = 0.3, -1.2, 0.7, 0.25
    g = a + b * xi[:, None] + c * yi[None, :] + d * xi[:, None] * yi[None, :]
    x = np.linspace(-0.1, math.pi + 0.1, 40)
    y = np.linspace(yi[0], yi[-1], 50)
    f = interp2d(xi, yi, g, x, y)
    want = a + b * x[:, None] + c * y[None, :] + d * x[:, None] * y[None, :]
    inside = (x >= 0.0) & (x <= math.pi)
    assert np.allclose(f[inside], want[inside], rtol=0, atol=1e-12)
    assert np.all(f[~inside] == 0.0)
    with pytest.raises(ValueError):
        interp2d(xi[::-1], yi, fi, x, y)


def test_gaussian_stream():
    g = GaussianStream(12345)
    x = g.take(200000)
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1.0) < 0.01
    assert abs(np.mean(np.abs(x)) - math.sqrt(2 / math.pi)) < 0.01
    # the stream continues across calls (second normal of a pair kept)
    a = GaussianStream(7).take(8)
    g2 = GaussianStream(7)
    b = np.concatenate([g2.take(3), g2.take(5)])
    assert np.array_equal(a, b)
    assert not np.array_equal(a, GaussianStream(8).take(8))


@pytest.mark.parametrize("model", [1, 2])
def test_rfe_frame_balance_poles_budget(model):
    t, p = W.uniform_theta(61), W.uniform_phi(121)
    o = Oracle(t, p, 2)
    rng = np.random.default_rng(3)
    normfac = math.pi / math.sqrt(2 * math.pi)
    rnd = normfac * rng.standard_normal((o.npm - 2, o.ntm))
    uf = [150.0, 40.0]
    fr = o.rfe_frame(rnd, model, uf)
    for r in range(2):
        m = fr[r]  # (npm, ntm)
        fp, fm = o.get_flux(m)
        assert abs(fp + fm) <= 1e-13 * fp  # balance_flux
        # pole rows are single values, seam columns periodic
        assert np.ptp(m[:, 0]) == 0.0 and np.ptp(m[:, -1]) == 0.0
        assert np.array_equal(m[0], m[-2]) and np.array_equal(m[-1], m[1])
        # E|rnd| = 1, so the unsigned flux per hour is the requested budget
        # (uf * 1e21 Mx / rsun_cm2 in code units) up to sampling noise
        want = uf[r] * 1e21 / RSUN_CM2
        assert abs((fp - fm) / want - 1.0) < 0.05, ((fp - fm) / want, model)
    # the two realizations differ only by their budgets (same random map)
    assert np.allclose(fr[1] * uf[0], fr[0] * uf[1], rtol=1e-12, atol=0)
