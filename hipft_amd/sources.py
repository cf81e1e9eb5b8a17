"""Source terms of a run: a constant source map read from file and random
flux emergence (RFE).

* interp2d restates the reference's bilinear interpolation of a 2-D map onto
  the run's mesh (interp2d and interp, src/hipft.f90:7992-8063, 8178-8221),
  used by load_source for source_from_file (7074-7124).
* GaussianStream is the Marsaglia-Bray polar method of get_gaussian
  (8616-8670), pairs of normals from pairs of uniforms, the second normal of a
  pair returned by the next call.  Its uniforms come from numpy's PCG64
  generator seeded with source_rfe_seed, not from the Fortran compiler's
  RANDOM_NUMBER, so the random maps (and the runs that use them) agree with
  the reference only in distribution.
* RandomFluxEmergence is the frame schedule of load_source / update_source
  (7160-7215, 3875-3947): two frames, the newer generated every
  source_rfe_lifetime hours, blended linearly in time; the frames themselves
  are built on the GPU from the random map (hft_rfe_generate).
"""
import math

import numpy as np


def _intervals(xs, xv):
    """First interval [xs[i], xs[i+1]] holding each xv and its interpolation
    factor (interp, 8178-8221); inside is False outside [xs[0], xs[-1]]."""
    xs = np.asarray(xs, dtype=np.float64)
    xv = np.asarray(xv, dtype=np.float64)
    inside = (xv >= xs[0]) & (xv <= xs[-1])
    if len(xs) == 1:
        i = np.zeros(xv.shape, dtype=np.int64)
        return i, i, np.zeros(xv.shape), inside & (xv == xs[0])
    i = np.clip(np.searchsorted(xs, xv, side="left") - 1, 0, len(xs) - 2)
    x0, x1 = xs[i], xs[i + 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        a = np.where(x0 == x1, 0.0, (xv - x0) / (x1 - x0))
    return i, i + 1, a, inside


def interp2d(xi, yi, fi, x, y):
    """fi[len(xi)][len(yi)] on xi x yi onto x x y: f[len(x)][len(y)], zero
    outside the input mesh; the reference's operation order."""
    xi, yi = (np.asarray(v, dtype=np.float64) for v in (xi, yi))
    if np.any(np.diff(xi) < 0) or np.any(np.diff(yi) < 0):
        raise ValueError("interp2d: scales are not monotonically increasing")
    fi = np.asarray(fi, dtype=np.float64)
    ii, iip1, ax, inx = _intervals(xi, x)
    jj, jjp1, ay, iny = _intervals(yi, y)
    one = 1.0
    AX, AY = ax[:, None], ay[None, :]
    f00 = fi[ii[:, None], jj[None, :]]
    f01 = fi[ii[:, None], jjp1[None, :]]
    f10 = fi[iip1[:, None], jj[None, :]]
    f11 = fi[iip1[:, None], jjp1[None, :]]
    f = (one - AX) * ((one - AY) * f00 + AY * f01) + AX * ((one - AY) * f10 + AY * f11)
    return np.where(inx[:, None] & iny[None, :], f, 0.0)


class GaussianStream:
    """Normal deviates by the polar method (get_gaussian, 8616-8670)."""

    def __init__(self, seed=None):
        self.rng = np.random.Generator(np.random.PCG64(None if seed in (None, -9999) else seed))
        self.pending = np.zeros(0)

    def take(self, n):
        out = [self.pending[:n]]
        have = len(out[0])
        self.pending = self.pending[have:]
        while have < n:
            m = max(1024, int(1.35 * (n - have) / 2) + 16)
            u = 2.0 * self.rng.random(m) - 1.0   # SCALE(u,1) - one
            v = 2.0 * self.rng.random(m) - 1.0
            s = u * u + v * v
            ok = (s < 1.0) & (s > 0.0)
            u, v, s = u[ok], v[ok], s[ok]
            sln = np.sqrt(-(2.0 * np.log(s)) / s)  # SQRT(-SCALE(LOG(s),1)/s)
            pairs = np.empty(2 * len(s))
            pairs[0::2] = u * sln
            pairs[1::2] = v * sln
            need = n - have
            out.append(pairs[:need])
            self.pending = pairs[need:]
            have += min(need, len(pairs))
        return np.concatenate(out)


class RandomFluxEmergence:
    """Frame schedule of random flux emergence for one model (FluxTransport)."""

    NORMFAC = math.pi / math.sqrt(2.0 * math.pi)  # generate_rfe, 8842

    def __init__(self, model, rfe_model, uf_per_hour, lifetime, seed, time):
        self.m = model
        self.rfe_model = int(rfe_model)
        self.uf = np.asarray(uf_per_hour, dtype=np.float64)
        self.lifetime = float(lifetime)
        self.g = GaussianStream(seed)
        self.first = 0           # slot of rfe1; rfe2 is the other
        self._generate(0)
        self._generate(1)
        self.time_next = time + self.lifetime
        self.time_prev = time

    def random_map(self):
        """rnd_nums(1:ntm, 2:npm-1) as [npm-2][ntm] (k outer, j inner, the
        reference's loop order)."""
        m = self.m
        return self.NORMFAC * self.g.take((m.npm - 2) * m.ntm).reshape(m.npm - 2, m.ntm)

    def _generate(self, slot):
        self.m.rfe_generate(slot, self.random_map(), self.rfe_model, self.uf)

    def update(self, time):
        """update_source at the start of a step: a new frame when its time has
        come, then the blend weight of the two frames."""
        if time >= self.time_next:
            self.time_prev = self.time_next
            self.time_next = self.time_prev + self.lifetime
            self.first = 1 - self.first
            self._generate(1 - self.first)
        alpha = (time - self.time_prev) / self.lifetime
        self.m.rfe_weights(self.first, alpha)
        return alpha

    def cut(self, time, dt):
        """update_timestep's RFE cut (4431-4442): (dt, cut or not)."""
        if time + dt >= self.time_next:
            return min(self.time_next - time, 0.5 * self.lifetime), True
        return dt, False
