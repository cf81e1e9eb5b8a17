"""Synthetic inputs for the benchmark and test workloads.

The reference reads initial maps and meshes from HDF5 files only
(src/hipft.f90:1820-1890, 2130-2132); it has no bipole generator and no grid
generator.  The workloads of this project are defined here:

* ``clustered_theta``: theta nodes on [0, pi] whose spacing near the poles is
  1/ratio of the equatorial spacing, blended by a tanh profile in |latitude|.
* ``bipole_map``: seeded bipolar active regions (two Gaussian polarities of
  great-circle width sigma, Joy's-law tilt, Hale polarity by hemisphere).
* ``meridional_profile``: v_lat = v0 sin(2 lat) cos(lat) as a theta profile of
  vt on the theta half mesh (code units, R_sun/hr).
* differential rotation Omega = 0.18 - 2.396 cos^2 - 1.787 cos^4 deg/day as the
  reference's analytic DR coefficients (m/s): 1 deg/day at 1 R_sun =
  6.96e10 cm * (pi/180) / 86400 s = 140.596 m/s.
"""
import math

import numpy as np

PI = 3.1415926535897932
TWOPI = 6.2831853071795864
M_S_TO_RS_HR = 5.172413793103448e-06  # src/hipft.f90:128
DEG_PER_DAY_TO_M_S = 6.96e8 * (PI / 180.0) / 86400.0
OMEGA_DEG_DAY = (0.18, -2.396, -1.787)


def dr_coefficients(scale=1.0):
    """(p0, p2, p4) in m/s of Omega(theta) * scale."""
    return tuple(scale * c * DEG_PER_DAY_TO_M_S for c in OMEGA_DEG_DAY)


def uniform_phi(npm1):
    return np.array([i * (TWOPI / (npm1 - 1)) for i in range(npm1)])


def uniform_theta(ntm):
    return np.array([i * (PI / (ntm - 1)) for i in range(ntm)])


def clustered_theta(ntm, ratio=2.0, lat_c=45.0, width=10.0):
    """Theta nodes with spacing h(lat) = 1 - (1 - 1/ratio) * P(|lat|),
    P = (1 + tanh((|lat| - lat_c)/width))/2, nodes equidistant in the
    stretched coordinate s = int dtheta / h."""
    n = 20001
    th = np.linspace(0.0, PI, n)
    lat = np.abs(90.0 - np.degrees(th))
    P = 0.5 * (1.0 + np.tanh((lat - lat_c) / width))
    dens = 1.0 / (1.0 - (1.0 - 1.0 / ratio) * P)
    s = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(th))])
    s /= s[-1]
    t = np.interp(np.linspace(0.0, 1.0, ntm), s, th)
    t[0], t[-1] = 0.0, PI
    return t


def bipole_map(t, p, nbip, seed, sigma_deg=(1.0, 4.0), bmax=(20.0, 500.0), lat_max=40.0,
               sep_deg=4.0, joy_deg=5.0, log_sigma=True, fixed_sigma=None, fixed_b=None):
    """PT map (ntm, npm1) [nt][np] in Gauss: sum of nbip seeded bipoles."""
    rng = np.random.default_rng(seed)
    t = np.asarray(t, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    out = np.zeros((len(t), len(p)))
    ct_all, st_all = np.cos(t), np.sin(t)
    for _ in range(nbip):
        lat = rng.uniform(-lat_max, lat_max)
        lon = rng.uniform(0.0, 360.0)
        if fixed_sigma is not None:
            sig = fixed_sigma
        elif log_sigma:
            sig = math.exp(rng.uniform(math.log(sigma_deg[0]), math.log(sigma_deg[1])))
        else:
            sig = rng.uniform(*sigma_deg)
        b = fixed_b if fixed_b is not None else rng.uniform(*bmax)
        tilt = math.radians(joy_deg * math.sin(math.radians(lat)))
        hemi = 1.0 if lat >= 0 else -1.0
        half = 0.5 * sep_deg
        # leading (westward, +lon) polarity nearer the equator by the tilt
        for sgn, dlon, dlat in ((hemi, half * math.cos(tilt), -hemi * half * math.sin(tilt)),
                                (-hemi, -half * math.cos(tilt), hemi * half * math.sin(tilt))):
            th0 = math.radians(90.0 - (lat + dlat))
            ph0 = math.radians(lon + dlon) % TWOPI
            # evaluate within 8 sigma of the centre (exp(-32) ~ 1e-14 relative)
            rad = 8.0 * math.radians(sig)
            jt = np.nonzero(np.abs(t - th0) <= rad)[0]
            if rad >= th0 or th0 + rad >= PI:
                kp = np.arange(len(p))  # the window contains a pole
            else:
                # generous phi half-width: twice the cap radius over the
                # smallest sin(theta) inside the cap
                smin = min(math.sin(th0 - rad), math.sin(th0 + rad))
                dphi = 2.0 * rad / smin
                if dphi >= PI:
                    kp = np.arange(len(p))
                else:
                    dd = np.abs(((p - ph0) + PI) % TWOPI - PI)
                    kp = np.nonzero(dd <= dphi)[0]
            if len(jt) == 0 or len(kp) == 0:
                continue
            ct, st = ct_all[jt][:, None], st_all[jt][:, None]
            cosd = ct * math.cos(th0) + st * math.sin(th0) * np.cos(p[kp][None, :] - ph0)
            d = np.arccos(np.clip(cosd, -1.0, 1.0))
            out[np.ix_(jt, kp)] += sgn * b * np.exp(-0.5 * (d / math.radians(sig)) ** 2)
    return out


def half_mesh_theta(t):
    """th on the theta half mesh, nt = ntm+1 (set_mesh, src/hipft.f90:6590-6596)."""
    ntm = len(t)
    th = np.zeros(ntm + 1)
    th[1:ntm] = 0.5 * (t[1:] + t[:-1])
    dth2 = t[1] - t[0]
    th[0] = th[1] - dth2
    th[ntm] = th[ntm - 1] + (t[ntm - 1] - t[ntm - 2])
    return th


def meridional_profile(t, v0):
    """vt on the theta half mesh (nt) in R_sun/hr for v_lat = v0 sin(2lat) cos(lat)
    (poleward in both hemispheres; vt = -v_lat)."""
    th = half_mesh_theta(t)
    sth = np.sin(np.abs(th))
    return -M_S_TO_RS_HR * v0 * 2.0 * np.cos(th) * sth * sth


def mu_map(t, p, lon_center_deg):
    """mu = cos(angle from disk centre) for an observer in the equatorial
    plane at longitude lon_center: mu = sin(theta) cos(phi - phi_c)."""
    T, P = np.meshgrid(t, p, indexing="ij")
    return np.sin(T) * np.cos(P - math.radians(lon_center_deg))
