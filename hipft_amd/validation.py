"""Built-in validation fields of the reference (validation_run = 1, 2).

validation_run = 1: the "soccer ball" Y_6^0 + sqrt(14/11) Y_6^5, an
eigenfunction of the surface Laplacian (l(l+1) = 42) that decays as
exp(-42 eta t) (src/hipft.f90:1892-1981, 3155-3160).
validation_run = 2: two Gaussian blobs moved by a rigid phi rotation and/or a
constant theta velocity (src/hipft.f90:2034-2101, 3161-3174).
"""
import math

import numpy as np

PI = 3.1415926535897932
PI_TWO = 1.5707963267948966
PI_FOUR = 0.7853981633974483
THREEPI_TWO = 4.71238898038469
TWOPI = 6.2831853071795864
KM_S_TO_RS_HR = 0.005172413793103448


def _ipow(x, n):
    res, sq, have = 1.0, x, False
    while n:
        if n & 1:
            res = res * sq if have else sq
            have = True
        n >>= 1
        if n:
            sq = sq * sq
    return res


def tesseral_harmonic(m, l, pvec, tvec):
    """Real spherical harmonic Y(nt, np) by the first modified forward column
    recursion of Holmes & Featherstone (2002), as in
    tesseral_spherical_harmonic (src/hipft.f90:2390-2533)."""
    pvec = np.asarray(pvec, dtype=np.float64)
    tvec = np.asarray(tvec, dtype=np.float64)
    abs_m = abs(m)
    ct = np.cos(tvec)
    pold = np.ones_like(tvec)
    poldold = np.zeros_like(tvec)
    if abs_m > 0:
        pold[:] = math.sqrt(3.0)
        for im in range(2, abs_m + 1):
            pold = math.sqrt((2.0 * im + 1.0) / (2.0 * im)) * pold
    for il in range(abs_m + 1, l + 1):
        anm = math.sqrt(((2.0 * il - 1.0) * (2.0 * il + 1.0)) / float((il - abs_m) * (il + abs_m)))
        bnm = math.sqrt(((2.0 * il + 1.0) * float(il + abs_m - 1) * float(il - abs_m - 1))
                        / (float((il - abs_m) * (il + abs_m)) * (2.0 * il - 3.0)))
        pl = anm * ct * pold - bnm * poldold
        poldold = pold
        pold = pl
    norm = 1.0 / math.sqrt(4.0 * 3.141592653589793)
    sgn = -1.0 if abs_m % 2 else 1.0
    st = np.sin(tvec)
    stm = np.array([_ipow(float(s), abs_m) for s in st])
    pold = sgn * stm * norm * pold
    pos = 1.0 if m + 1 >= 0 else 0.0
    ang = pos * np.cos(m * pvec) + (1.0 - pos) * np.sin(abs_m * pvec)
    return pold[:, None] * ang[None, :]


def soccer_ball(pvec, tvec):
    """Initial map (nt, np) of validation_run = 1."""
    y0 = tesseral_harmonic(0, 6, pvec, tvec)
    y5 = tesseral_harmonic(5, 6, pvec, tvec)
    return y0 + math.sqrt(14.0 / 11.0) * y5


def uniform_scales(res_nt, res_np):
    """Map scales of a resolution-only run (src/hipft.f90:1793-1813)."""
    d1 = TWOPI / (res_np - 1)
    d2 = PI / (res_nt - 1)
    return (np.array([i * d1 for i in range(res_np)]), np.array([i * d2 for i in range(res_nt)]))


def gaussians(pvec, tvec, width=0.03, t_center=PI_FOUR, p1=PI_TWO, p2=THREEPI_TWO):
    """The two-blob field of validation_run = 2 on a (nt, np) grid."""
    tvec = np.asarray(tvec, dtype=np.float64)
    pvec = np.asarray(pvec, dtype=np.float64)
    n = len(tvec)
    idx = np.arange(n)
    with np.errstate(divide="ignore"):
        st_i = np.where((idx > 0) & (idx < n - 1), 1.0 / np.sin(tvec), 0.0)
    T, P = np.meshgrid(tvec, pvec, indexing="ij")
    return (-st_i[:, None] * np.exp(-(T - t_center) ** 2 / width - (P - p1) ** 2 / width)
            + st_i[:, None] * np.exp(-(T - t_center) ** 2 / width - (P - p2) ** 2 / width))


def gaussians_at(t, p, st_i, width, vt_const, vp_omega, elapsed):
    """Exact validation_run = 2 field at a later time on the run mesh
    (src/hipft.f90:3162-3174); t, st_i on the theta mesh, p the phi nodes."""
    vtt = vt_const * KM_S_TO_RS_HR * elapsed
    vpt = vp_omega * KM_S_TO_RS_HR * elapsed
    p1 = (PI_TWO + vpt) - math.floor((PI_TWO + vpt) / TWOPI) * TWOPI
    p2 = (THREEPI_TWO + vpt) - math.floor((THREEPI_TWO + vpt) / TWOPI) * TWOPI
    T, P = np.meshgrid(t, p, indexing="ij")
    st_i = np.asarray(st_i)
    return (-st_i[:, None] * np.exp(-(T - PI_FOUR - vtt) ** 2 / width - (P - p1) ** 2 / width)
            + st_i[:, None] * np.exp(-(T - PI_FOUR - vtt) ** 2 / width - (P - p2) ** 2 / width))
