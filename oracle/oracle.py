"""ctypes front end of the CPU oracle, plus a restatement of HipFT's main loop.

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import
this module.

``Oracle`` wraps ``liboracle.so`` (hft_oracle.c).  ``run_case`` restates the
reference's time loop (src/hipft.f90:903-948: update_step 3338,
update_timestep 4308-4461, flux_transport_step 4463-4507,
apply_data_assimilation 4148, analysis_step 3098, output_histories 2607) with
``verbose=2`` history semantics (extra rows inside diffusion subcycles,
src/hipft.f90:4659-4669), so its histories compare row by row with the
reference's testsuite fixtures.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# Reference constants (src/hipft.f90:83-136).
PI = 3.1415926535897932
PI_TWO = 1.5707963267948966
PI_FOUR = 0.7853981633974483
THREEPI_TWO = 4.71238898038469
TWOPI = 6.2831853071795864
KM_S_TO_RS_HR = 0.005172413793103448
DIFF_KM2_S_TO_RS2_HR = 7.43162901307967e-09

_D = ctypes.POINTER(ctypes.c_double)
_L = ctypes.POINTER(ctypes.c_long)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int]
        for name in ("orc_destroy", "orc_flow_reset", "orc_set_lf_alpha"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_mesh.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.orc_weno.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_alpha.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_coef.argtypes = [ctypes.c_void_p, _D]
        L.orc_pole_fac.argtypes = [ctypes.c_void_p, _D]
        L.orc_set_f.argtypes = [ctypes.c_void_p, _D]
        L.orc_get_f.argtypes = [ctypes.c_void_p, _D]
        L.orc_ingest_pt.argtypes = [ctypes.c_void_p, _D, ctypes.c_double]
        L.orc_flow_add_mf.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.orc_flow_add_dr.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.orc_flow_add_rigid.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_flow_add_vt_const.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_flow_set.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_flow_get.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_flow_attenuate.argtypes = [ctypes.c_void_p, _D]
        L.orc_flow_dtmax.restype = ctypes.c_double
        L.orc_flow_dtmax.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_advection_operator.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_advection_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_load_diffusion.argtypes = [ctypes.c_void_p, _D, ctypes.c_double]
        L.orc_dt_euler.restype = ctypes.c_double
        L.orc_dt_euler.argtypes = [ctypes.c_void_p]
        L.orc_dt_euler_r.restype = ctypes.c_double
        L.orc_dt_euler_r.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_diffusion_operator.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_ptl.restype = ctypes.c_double
        L.orc_ptl.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_sts_rkg2.restype = ctypes.c_long
        L.orc_sts_rkg2.argtypes = [ctypes.c_double, ctypes.c_double, _D, _D, _D, _D, ctypes.c_long]
        L.orc_diffusion_sts.restype = ctypes.c_long
        L.orc_diffusion_sts.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_diffusion_step.restype = ctypes.c_int
        L.orc_diffusion_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, _D, _D, _L, _L, _D, _D]
        L.orc_analysis.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_assimilate.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D, _D, _D,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_get_flux.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.orc_balance_flux.argtypes = [ctypes.c_void_p, _D]
        L.orc_set_weno_eps.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_set_pole_flux_lat_limit.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_rfe_frame.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D, _D]
        L.orc_source_step.argtypes = [ctypes.c_void_p, ctypes.c_double, _D]
        L.orc_tesseral.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int,
                                   ctypes.c_int, _D]
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_D)


def _vec(x, n):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)))
    return a


MESH = ["t", "p", "th", "ph", "dt", "dth", "dp", "dph", "st", "sth", "ct", "cth", "st_i",
        "sth_i", "dt_i", "dth_i", "dp_i", "dph_i"]


class Oracle:
    """One oracle context: mesh, state f, flows and diffusion coefficients.

    Arrays cross the boundary in the reference layout f(ntm,npm,nr) (Fortran
    order), i.e. numpy C-order arrays of shape (nr, npm, ntm).
    """

    def __init__(self, t, p, nr):
        t = np.ascontiguousarray(t, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        self.ntm, self.npm1, self.nr = len(t), len(p), nr
        self.npm = self.np_ = self.npm1 + 1
        self.nt = self.ntm + 1
        self._h = lib().orc_create(self.ntm, self.npm1, _p(t), _p(p), nr)
        self.mesh = {}
        for i, name in enumerate(MESH):
            buf = np.zeros(max(self.nt, self.np_) + 1)
            n = lib().orc_mesh(self._h, i, _p(buf))
            self.mesh[name] = buf[:n].copy()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    @property
    def shape(self):
        return (self.nr, self.npm, self.ntm)

    def get_f(self):
        out = np.empty(self.shape)
        lib().orc_get_f(self._h, _p(out))
        return out

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.shape)
        lib().orc_set_f(self._h, _p(f))

    def ingest_pt(self, map_pt, mult_fac=1.0):
        m = np.ascontiguousarray(map_pt, dtype=np.float64)
        if m.ndim == 2:
            m = np.ascontiguousarray(np.broadcast_to(m, (self.nr,) + m.shape))
        assert m.shape == (self.nr, self.ntm, self.npm1), m.shape
        lib().orc_ingest_pt(self._h, _p(m), mult_fac)

    # flows
    def flow_reset(self):
        lib().orc_flow_reset(self._h)

    def flow_add_mf(self, p1, p3, p5):
        n = self.nr
        lib().orc_flow_add_mf(self._h, _p(_vec(p1, n)), _p(_vec(p3, n)), _p(_vec(p5, n)))

    def flow_add_dr(self, p0, p2, p4):
        n = self.nr
        lib().orc_flow_add_dr(self._h, _p(_vec(p0, n)), _p(_vec(p2, n)), _p(_vec(p4, n)))

    def flow_add_rigid(self, omega_kms):
        lib().orc_flow_add_rigid(self._h, float(omega_kms))

    def flow_add_vt_const(self, vt_kms):
        lib().orc_flow_add_vt_const(self._h, float(vt_kms))

    def flow_set(self, vt, vp):
        vt = np.ascontiguousarray(vt, dtype=np.float64).reshape(self.nr, self.npm, self.nt)
        vp = np.ascontiguousarray(vp, dtype=np.float64).reshape(self.nr, self.np_, self.ntm)
        lib().orc_flow_set(self._h, _p(vt), _p(vp))

    def flow_get(self):
        vt = np.empty((self.nr, self.npm, self.nt))
        vp = np.empty((self.nr, self.np_, self.ntm))
        lib().orc_flow_get(self._h, _p(vt), _p(vp))
        return vt, vp

    def weno_weights(self):
        """WENO3 grid weights: dict theta/phi -> [CP, MC, M_T, CP_T, P_T, MC_T, eps]."""
        out = {"theta": [], "phi": []}
        for w in range(14):
            buf = np.zeros(max(self.nt, self.np_) + 1)
            n = lib().orc_weno(self._h, w, _p(buf))
            out["theta" if w < 7 else "phi"].append(buf[:n].copy())
        return out

    def alphas(self):
        at, ap = np.empty(self.shape), np.empty(self.shape)
        lib().orc_alpha(self._h, _p(at), _p(ap))
        return at, ap

    def diffusion_coef(self):
        """coef(2:ntm-1, 2:npm-1, 5, nr) as a C-order array (nr, 5, npm-2, ntm-2)."""
        c = np.empty((self.nr, 5, self.npm - 2, self.ntm - 2))
        lib().orc_coef(self._h, _p(c))
        return c

    def pole_factors(self):
        out = np.zeros(4)
        lib().orc_pole_fac(self._h, _p(out))
        return dict(flow_n=out[0], flow_s=out[1], diff_n=out[2], diff_s=out[3])

    def flow_attenuate(self, values):
        lib().orc_flow_attenuate(self._h, _p(_vec(values, self.nr)))

    def set_lf_alpha(self):
        lib().orc_set_lf_alpha(self._h)

    def flow_dtmax(self, strang):
        return lib().orc_flow_dtmax(self._h, int(bool(strang)))

    def advection_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.shape)
        y = np.zeros(self.shape)
        lib().orc_advection_operator(self._h, _p(x), _p(y))
        return y

    def advection_step(self, dt):
        lib().orc_advection_step(self._h, float(dt))

    # diffusion
    def load_diffusion(self, eta, factor=DIFF_KM2_S_TO_RS2_HR):
        lib().orc_load_diffusion(self._h, _p(_vec(eta, self.nr)), float(factor))

    @property
    def dt_euler(self):
        return lib().orc_dt_euler(self._h)

    def dt_euler_r(self, r):
        return lib().orc_dt_euler_r(self._h, r)

    def diffusion_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.shape)
        y = np.zeros(self.shape)
        lib().orc_diffusion_operator(self._h, _p(x), _p(y))
        return y

    def ptl(self, dt_in):
        return lib().orc_ptl(self._h, float(dt_in))

    def diffusion_sts(self, dt):
        return lib().orc_diffusion_sts(self._h, float(dt))

    def diffusion_step(self, dt, subcycles=0, record_history=False, fval_pt=None,
                       max_cycles=64):
        ct = np.zeros(max_cycles)
        cdt = np.zeros(max_cycles)
        cs = np.zeros(max_cycles, dtype=np.int64)
        csc = np.zeros(max_cycles, dtype=np.int64)
        hist = np.zeros((max_cycles, self.nr, 14)) if record_history else None
        fv = None if fval_pt is None else np.ascontiguousarray(fval_pt, dtype=np.float64)
        n = lib().orc_diffusion_step(self._h, float(dt), int(subcycles), max_cycles, _p(ct),
                                     _p(cdt), cs.ctypes.data_as(_L), csc.ctypes.data_as(_L),
                                     _p(hist), _p(fv))
        m = min(n, max_cycles)
        return dict(ncycles=n, time=ct[:m], dt=cdt[:m], s=cs[:m], sc=csc[:m],
                    hist=None if hist is None else hist[:m])

    def analysis(self, fval_pt=None):
        out = np.zeros((self.nr, 14))
        fv = None if fval_pt is None else np.ascontiguousarray(fval_pt, dtype=np.float64)
        lib().orc_analysis(self._h, _p(fv), _p(out))
        return out

    def assimilate(self, data, custom_from_mu=True, mu_power=4.0, mu_limit=0.1, lat_limit=0.0,
                   mult_fac=1.0, add_mode=False, balance=False):
        d = np.ascontiguousarray(data, dtype=np.float64)
        assert d.shape == (self.nr, 3, self.ntm, self.npm1), d.shape
        n = self.nr
        lib().orc_assimilate(self._h, _p(d), int(custom_from_mu), _p(_vec(mu_power, n)),
                             _p(_vec(mu_limit, n)), _p(_vec(lat_limit, n)), float(mult_fac),
                             int(add_mode), int(balance))

    def rfe_frame(self, rnd, model, uf_per_hour):
        """Random flux emergence frame (nr, npm, ntm) from rnd [npm-2][ntm]."""
        r = np.ascontiguousarray(rnd, dtype=np.float64)
        assert r.shape == (self.npm - 2, self.ntm), r.shape
        out = np.zeros((self.nr, self.npm, self.ntm))
        lib().orc_rfe_frame(self._h, _p(r), int(model), _p(_vec(uf_per_hour, self.nr)), _p(out))
        return out

    def source_step(self, dt, src):
        """f += dt * src, src (nr, npm, ntm)."""
        a = np.ascontiguousarray(np.broadcast_to(src, (self.nr, self.npm, self.ntm)),
                                 dtype=np.float64)
        lib().orc_source_step(self._h, float(dt), _p(a))

    def set_weno_eps(self, eps):
        """weno_eps (load_weno 6895-6901): <= 0 gives the cell widths."""
        lib().orc_set_weno_eps(self._h, float(eps))

    def set_pole_flux_lat_limit(self, deg):
        lib().orc_set_pole_flux_lat_limit(self._h, float(deg))

    def balance_flux(self, map_tp):
        """balance_flux (8752-8801) of one map (npm, ntm), returned."""
        a = np.array(map_tp, dtype=np.float64, order="C")
        lib().orc_balance_flux(self._h, _p(a))
        return a

    def get_flux(self, map_tp):
        """(positive, negative) flux of one map (npm, ntm), get_flux order."""
        a = np.ascontiguousarray(map_tp, dtype=np.float64)
        fp, fm = np.zeros(1), np.zeros(1)
        lib().orc_get_flux(self._h, _p(a), _p(fp), _p(fm))
        return float(fp[0]), float(fm[0])


def set_threads(n):
    """OpenMP threads of the oracle's loops over points."""
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def sts_rkg2(dt, dt_euler):
    """RKG2 stage count and coefficient vectors (src/hipft.f90:5361-5475)."""
    s = lib().orc_sts_rkg2(dt, dt_euler, None, None, None, None, 0)
    u, v, ub, g = (np.zeros(s) for _ in range(4))
    lib().orc_sts_rkg2(dt, dt_euler, _p(u), _p(v), _p(ub), _p(g), s)
    return s, u, v, ub, g


def tesseral(m, l, pvec, tvec):
    pvec = np.ascontiguousarray(pvec, dtype=np.float64)
    tvec = np.ascontiguousarray(tvec, dtype=np.float64)
    Y = np.zeros((len(tvec), len(pvec)))  # Fortran Y(np,nt) == C [nt][np]
    lib().orc_tesseral(m, l, _p(pvec), _p(tvec), len(pvec), len(tvec), _p(Y))
    return Y


# ---------------------------------------------------------------------------
# Reference main loop restated over the oracle.
# ---------------------------------------------------------------------------

DEFAULTS = dict(
    verbose=2, res_nt=0, res_np=0, n_realizations=1, validation_run=0,
    validation_run_width=0.03, time_start=0.0, time_end=1.0, dt_max=1.7976931348623157e308,
    strang_splitting=False, advance_flow=False, flow_vp_rigid_omega=0.0, flow_vt_const=0.0,
    flow_attenuate=False, flow_attenuate_values=None, flow_attenuate_value=500.0,
    flow_dr_model=0, flow_mf_model=0,
    flow_dr_coef_p0_values=None, flow_dr_coef_p2_values=None, flow_dr_coef_p4_values=None,
    flow_dr_coef_p0_value=46.0, flow_dr_coef_p2_value=-262.0, flow_dr_coef_p4_value=-379.0,
    flow_mf_coef_p1_values=None, flow_mf_coef_p3_values=None, flow_mf_coef_p5_values=None,
    flow_mf_coef_p1_value=22.0, flow_mf_coef_p3_value=11.0, flow_mf_coef_p5_value=-28.0,
    advance_diffusion=False, diffusion_coef_constant=0.0, diffusion_coef_constants=None,
    diffusion_coef_factor=DIFF_KM2_S_TO_RS2_HR, diffusion_subcycles=0,
    output_history_time_cadence=0.0, output_map_time_cadence=0.0, output_map_idx_start=1,
    initial_map=None, initial_scales=None, initial_map_mult_fac=1.0,
)


def _per_real(cfg, name, nr):
    vals = cfg.get(name + "s")
    base = cfg[name]
    out = np.full(nr, float(base))
    if vals is not None:
        for i, v in enumerate(vals[:nr]):
            out[i] = v
    return out


def make_case(cfg):
    """Build an Oracle plus the validation machinery for a namelist-like dict."""
    c = dict(DEFAULTS)
    c.update(cfg)
    nr = c["n_realizations"]
    vfield = None
    if c["initial_map"] is None:
        n1, n2 = c["res_np"], c["res_nt"]
        d1, d2 = TWOPI / (n1 - 1), PI / (n2 - 1)
        s1 = np.array([(i) * d1 for i in range(n1)])
        s2 = np.array([(i) * d2 for i in range(n2)])
        f_local = np.zeros((n2, n1))  # PT stored as C [nt][np]
    else:
        f_local = np.asarray(c["initial_map"], dtype=np.float64)
        s1, s2 = (np.asarray(x, dtype=np.float64) for x in c["initial_scales"])
    if c["validation_run"] == 1:
        y0 = tesseral(0, 6, s1, s2)
        y5 = tesseral(5, 6, s1, s2)
        fval_u0 = y0 + math.sqrt(14.0 / 11.0) * y5
        f_local = fval_u0.copy()
        vfield = ("soccer", fval_u0)
    elif c["validation_run"] == 2:
        if abs(c["flow_vt_const"]) > 0:
            c["flow_vt_const"] = (PI_TWO / (c["time_end"] - c["time_start"])) / KM_S_TO_RS_HR
        if abs(c["flow_vp_rigid_omega"]) > 0:
            c["flow_vp_rigid_omega"] = (TWOPI / (c["time_end"] - c["time_start"])) / KM_S_TO_RS_HR
        w = c["validation_run_width"]
        st_i = np.where((np.arange(len(s2)) > 0) & (np.arange(len(s2)) < len(s2) - 1),
                        1.0 / np.sin(s2), 0.0)
        T, P = np.meshgrid(s2, s1, indexing="ij")
        f_local = (-st_i[:, None] * np.exp(-(T - PI_FOUR) ** 2 / w - (P - PI_TWO) ** 2 / w)
                   + st_i[:, None] * np.exp(-(T - PI_FOUR) ** 2 / w - (P - THREEPI_TWO) ** 2 / w))
        vfield = ("gauss", None)
    o = Oracle(s2, s1, nr)
    o.ingest_pt(f_local, c["initial_map_mult_fac"])
    c["_eta"] = _per_real(c, "diffusion_coef_constant", nr)
    return c, o, vfield


def _fval(c, o, vfield, time):
    if vfield is None:
        return None
    kind, u0 = vfield
    nr = o.nr
    if kind == "soccer":
        fac = np.exp(-42.0 * c["_eta"] * c["diffusion_coef_factor"] * (time - c["time_start"]))
        return np.ascontiguousarray(u0[None, :, :] * fac[:, None, None])
    t, p, st_i = o.mesh["t"], o.mesh["p"][: o.npm1], o.mesh["st_i"]
    w = c["validation_run_width"]
    vtt = c["flow_vt_const"] * KM_S_TO_RS_HR * (time - c["time_start"])
    vpt = c["flow_vp_rigid_omega"] * KM_S_TO_RS_HR * (time - c["time_start"])
    p1 = (PI_TWO + vpt) - math.floor((PI_TWO + vpt) / TWOPI) * TWOPI
    p2 = (THREEPI_TWO + vpt) - math.floor((THREEPI_TWO + vpt) / TWOPI) * TWOPI
    T, P = np.meshgrid(t, p, indexing="ij")
    fv = (-st_i[:, None] * np.exp(-(T - PI_FOUR - vtt) ** 2 / w - (P - p1) ** 2 / w)
          + st_i[:, None] * np.exp(-(T - PI_FOUR - vtt) ** 2 / w - (P - p2) ** 2 / w))
    return np.ascontiguousarray(np.broadcast_to(fv, (nr,) + fv.shape))


def run_case(cfg, max_steps=None):
    """Run a namelist-like case; returns (sol, num) history arrays [nr][rows][cols]."""
    c, o, vfield = make_case(cfg)
    nr = o.nr
    huge = 1.7976931348623157e308
    st = dict(time=c["time_start"], ntime=0, dtime=0.0, dtmax_flow=0.0, dt_adv_used=0.0,
              dt_euler=0.0, sts_s=0, subcycles=c["diffusion_subcycles"], idx_hist=0,
              idx_out=c["output_map_idx_start"] - 1, ts_needs=True, tsflow_needs=True,
              flow_needs=True, changed=False)
    sol = [[] for _ in range(nr)]
    num = [[] for _ in range(nr)]
    if c["advance_diffusion"]:
        o.load_diffusion(c["_eta"], c["diffusion_coef_factor"])
        st["dt_euler"] = o.dt_euler
    p0 = _per_real(c, "flow_dr_coef_p0_value", nr)
    p2 = _per_real(c, "flow_dr_coef_p2_value", nr)
    p4 = _per_real(c, "flow_dr_coef_p4_value", nr)
    p1 = _per_real(c, "flow_mf_coef_p1_value", nr)
    p3 = _per_real(c, "flow_mf_coef_p3_value", nr)
    p5 = _per_real(c, "flow_mf_coef_p5_value", nr)
    att = _per_real(c, "flow_attenuate_value", nr)

    def want_hist(time):
        cad = c["output_history_time_cadence"]
        return (time >= c["time_start"] + st["idx_hist"] * cad) if cad > 0 else True

    def emit(time, hist_rows=None, sts_s=None, subcycles=None):
        if not want_hist(time):
            return
        st["idx_hist"] += 1
        h = hist_rows if hist_rows is not None else o.analysis(_fval(c, o, vfield, time))
        for r in range(nr):
            sol[r].append([st["ntime"], time] + list(h[r]))
            num[r].append([st["ntime"], time, st["dtime"], st["dtmax_flow"], st["dt_adv_used"],
                           st["dt_euler"], st["sts_s"] if sts_s is None else sts_s,
                           st["subcycles"] if subcycles is None else subcycles])

    emit(st["time"])
    # output_map at setup advances idx_out when a map cadence is set (2732-2741).
    if c["output_map_time_cadence"] > 0 and st["time"] >= c["time_start"] + (
            st["idx_out"] - (c["output_map_idx_start"] - 1)) * c["output_map_time_cadence"]:
        st["idx_out"] += 1
    steps = 0
    while True:
        st["ntime"] += 1
        time = st["time"]
        # update_flow (3371-3523)
        if c["advance_flow"] and st["flow_needs"]:
            st["flow_needs"] = False
            st["tsflow_needs"] = True
            st["ts_needs"] = True
            o.flow_reset()
            if c["flow_mf_model"] == 1:
                o.flow_add_mf(p1, p3, p5)
            if c["flow_dr_model"] == 1:
                o.flow_add_dr(p0, p2, p4)
            if c["flow_vp_rigid_omega"] != 0.0:
                o.flow_add_rigid(c["flow_vp_rigid_omega"])
            if c["flow_vt_const"] != 0.0:
                o.flow_add_vt_const(c["flow_vt_const"])
            if c["flow_attenuate"]:
                o.flow_attenuate(att)
                st["flow_needs"] = True
            o.set_lf_alpha()
        # update_timestep (4308-4461)
        if st["ts_needs"]:
            st["dtime"] = c["time_end"] - time
            st["changed"] = True
            if c["advance_flow"]:
                st["dtmax_flow"] = huge
                if st["tsflow_needs"]:
                    st["dtmax_flow"] = o.flow_dtmax(c["strang_splitting"])
                    st["tsflow_needs"] = False
                if st["dtmax_flow"] < st["dtime"]:
                    st["dtime"] = st["dtmax_flow"]
            if c["dt_max"] < st["dtime"]:
                st["dtime"] = c["dt_max"]
            st["ts_needs"] = False
        cad = c["output_map_time_cadence"]
        if cad > 0:
            tnext = c["time_start"] + (st["idx_out"] - (c["output_map_idx_start"] - 1)) * cad
            if time + st["dtime"] >= tnext:
                st["dtime"] = tnext - time
                st["ts_needs"] = True
                st["tsflow_needs"] = True
                st["changed"] = True
        if time + st["dtime"] >= c["time_end"]:
            st["dtime"] = c["time_end"] - time
            st["ts_needs"] = False
            st["changed"] = True
        dt = st["dtime"]
        # flux_transport_step (4463-4507)
        if c["strang_splitting"]:
            if c["advance_flow"]:
                st["dt_adv_used"] = dt * 0.5
                o.advection_step(dt * 0.5)
            if c["advance_diffusion"]:
                _diffuse(c, o, st, vfield, emit, dt, time)
            if c["advance_flow"]:
                o.advection_step(dt * 0.5)
        else:
            if c["advance_flow"]:
                st["dt_adv_used"] = dt
                o.advection_step(dt)
            if c["advance_diffusion"]:
                _diffuse(c, o, st, vfield, emit, dt, time)
        st["time"] = time + dt
        emit(st["time"])
        if c["output_map_time_cadence"] > 0 and st["time"] >= c["time_start"] + (
                st["idx_out"] - (c["output_map_idx_start"] - 1)) * c["output_map_time_cadence"]:
            st["idx_out"] += 1
        steps += 1
        if st["time"] >= c["time_end"] or (max_steps is not None and steps >= max_steps):
            break
    return (np.array(sol), np.array(num), o)


def _diffuse(c, o, st, vfield, emit, dt, time):
    """diffusion_step (src/hipft.f90:4551-4695) with verbose=2 cycle rows."""
    auto_sc = c["diffusion_subcycles"] == 0 if "_sc_input" not in st else st["_sc_input"] == 0
    st.setdefault("_sc_input", c["diffusion_subcycles"])
    sc_in = st["_sc_input"]
    i, done, stepped = 1, False, 0.0
    dt2 = 0.0 if auto_sc else dt / sc_in
    while True:
        if auto_sc:
            dt2 = o.ptl(dt2)
        if auto_sc and stepped + dt2 >= dt:
            dt2 = dt - stepped
            done = True
        if auto_sc and i == 60 - 1:
            dt2 = dt - stepped
            done = True
        if auto_sc and c["verbose"] > 0:
            q = math.ceil((dt - stepped) / dt2)
            st["subcycles"] = int(min(q, 2 ** 31 - 1))
        st["sts_s"] = o.diffusion_sts(dt2)
        stepped = stepped + dt2
        if c["verbose"] > 1:
            emit(time + stepped)
        if not auto_sc and i == sc_in:
            done = True
        if done:
            break
        i += 1
    if auto_sc:
        st["subcycles"] = i
