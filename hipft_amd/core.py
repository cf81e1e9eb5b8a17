"""ctypes binding of libhft (include/hft.h): the MI355X flux-transport core.

``FluxTransport`` holds one library context: mesh, the batch of realizations
and their state, flows and diffusion.  Its methods are the reference's
hot-path subroutines (src/hipft.f90): ``advection_step`` (4509),
``diffusion_step`` (4551), ``flux_transport_step`` (4463), ``analysis``
(analysis_step 3098), ``assimilate`` (apply_data_assimilation 4148) and the
set-up routines they depend on.  Every call runs on the GPU through libhft;
there is no CPU fallback: a missing or unloadable library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhft.so")

DIFF_KM2_S_TO_RS2_HR = 7.43162901307967e-09  # src/hipft.f90:124
NHIST = 14
HIST_COLUMNS = ["BR_MIN", "BR_MAX", "BR_ABS_MIN", "FLUX_POSITIVE", "FLUX_NEGATIVE",
                "NPOLE_FLUX_POSITIVE", "NPOLE_FLUX_NEGATIVE", "SPOLE_FLUX_POSITIVE",
                "SPOLE_FLUX_NEGATIVE", "NPOLE_AREA", "SPOLE_AREA", "EQ_DIPOLE", "AX_DIPOLE",
                "VALIDATION_ERR_HHabs"]

LAYOUT_TP, LAYOUT_PT = 0, 1
MODE_COUPLED, MODE_INDEPENDENT = 0, 1

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)


class HftError(RuntimeError):
    pass


class Cycle(ctypes.Structure):
    _fields_ = [("time_stepped", ctypes.c_double), ("dt", ctypes.c_double),
                ("sts_s", ctypes.c_int64), ("subcycles", ctypes.c_int64)]


CYCLE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Cycle))

# (name, restype, argtypes): every symbol include/hft.h declares.
_SIGS = [
    ("hft_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_void_p)]),
    ("hft_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("hft_last_error", ctypes.c_char_p, []),
    ("hft_dims", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    ("hft_set_state", ctypes.c_int, [ctypes.c_void_p, _D, ctypes.c_int, ctypes.c_double]),
    ("hft_get_state", ctypes.c_int, [ctypes.c_void_p, _D, ctypes.c_int]),
    ("hft_set_flow_analytic", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D, _D, _D,
                                             ctypes.c_double, ctypes.c_double]),
    ("hft_set_flow_profiles", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("hft_set_flow_arrays", ctypes.c_int, [ctypes.c_void_p, _D, _D, ctypes.c_int]),
    ("hft_set_source", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("hft_rfe_generate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, _D]),
    ("hft_rfe_weights", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    ("hft_get_rfe", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    ("hft_source_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("hft_set_flow_attenuation", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("hft_update_flow", ctypes.c_int, [ctypes.c_void_p]),
    ("hft_flow_dtmax", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D]),
    ("hft_set_diffusion", ctypes.c_int, [ctypes.c_void_p, _D, ctypes.c_double]),
    ("hft_dt_euler", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("hft_flux_transport_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int, CYCLE_FN, ctypes.c_void_p]),
    ("hft_advection_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("hft_diffusion_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                          CYCLE_FN, ctypes.c_void_p]),
    ("hft_diffusion_stats", ctypes.c_int, [ctypes.c_void_p, _I64, _I64]),
    ("hft_apply_diffusion_operator", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("hft_apply_advection_operator", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("hft_assimilate", ctypes.c_int, [ctypes.c_void_p, _D, ctypes.c_int, ctypes.c_int, _D, _D, _D,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    ("hft_assimilate_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_int, _D, _D, _D, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_int]),
    ("hft_set_validation_field", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("hft_set_weno_eps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("hft_set_pole_flux_lat_limit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("hft_balance_flux", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("hft_analysis", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D]),
    ("hft_analysis_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                            ctypes.c_void_p]),
    ("hft_state_device", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_int)]),
    ("hft_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("hft_stream", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("hft_set_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    ("hft_get_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p,
                                      ctypes.POINTER(ctypes.c_int)]),
    ("hft_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hft_kernel_time", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _D, _I64, _D]),
    ("hft_kernel_stage_bytes", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _D]),
    ("hft_set_state_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    ("hft_get_state_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    ("hft_set_stream", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    ("hft_step_adjoint", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, _D,
                                        ctypes.c_int]),
    ("hft_steps_adjoint", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_int64, _D, ctypes.c_int]),
]

ADJ_FROZEN_WENO = 1

_LIB = None


def lib():
    """Load libhft.so (raises HftError when it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HftError(f"libhft.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return [s[0] for s in _SIGS]


def _check(rc, what):
    if rc != 0:
        msg = lib().hft_last_error().decode(errors="replace")
        raise HftError(f"{what} failed ({rc}): {msg}")


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _vec(x, n):
    if x is None:
        return None
    return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)))


class FluxTransport:
    """A batch of ``nr`` realizations on one theta-phi mesh, on the current GPU.

    t: theta nodes (ntm) on [0, pi]; p: phi nodes (npm1) on [0, 2 pi] with the
    one-point overlap, i.e. the scales of a PT map file (src/hipft.f90:2130-2132).
    mode: "coupled" (one controller for the batch, reference semantics) or
    "independent" (each realization behaves as its own single run).
    """

    def __init__(self, t, p, nr=1, mode="coupled"):
        t = np.ascontiguousarray(t, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        self.ntm, self.npm1, self.nr = len(t), len(p), int(nr)
        self.npm = self.npm1 + 1
        self.nt = self.ntm + 1
        self.mode = mode
        m = {"coupled": MODE_COUPLED, "independent": MODE_INDEPENDENT}[mode]
        h = ctypes.c_void_p()
        _check(lib().hft_create(self.ntm, self.npm1, _p(t), _p(p), self.nr, m, ctypes.byref(h)),
               "hft_create")
        self._h = h
        self._cb_keep = None

    def close(self):
        if getattr(self, "_h", None):
            lib().hft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ---------------------------------------------------------
    @property
    def tp_shape(self):
        return (self.nr, self.npm, self.ntm)

    @property
    def pt_shape(self):
        return (self.nr, self.ntm, self.npm1)

    def set_state(self, f, layout="tp", mult_fac=1.0):
        if layout == "tp":
            a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.tp_shape)
            _check(lib().hft_set_state(self._h, _p(a), LAYOUT_TP, mult_fac), "hft_set_state")
        else:
            a = np.asarray(f, dtype=np.float64)
            if a.ndim == 2:
                a = np.broadcast_to(a, (self.nr,) + a.shape)
            a = np.ascontiguousarray(a).reshape(self.pt_shape)
            _check(lib().hft_set_state(self._h, _p(a), LAYOUT_PT, mult_fac), "hft_set_state")

    def get_state(self, layout="tp"):
        shape = self.tp_shape if layout == "tp" else self.pt_shape
        out = np.empty(shape)
        _check(lib().hft_get_state(self._h, _p(out), LAYOUT_TP if layout == "tp" else LAYOUT_PT),
               "hft_get_state")
        return out

    # ---- sources ----------------------------------------------------------
    def set_source(self, src_tp=None):
        """Constant source map (G/hr), reference storage (npm, ntm) as a C
        array (Fortran src(ntm,npm)), shared by all realizations; None
        removes it."""
        if src_tp is None:
            _check(lib().hft_set_source(self._h, None), "hft_set_source")
            return
        a = np.ascontiguousarray(src_tp, dtype=np.float64)
        if a.shape != (self.npm, self.ntm):
            raise ValueError(f"set_source: shape {a.shape}, expected ({self.npm}, {self.ntm})")
        _check(lib().hft_set_source(self._h, _p(a)), "hft_set_source")

    def rfe_generate(self, slot, rnd, model, uf_per_hour):
        """Random flux emergence frame into slot 0/1 from the random map
        rnd [npm-2][ntm] (C order of rnd_nums(1:ntm, 2:npm-1))."""
        a = np.ascontiguousarray(rnd, dtype=np.float64)
        if a.shape != (self.npm - 2, self.ntm):
            raise ValueError(f"rfe_generate: shape {a.shape}, expected ({self.npm - 2}, {self.ntm})")
        _check(lib().hft_rfe_generate(self._h, int(slot), _p(a), int(model),
                                      _p(_vec(uf_per_hour, self.nr))), "hft_rfe_generate")

    def rfe_weights(self, first_slot, alpha):
        _check(lib().hft_rfe_weights(self._h, int(first_slot), float(alpha)), "hft_rfe_weights")

    def get_rfe(self, slot):
        out = np.empty(self.tp_shape)
        _check(lib().hft_get_rfe(self._h, int(slot), _p(out)), "hft_get_rfe")
        return out

    def source_step(self, dt):
        _check(lib().hft_source_step(self._h, float(dt)), "hft_source_step")

    # ---- flows ------------------------------------------------------------
    def set_flow_analytic(self, p0=None, p2=None, p4=None, p1=None, p3=None, p5=None,
                          vp_rigid_omega=0.0, vt_const=0.0):
        n = self.nr
        _check(lib().hft_set_flow_analytic(self._h, _p(_vec(p0, n)), _p(_vec(p2, n)),
                                           _p(_vec(p4, n)), _p(_vec(p1, n)), _p(_vec(p3, n)),
                                           _p(_vec(p5, n)), float(vp_rigid_omega),
                                           float(vt_const)), "hft_set_flow_analytic")
        self._flow = True

    def set_flow_profiles(self, vt=None, vp=None):
        a = None if vt is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(vt, dtype=np.float64), (self.nr, self.nt)))
        b = None if vp is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(vp, dtype=np.float64), (self.nr, self.ntm)))
        _check(lib().hft_set_flow_profiles(self._h, _p(a), _p(b)), "hft_set_flow_profiles")
        self._flow = True

    def set_flow_arrays(self, vt=None, vp=None):
        """2-D base flows (flows from files) in code units, in the reference's
        storage order as numpy arrays: vt [nrf][npm][nt] and vp [nrf][npm][ntm]
        (C order, i.e. Fortran vt(nt,npm,nrf)), nrf = 1 (shared by every
        realization) or nr.  None removes them."""
        if vt is None and vp is None:
            _check(lib().hft_set_flow_arrays(self._h, None, None, 0), "hft_set_flow_arrays")
            return
        vt = np.asarray(vt, dtype=np.float64)
        vp = np.asarray(vp, dtype=np.float64)
        if vt.ndim == 2:
            vt = vt[None]
        if vp.ndim == 2:
            vp = vp[None]
        nrf = vt.shape[0]
        if nrf not in (1, self.nr) or vp.shape[0] != nrf:
            raise ValueError(f"set_flow_arrays: {nrf} flows for {self.nr} realizations")
        if vt.shape[1:] != (self.npm, self.nt) or vp.shape[1:] != (self.npm, self.ntm):
            raise ValueError(f"set_flow_arrays: vt {vt.shape}, vp {vp.shape}; expected "
                             f"[nrf][{self.npm}][{self.nt}] and [nrf][{self.npm}][{self.ntm}]")
        a = np.ascontiguousarray(vt)
        b = np.ascontiguousarray(vp)
        _check(lib().hft_set_flow_arrays(self._h, _p(a), _p(b), int(nrf == self.nr and self.nr > 1)),
               "hft_set_flow_arrays")
        self._flow = True

    def has_flow(self):
        return getattr(self, "_flow", False)

    def set_flow_attenuation(self, values):
        _check(lib().hft_set_flow_attenuation(self._h, _p(_vec(values, self.nr))),
               "hft_set_flow_attenuation")

    def update_flow(self):
        _check(lib().hft_update_flow(self._h), "hft_update_flow")

    def flow_dtmax(self, strang):
        dt = ctypes.c_double()
        per = np.zeros(self.nr)
        _check(lib().hft_flow_dtmax(self._h, int(bool(strang)), ctypes.byref(dt), _p(per)),
               "hft_flow_dtmax")
        return dt.value, per

    # ---- diffusion --------------------------------------------------------
    def set_diffusion(self, eta, factor=DIFF_KM2_S_TO_RS2_HR):
        _check(lib().hft_set_diffusion(self._h, _p(_vec(eta, self.nr)), float(factor)),
               "hft_set_diffusion")

    def dt_euler(self):
        dt = ctypes.c_double()
        per = np.zeros(self.nr)
        _check(lib().hft_dt_euler(self._h, ctypes.byref(dt), _p(per)), "hft_dt_euler")
        return dt.value, per

    # ---- steps --------------------------------------------------------------
    def _wrap_cb(self, callback):
        if callback is None:
            return CYCLE_FN(0)
        nr = self.nr

        def fn(user, cycle, recs):
            rows = [(recs[r].time_stepped, recs[r].dt, recs[r].sts_s, recs[r].subcycles)
                    for r in range(nr)]
            callback(cycle, rows)

        cb = CYCLE_FN(fn)
        self._cb_keep = cb
        return cb

    def flux_transport_step(self, dt, strang=False, subcycles=0, callback=None):
        cb = self._wrap_cb(callback)
        _check(lib().hft_flux_transport_step(self._h, float(dt), int(bool(strang)), int(subcycles),
                                             cb, None), "hft_flux_transport_step")

    def advection_step(self, dt):
        _check(lib().hft_advection_step(self._h, float(dt)), "hft_advection_step")

    def diffusion_step(self, dt, subcycles=0, callback=None):
        cb = self._wrap_cb(callback)
        _check(lib().hft_diffusion_step(self._h, float(dt), int(subcycles), cb, None),
               "hft_diffusion_step")

    def diffusion_stats(self):
        s = np.zeros(self.nr, dtype=np.int64)
        c = np.zeros(self.nr, dtype=np.int64)
        _check(lib().hft_diffusion_stats(self._h, s.ctypes.data_as(_I64), c.ctypes.data_as(_I64)),
               "hft_diffusion_stats")
        return s, c

    # ---- operators ------------------------------------------------------------
    def apply_diffusion_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.tp_shape)
        y = np.empty(self.tp_shape)
        _check(lib().hft_apply_diffusion_operator(self._h, _p(x), _p(y)),
               "hft_apply_diffusion_operator")
        return y

    def apply_advection_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.tp_shape)
        y = np.empty(self.tp_shape)
        _check(lib().hft_apply_advection_operator(self._h, _p(x), _p(y)),
               "hft_apply_advection_operator")
        return y

    # ---- assimilation ---------------------------------------------------------
    def assimilate(self, data, custom_from_mu=True, mu_power=4.0, mu_limit=0.1, lat_limit=0.0,
                   mult_fac=1.0, add_mode=False, balance=False):
        """data: (3, ntm, npm1) shared or (nr, 3, ntm, npm1): slices data, weight, mu."""
        d = np.ascontiguousarray(data, dtype=np.float64)
        per = d.ndim == 4
        exp = (self.nr, 3, self.ntm, self.npm1) if per else (3, self.ntm, self.npm1)
        if d.shape != exp:
            raise HftError(f"assimilation map shape {d.shape}, expected {exp}")
        n = self.nr
        _check(lib().hft_assimilate(self._h, _p(d), int(per), int(custom_from_mu),
                                    _p(_vec(mu_power, n)), _p(_vec(mu_limit, n)),
                                    _p(_vec(lat_limit, n)), float(mult_fac), int(add_mode),
                                    int(balance)), "hft_assimilate")

    def assimilate_device(self, ptr, per_realization=False, custom_from_mu=True, mu_power=4.0,
                          mu_limit=0.1, lat_limit=0.0, mult_fac=1.0, add_mode=False,
                          balance=False):
        n = self.nr
        _check(lib().hft_assimilate_device(self._h, ctypes.c_void_p(ptr), int(per_realization),
                                           int(custom_from_mu), _p(_vec(mu_power, n)),
                                           _p(_vec(mu_limit, n)), _p(_vec(lat_limit, n)),
                                           float(mult_fac), int(add_mode), int(balance)),
               "hft_assimilate_device")

    # ---- analysis ---------------------------------------------------------------
    def set_validation_field(self, fval_pt):
        a = None if fval_pt is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(fval_pt, dtype=np.float64), self.pt_shape))
        _check(lib().hft_set_validation_field(self._h, _p(a)), "hft_set_validation_field")

    def set_weno_eps(self, eps):
        """weno_eps (load_weno): <= 0 uses the cell widths (the default)."""
        _check(lib().hft_set_weno_eps(self._h, float(eps)), "hft_set_weno_eps")

    def set_pole_flux_lat_limit(self, deg):
        _check(lib().hft_set_pole_flux_lat_limit(self._h, float(deg)),
               "hft_set_pole_flux_lat_limit")

    def balance_flux(self, copy=False):
        """balance_flux of every realization: in place, or (copy=True) a
        balanced PT copy of the state, returned, with the state unchanged."""
        if not copy:
            _check(lib().hft_balance_flux(self._h, None), "hft_balance_flux")
            return None
        out = np.zeros(self.pt_shape)
        _check(lib().hft_balance_flux(self._h, _p(out)), "hft_balance_flux")
        return out

    def analysis(self, fval_pt=None, fval_scale=None):
        out = np.zeros((self.nr, NHIST))
        a = None if fval_pt is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(fval_pt, dtype=np.float64), self.pt_shape))
        s = _vec(fval_scale, self.nr)
        _check(lib().hft_analysis(self._h, _p(a), _p(s), _p(out)), "hft_analysis")
        return out

    def analysis_noalloc(self, out_ptr, f_ptr=None, ld=0):
        """hft_analysis_noalloc: diagnostics of the state (or of a device batch
        f_ptr with row length ld) into the device array out_ptr [nr][14], on
        the context's stream, allocating nothing."""
        _check(lib().hft_analysis_noalloc(self._h, f_ptr, int(ld), out_ptr),
               "hft_analysis_noalloc")

    # ---- device access / timing -------------------------------------------------
    def state_device(self):
        ptr = ctypes.c_void_p()
        pitch = ctypes.c_int()
        _check(lib().hft_state_device(self._h, ctypes.byref(ptr), ctypes.byref(pitch)),
               "hft_state_device")
        return ptr.value, pitch.value

    def stream(self):
        s = ctypes.c_void_p()
        _check(lib().hft_stream(self._h, ctypes.byref(s)), "hft_stream")
        return s.value

    def synchronize(self):
        _check(lib().hft_synchronize(self._h), "hft_synchronize")

    def set_tuning(self, **kw):
        for k, v in kw.items():
            _check(lib().hft_set_tuning(self._h, k.encode(), int(v)), "hft_set_tuning")

    def get_tuning(self, key):
        v = ctypes.c_int()
        _check(lib().hft_get_tuning(self._h, key.encode(), ctypes.byref(v)), "hft_get_tuning")
        return v.value

    def set_timing(self, enable, phases=False):
        """Reset the timers; enable: events around every launch, phases:
        events at the boundaries of the step's phases (the phase_* groups)."""
        _check(lib().hft_set_timing(self._h, int(bool(enable)) | (2 if phases else 0)),
               "hft_set_timing")

    # ---- device-resident state and the adjoint -----------------------------
    def set_state_device(self, ptr, ld):
        """Copy a device array [nr][ntm][ld] (unique phi columns) into the state."""
        _check(lib().hft_set_state_device(self._h, ctypes.c_void_p(ptr), int(ld)),
               "hft_set_state_device")

    def get_state_device(self, ptr, ld):
        _check(lib().hft_get_state_device(self._h, ctypes.c_void_p(ptr), int(ld)),
               "hft_get_state_device")

    def set_stream(self, stream_ptr):
        """Launch on a caller's HIP stream (0 is the null stream, as torch's
        default stream reports it); None: the context's own stream again."""
        if stream_ptr is None:
            _check(lib().hft_set_stream(self._h, None, 1), "hft_set_stream")
        else:
            _check(lib().hft_set_stream(self._h, ctypes.c_void_p(stream_ptr or None), 0),
                   "hft_set_stream")

    def step_adjoint(self, dt, lam_ptr, ld, strang=True, subcycles=1, frozen_weno=False):
        """Adjoint of one flux_transport_step from the current state.  lam (device
        [nr][ntm][ld]) goes from dJ/dx_{n+1} to dJ/dx_n in place; returns
        [nr][3] = dJ/dlog(eta), dJ/dlog(vt scale), dJ/dlog(vp scale)."""
        g = np.zeros((self.nr, 3))
        _check(lib().hft_step_adjoint(self._h, float(dt), int(bool(strang)), int(subcycles),
                                      ctypes.c_void_p(lam_ptr), int(ld), _p(g),
                                      ADJ_FROZEN_WENO if frozen_weno else 0), "hft_step_adjoint")
        return g

    def steps_adjoint(self, nsteps, dt, lam_ptr, ld, strang=True, subcycles=1,
                      frozen_weno=False):
        """Adjoint of nsteps steps from the current state (which ends at the
        state after them); lam in place, returns [nr][3] as step_adjoint."""
        g = np.zeros((self.nr, 3))
        _check(lib().hft_steps_adjoint(self._h, int(nsteps), float(dt), int(bool(strang)),
                                       int(subcycles), ctypes.c_void_p(lam_ptr), int(ld), _p(g),
                                       ADJ_FROZEN_WENO if frozen_weno else 0),
               "hft_steps_adjoint")
        return g

    def kernel_time(self, group):
        """(total device ms, launches, algorithmic bytes) of a kernel group."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        b = ctypes.c_double()
        _check(lib().hft_kernel_time(self._h, group.encode(), ctypes.byref(ms), ctypes.byref(n),
                                     ctypes.byref(b)), "hft_kernel_time")
        return ms.value, n.value, b.value

    def kernel_stage_bytes(self, group):
        """Bytes the group's work would move one RKG2 stage per launch."""
        b = ctypes.c_double()
        _check(lib().hft_kernel_stage_bytes(self._h, group.encode(), ctypes.byref(b)),
               "hft_kernel_stage_bytes")
        return b.value
