"""Differentiable FP64 CPU transcription of the flux-transport step (PyTorch
autograd).  TEST INFRASTRUCTURE ONLY: it is the independent check of the
hand-written adjoint in libhft and is never used by the library.

It restates, with torch tensor operations, the same discrete step as the C
oracle (hft_oracle.c), taking the operator data (mesh, WENO3 weights,
Lax-Friedrichs alphas, diffusion coefficients, pole factors) from an oracle
context so that both describe one discretisation:

  * advection_operator_weno3   src/hipft.f90:7711-7925 (W below)
  * advection_step_ssprk43     src/hipft.f90:7515-7582
  * diffusion_operator_cd      src/hipft.f90:7927-7990 (L below)
  * diffusion_step_sts_cd      src/hipft.f90:5163-5242 (RKG2, stage counts
    from load_sts_rkg2 5361-5475)
  * flux_transport_step        src/hipft.f90:4463-4507 (Strang or A then D)

State layout: tensors (nr, ntm, nu), nu = npm-2 unique phi columns (unique
column c is reference column c+2; the seam columns are implied by periodic
rolls).  Parameters, per realization, all differentiable:
  logD  -- L = exp(logD) * L(eta), the diffusivity log-scale,
  st    -- scale of vt (and of alpha_t, which scales with |vt|),
  sp    -- scale of vp (and alpha_p).
At logD = 0, st = sp = 1 the forward values agree with the C oracle to
rounding (pole sums are summed in a different order).
"""
import numpy as np
import torch

from .oracle import sts_rkg2

F64 = torch.float64


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=F64)


def weno_face(LNm2, LNm1, LN0, LPm1, LP0, LPp1, w, frozen=False):
    """Face flux from split fluxes (src/hipft.f90:7775-7801), operation order
    of the reference.  frozen: the nonlinear weights are constants for
    differentiation (the linearisation with frozen smoothness weights)."""
    CPm, MCm, CP0, MC0, PTm, MCTm, MT0, CPT0, em, e0 = w
    p0m = (1.0 + MCm) * LNm1 - MCm * LNm2
    p1m = MC0 * LNm1 + CPm * LN0
    p0p = (1.0 + CP0) * LP0 - CP0 * LPp1
    p1p = CPm * LP0 + MC0 * LPm1
    t0 = MCm * (LNm1 - LNm2)
    t1 = CPm * (LN0 - LNm1)
    t2 = CP0 * (LPp1 - LP0)
    t3 = MC0 * (LP0 - LPm1)
    e0m = em + 4.0 * (t0 * t0)
    e1m = em + 4.0 * (t1 * t1)
    e2p = e0 + 4.0 * (t2 * t2)
    e3p = e0 + 4.0 * (t3 * t3)
    w0m = PTm / (e0m * e0m)
    w1m = MCTm / (e1m * e1m)
    w0p = MT0 / (e2p * e2p)
    w1p = CPT0 / (e3p * e3p)
    if frozen:
        w0m, w1m, w0p, w1p = (a.detach() for a in (w0m, w1m, w0p, w1p))
    wm = w0m + w1m
    wp = w0p + w1p
    um = (w0m / wm) * p0m + (w1m / wm) * p1m
    up = (w0p / wp) * p0p + (w1p / wp) * p1p
    return up + um


class TorchStep:
    """Differentiable step built from an oracle context whose flows
    (set_lf_alpha done) and diffusion (load_diffusion) are set."""

    def __init__(self, o, D=None, flow=True, diffusion=True, frozen_weno=False):
        self.o = o
        self.frozen = frozen_weno
        nr, ntm, npm, nt = o.nr, o.ntm, o.npm, o.nt
        nu = npm - 2
        self.nr, self.ntm, self.nu = nr, ntm, nu
        m = o.mesh
        self.sth = _t(m["sth"])      # [nt]  sth(j+1)
        self.st_i = _t(m["st_i"])    # [ntm]
        self.dt_i = _t(m["dt_i"])    # [ntm]
        self.dp = _t(m["dp"])        # [npm]
        self.dp_i = _t(m["dp_i"])    # [npm]
        pf = o.pole_factors()
        self.pf = pf
        self.flow = flow
        self.diffusion = diffusion
        if flow:
            vt, vp = o.flow_get()          # (nr, npm, nt), (nr, np, ntm)
            at, ap = o.alphas()            # (nr, npm, ntm)
            cols = slice(1, npm - 1)       # reference k = 2..npm-1
            tr = lambda a: _t(a).transpose(1, 2).contiguous()
            self.vta = tr(vt[:, cols, 0:ntm])      # VT(j) of row j
            self.vtb = tr(vt[:, cols, 1:nt])       # VT(j+1)
            self.at = tr(at[:, cols, :])
            self.vpa = tr(vp[:, cols, :])          # VP(j,k)
            self.vpb = tr(vp[:, 2:npm, :])         # VP(j,k+1)
            self.ap = tr(ap[:, cols, :])
            self.v_n = _t(vt[:, cols, 1])          # VT(2)   (nr, nu)
            self.v_s = _t(vt[:, cols, nt - 2])     # VT(nt-1)
            ww = o.weno_weights()
            # theta faces q = 1..ntm-3 use weight indices q (m) and q+1 (0)
            th = [_t(a) for a in ww["theta"]]
            CP, MC, MT, CPT, PT, MCT, eps = th
            self.wt = (CP[1:ntm - 2, None], MC[1:ntm - 2, None], CP[2:ntm - 1, None],
                       MC[2:ntm - 1, None], PT[1:ntm - 2, None], MCT[1:ntm - 2, None],
                       MT[2:ntm - 1, None], CPT[2:ntm - 1, None], eps[1:ntm - 2, None],
                       eps[2:ntm - 1, None])
            ph = [_t(a) for a in ww["phi"]]
            CP, MC, MT, CPT, PT, MCT, eps = ph
            self.wp = (CP[0:nu], MC[0:nu], CP[1:nu + 1], MC[1:nu + 1], PT[0:nu], MCT[0:nu],
                       MT[1:nu + 1], CPT[1:nu + 1], eps[0:nu], eps[1:nu + 1])
        if diffusion:
            c = o.diffusion_coef()                 # (nr, 5, npm-2, ntm-2)
            self.coef = _t(c).transpose(2, 3).contiguous()   # (nr, 5, ntm-2, nu)
            assert D is not None, "D (diffusivity in code units per realization) required"
            self.D2 = _t(np.asarray(D, dtype=np.float64) * 2.0)
            self.dte = np.array([o.dt_euler_r(r) for r in range(nr)])

    # ---- advection -------------------------------------------------------
    def W(self, x, st, sp):
        nr, ntm, nu = self.nr, self.ntm, self.nu
        s_t = st[:, None, None]
        s_p = sp[:, None, None]
        LN = 0.5 * x * (s_t * self.vta + s_t * self.at)
        LP = 0.5 * x * (s_t * self.vtb - s_t * self.at)
        Fint = weno_face(LN[:, 0:ntm - 3], LN[:, 1:ntm - 2], LN[:, 2:ntm - 1],
                         LP[:, 1:ntm - 2], LP[:, 2:ntm - 1], LP[:, 3:ntm], self.wt,
                         self.frozen)
        vn = st[:, None] * self.v_n
        vs = st[:, None] * self.v_s
        cn = torch.where(torch.signbit(vn), -1.0, 1.0).to(F64)
        cs = torch.where(torch.signbit(vs), -1.0, 1.0).to(F64)
        Fn = vn * 0.5 * ((1.0 - cn) * x[:, 1] + (1.0 + cn) * x[:, 0])
        Fs = vs * 0.5 * ((1.0 - cs) * x[:, ntm - 1] + (1.0 + cs) * x[:, ntm - 2])
        FT = torch.cat([Fn[:, None], Fint, Fs[:, None]], dim=1)    # faces 0..ntm-2
        LNp = 0.5 * x * (s_p * self.vpa + s_p * self.ap)
        LPp = 0.5 * x * (s_p * self.vpb - s_p * self.ap)
        r = lambda a, k: torch.roll(a, k, dims=2)   # r(a, k)[c] = a[c-k]
        FP = weno_face(r(LNp, 2), r(LNp, 1), LNp, r(LPp, 1), LPp, r(LPp, -1), self.wp,
                       self.frozen)
        j = slice(1, ntm - 1)
        sth, st_i, dt_i = self.sth, self.st_i, self.dt_i
        a = (sth[2:ntm, None] * FT[:, 1:ntm - 1] - sth[1:ntm - 1, None] * FT[:, 0:ntm - 2])
        a = a * st_i[1:ntm - 1, None] * dt_i[1:ntm - 1, None]
        b = (r(FP, -1)[:, j] - FP[:, j]) * st_i[1:ntm - 1, None] * self.dp_i[1:nu + 1]
        inner = a + b
        fn = (FT[:, 0] * self.dp[1:nu + 1]).sum(dim=1)
        fs = (FT[:, ntm - 2] * self.dp[1:nu + 1]).sum(dim=1)
        top = (fn * self.pf["flow_n"])[:, None, None].expand(nr, 1, nu)
        bot = (-fs * self.pf["flow_s"])[:, None, None].expand(nr, 1, nu)
        return torch.cat([top, inner, bot], dim=1)

    def ssprk43(self, x, dtl, st, sp):
        f1 = x - 0.5 * dtl * self.W(x, st, sp)
        f1 = f1 - 0.5 * dtl * self.W(f1, st, sp)
        f1 = 0.66666666666666666 * x + 0.33333333333333333 * f1 - \
            0.16666666666666666 * dtl * self.W(f1, st, sp)
        return f1 - 0.5 * dtl * self.W(f1, st, sp)

    # ---- diffusion -------------------------------------------------------
    def L(self, x, logD):
        nr, ntm, nu = self.nr, self.ntm, self.nu
        c = self.coef
        j = slice(1, ntm - 1)
        xc = x[:, j]
        xw = torch.roll(x, 1, dims=2)[:, j]
        xe = torch.roll(x, -1, dims=2)[:, j]
        xn = x[:, 0:ntm - 2]
        xs = x[:, 2:ntm]
        y = c[:, 0] * xw + c[:, 1] * xn + c[:, 2] * xc + c[:, 3] * xs + c[:, 4] * xe
        D2 = self.D2[:, None]
        fn = ((D2 * (x[:, 1] - x[:, 0])) * self.dp[1:nu + 1]).sum(dim=1)
        fs = ((D2 * (x[:, ntm - 1] - x[:, ntm - 2])) * self.dp[1:nu + 1]).sum(dim=1)
        top = (fn * self.pf["diff_n"])[:, None, None].expand(nr, 1, nu)
        bot = (-fs * self.pf["diff_s"])[:, None, None].expand(nr, 1, nu)
        out = torch.cat([top, y, bot], dim=1)
        return torch.exp(logD)[:, None, None] * out

    def rkg2_cycle(self, x, dt2, s_r, logD):
        """One RKG2 super-step per realization with its own stage count."""
        outs = []
        for r in range(self.nr):
            s, uj, vj, ubj, gj = sts_rkg2(dt2, self.dte[r]) if s_r is None else \
                _coefs(int(s_r[r]))
            xr = x[r:r + 1]
            sub = _Sub(self, r)
            dty0 = dt2 * sub.L(xr, logD[r:r + 1])
            u0 = xr
            ukm2 = xr
            ukm1 = xr + ubj[0] * dty0
            for k in range(2, s + 1):
                a, b, cc, g = uj[k - 1], vj[k - 1], ubj[k - 1], gj[k - 1]
                fk = a * ukm1 + b * ukm2 + (1.0 - a - b) * u0 + (cc * dt2) * sub.L(
                    ukm1, logD[r:r + 1]) + g * dty0
                ukm2, ukm1 = ukm1, fk
            outs.append(ukm1)
        return torch.cat(outs, dim=0)

    def step(self, x, logD, st, sp, dt, strang=True, subcycles=1, s_r=None):
        if strang:
            if self.flow:
                x = self.ssprk43(x, 0.5 * dt, st, sp)
            if self.diffusion:
                for _ in range(subcycles):
                    x = self.rkg2_cycle(x, dt / subcycles, s_r, logD)
            if self.flow:
                x = self.ssprk43(x, 0.5 * dt, st, sp)
        else:
            if self.flow:
                x = self.ssprk43(x, dt, st, sp)
            if self.diffusion:
                for _ in range(subcycles):
                    x = self.rkg2_cycle(x, dt / subcycles, s_r, logD)
        return x

    def stage_counts(self, dt, subcycles=1):
        return [sts_rkg2(dt / subcycles, self.dte[r])[0] for r in range(self.nr)]


class _Sub:
    """Diffusion operator of one realization of a TorchStep."""

    def __init__(self, ts, r):
        self.ts, self.r = ts, r

    def L(self, x, logD):
        ts, r = self.ts, self.r
        full = TorchStep.__new__(TorchStep)
        full.nr, full.ntm, full.nu = 1, ts.ntm, ts.nu
        full.coef = ts.coef[r:r + 1]
        full.D2 = ts.D2[r:r + 1]
        full.dp, full.pf = ts.dp, ts.pf
        return TorchStep.L(full, x, logD)


def _coefs(s):
    """RKG2 vectors for a given odd stage count s >= 3."""
    # dt/dt_euler that yields exactly s: s_real = (sqrt(25+24 q) - 3)/2 in (s-2, s]
    q = (((2.0 * s - 0.5) + 3.0) ** 2 - 25.0) / 24.0
    s2, u, v, ub, g = sts_rkg2(q, 1.0)
    assert s2 == s, (s2, s)
    return s2, u, v, ub, g
