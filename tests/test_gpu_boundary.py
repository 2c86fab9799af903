"""The fused boundary entry points through the C ABI, for an engine that keeps
its own call sequence:

  um_field_torque   SetTorque after SetDemagField (engine/torque.go:124-133,
                    engine/effectivefield.go:14-24) in one pass per cell;
  um_rk45dp_stage   MaddN + Normalize of one RK stage (engine/rk45dp.go:50-87);
  um_rk45dp_error   integralController's err, rlerr and maxdm
                    (engine/stepcontroller.go:35-62) from the stage torques.

Each is checked against the FP64 oracle: fields and torques normwise and per
cell (|dB_i| / |B_i|, torque error over |B_i|) to 1e-5 in FP32 and 1e-12 in
FP64; stage sums to 1e-6 / 1e-14; controller norms to 1e-6 / 1e-13.  An
RK45DP integrator written the way engine/rk45dp.go and stepcontroller.go
drive the reference's opencl package, but calling these entry points, is run
next to the simulation handle (um_sim_steps): adaptive and fixed steps agree
bit for bit, and the fixed-step run stays within FP32 rounding of the oracle.
"""
import math
import sys

import numpy as np
import pytest

from cases import GAMMA, percell_relerr, percell_torque_relerr, random_unit, relerr, setup, strip_m

pytestmark = pytest.mark.gpu

KB = 1.380650424e-23


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    from umagnus_amd import umhip
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    umhip.init(0)
    return umhip


def _t(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype)).cuda()


def _h3(ts):
    return np.stack([t.detach().cpu().numpy() for t in ts])


def _vec(a, dtype):
    return [_t(a[c], dtype) for c in range(3)]


def _zeros3(shape, dtype):
    return [_t(np.zeros(shape), dtype) for _ in range(3)]


# ---- um_field_torque ----------------------------------------------------------------------
def _case(name):
    """(N, cs, setup() parameters, per-cell Ms or None, STT parameters or None)."""
    if name == "film":
        return (32, 24, 6), (2e-9,) * 3, dict(Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1), None, None
    if name == "four_cell":
        return (128, 64, 64), (2e-9,) * 3, dict(Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1), None, None
    if name == "regions_thermal":
        N = (20, 16, 3)
        rng = np.random.default_rng(8)
        shape = (N[2], N[1], N[0])
        regions = rng.integers(0, 4, shape).astype(np.uint8)
        Ms = rng.uniform(6e5, 1e6, shape)
        lut = [np.zeros(256), np.zeros(256), np.zeros(256)]
        lut[0][1], lut[2][2] = 0.05, -0.02
        frozen = np.zeros(256)
        frozen[3] = 1
        p = dict(Ms=Ms, Aex={0: 1e-11, 1: 2e-11, 2: 1.5e-11, 3: 1.2e-11}, Kc=(4e5, -1e5, 0.0), c1=(1, 1, 0),
                 c2=(-1, 1, 0), alpha=0.05, regions=regions, bext_lut=lut, frozen=frozen, temp=300.0, seed=99,
                 Bext=(0.01, 0.0, 0.02), Ku1=2e5, Ku2=-5e4, anisU=(0.3, 0.4, 0.866))
        return N, (2e-9, 2.5e-9, 3e-9), p, Ms, None
    if name == "stt":
        stt = dict(J=(2e11, 0.0, 1e11), xi=0.05, pol=0.5, lambda_=1.0, eps_prime=0.1, free_layer_thickness=0.0,
                   fixed_layer=(0.0, 0.0, 1.0), fixed_layer_top=1.0, zhangli=1, slonczewski=1)
        return (16, 12, 4), (2e-9, 3e-9, 2.5e-9), dict(Ms=8e5, Aex=1.3e-11, Dind=1e-3, alpha=0.1), None, stt
    raise KeyError(name)


def _oracle_stt(so, oracle, stt):
    so.set_vector(oracle.P_J, stt["J"])
    so.set_scalar(oracle.P_XI, stt["xi"])
    so.set_scalar(oracle.P_POL, stt["pol"])
    so.set_scalar(oracle.P_LAMBDA, stt["lambda_"])
    so.set_scalar(oracle.P_EPSPRIME, stt["eps_prime"])
    so.set_scalar(oracle.P_FREELAYER, stt["free_layer_thickness"])
    so.set_vector(oracle.P_FIXEDLAYER, stt["fixed_layer"])
    so.set_stt_flags(0, 0, True)


def _gpu_field(gpu, N, cs, p, Ms_cells, stt, m, dtype, dt=1e-14, badd=None, tadd=None):
    """B_demag through the demag plan, then um_field_torque -> (B_eff, torque) host arrays."""
    from umagnus_amd.params import exchange_lut
    shape = (N[2], N[1], N[0])
    prec = 0 if dtype == np.float32 else 1
    mt = _vec(m, dtype)
    Ms = _t(Ms_cells, dtype) if Ms_cells is not None else p["Ms"]
    conv = gpu.DemagConvolution(N, cs, precision=prec)
    bd = _zeros3(shape, dtype)
    conv.Exec(bd, mt, None, Ms)
    B, tau = _zeros3(shape, dtype), _zeros3(shape, dtype)
    Dind = p.get("Dind", 0.0)
    kw = dict(Msat=Ms, alpha=p["alpha"], bdemag=bd, beff=B, torque=tau,
              aex_lut=_t(exchange_lut(p["Aex"], dtype=dtype), dtype),
              dmi_lut=_t(exchange_lut(Dind, dtype=dtype), dtype) if (not np.isscalar(Dind) or Dind != 0) else None,
              Ku1=p.get("Ku1", 0.0), Ku2=p.get("Ku2", 0.0), anisU=p.get("anisU", (0, 0, 1)),
              Kc1=p.get("Kc", (0, 0, 0))[0], Kc2=p.get("Kc", (0, 0, 0))[1], Kc3=p.get("Kc", (0, 0, 0))[2],
              anisC1=p.get("c1", (1, 0, 0)), anisC2=p.get("c2", (0, 1, 0)), Bext=p.get("Bext", (0, 0, 0)),
              badd=badd, tadd=tadd)
    if p.get("regions") is not None:
        kw["regions"] = _t(p["regions"], np.uint8)
    if p.get("bext_lut") is not None:
        kw["bext_lut"] = [_t(l, dtype) for l in p["bext_lut"]]
    if p.get("frozen") is not None:
        kw["frozen_lut"] = _t(p["frozen"], dtype)
    if p.get("temp", 0.0):
        k = 2 * KB / (GAMMA * cs[0] * cs[1] * cs[2] * dt)
        kw["thermal"] = (p.get("seed", 0), 0, 0, p["temp"], k)
    if stt is not None:
        kw["stt"] = stt
    gpu.FieldTorque(mt, cs, **kw)
    return _h3(B), _h3(tau)


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("name", ["film", "four_cell", "regions_thermal", "stt"])
def test_field_torque_vs_oracle(gpu, oracle, name, prec):
    N, cs, p, Ms_cells, stt = _case(name)
    _say("field_torque", name, prec)
    dtype = np.float32 if prec == 0 else np.float64
    so = oracle.Sim(N, cs, dtype=np.float64)
    setup(so, oracle, dtype=np.float64, **p)
    if stt is not None:
        _oracle_stt(so, oracle, stt)
    so.set_solver_params(dt=1e-14)
    m = random_unit((N[2], N[1], N[0]), 17)
    so.set_m(m)
    b_o, t_o = so.beff(), so.torque()
    b_g, t_g = _gpu_field(gpu, N, cs, p, Ms_cells, stt, m, dtype)
    tol = 1e-5 if prec == 0 else 1e-12
    eb, et = percell_relerr(b_g, b_o), percell_torque_relerr(t_g, t_o, b_o)
    _say("  per cell B", eb, "tau", et, "normwise", relerr(b_g, b_o), relerr(t_g, t_o))
    assert relerr(b_g, b_o) <= tol and relerr(t_g, t_o) <= tol
    assert eb <= tol and et <= tol


@pytest.mark.parametrize("prec", [1, 0])
def test_field_torque_custom_terms(gpu, prec):
    """badd enters B_eff after every other term (AddCustomField) and tadd the
    torque before FreezeSpins (AddCustomTorques): B_eff(badd) = B_eff + badd and
    tau(tadd) = tau + tadd bit for bit, frozen cells stay zero."""
    import torch
    N, cs, p, Ms_cells, stt = _case("regions_thermal")
    dtype = np.float32 if prec == 0 else np.float64
    shape = (N[2], N[1], N[0])
    m = random_unit(shape, 4)
    rng = np.random.default_rng(3)
    badd_h = rng.normal(0, 0.05, (3,) + shape)
    tadd_h = rng.normal(0, 0.05, (3,) + shape)
    b0, t0 = _gpu_field(gpu, N, cs, p, Ms_cells, stt, m, dtype)
    b1, t1 = _gpu_field(gpu, N, cs, p, Ms_cells, stt, m, dtype, badd=_vec(badd_h, dtype), tadd=_vec(tadd_h, dtype))
    assert np.array_equal(b1, (b0 + badd_h.astype(dtype)).astype(dtype))
    frozen = np.isin(p["regions"], [3])
    # the torque with badd differs through B; compare tadd against a run with badd only
    b2, t2 = _gpu_field(gpu, N, cs, p, Ms_cells, stt, m, dtype, badd=_vec(badd_h, dtype))
    want = (t2 + tadd_h.astype(dtype)).astype(dtype)
    want[:, frozen] = 0
    assert np.array_equal(t1, want)
    assert np.all(t1[:, frozen] == 0)
    torch.cuda.synchronize()


@pytest.mark.parametrize("name", ["film", "four_cell"])
def test_field_torque_in_place_over_bdemag(gpu, name):
    """SetLLTorque writes the torque over B_eff in dst (engine/torque.go:136-144):
    torque aliasing bdemag gives the out-of-place result bit for bit."""
    from umagnus_amd.params import exchange_lut
    N, cs, p, _, _ = _case(name)
    shape = (N[2], N[1], N[0])
    dtype = np.float32
    mt = _vec(random_unit(shape, 5), dtype)
    conv = gpu.DemagConvolution(N, cs, precision=0)
    bd = _zeros3(shape, dtype)
    conv.Exec(bd, mt, None, p["Ms"])
    bd2 = [b.clone() for b in bd]
    tau = _zeros3(shape, dtype)
    kw = dict(Msat=p["Ms"], alpha=p["alpha"], aex_lut=_t(exchange_lut(p["Aex"], dtype=dtype), dtype),
              dmi_lut=_t(exchange_lut(p["Dind"], dtype=dtype), dtype), Ku1=p["Ku1"], anisU=(0, 0, 1))
    gpu.FieldTorque(mt, cs, bdemag=bd, torque=tau, **kw)
    gpu.FieldTorque(mt, cs, bdemag=bd2, torque=bd2, **kw)
    assert np.array_equal(_h3(tau), _h3(bd2))


# ---- um_rk45dp_stage / um_rk45dp_error --------------------------------------------------------
def _oracle_madd(oracle, dtype, srcs, facs):
    n = srcs[0].size
    dst = np.zeros(n, dtype)
    s = [np.ascontiguousarray(x, dtype).reshape(-1) for x in srcs]
    f = np.ascontiguousarray(facs, dtype)
    arr = (oracle.C.c_void_p * len(s))(*[oracle.ptr(x) for x in s])
    oracle.fn("madd", dtype)(oracle.ptr(dst), len(s), arr, oracle.ptr(f), n)
    return dst


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("nsrc,normalize", [(2, 1), (3, 1), (6, 1), (2, 0)])
def test_rk45dp_stage_vs_oracle(gpu, oracle, prec, nsrc, normalize):
    dtype = np.float32 if prec == 0 else np.float64
    shape = (3, 5, 40, 33)
    rng = np.random.default_rng(nsrc + 10 * normalize)
    srcs = [random_unit(shape[1:], 50)] + [rng.normal(0, 1.0, shape) for _ in range(nsrc - 1)]
    srcs = [s.astype(dtype) for s in srcs]
    facs = [1.0] + list(rng.uniform(-0.3, 0.3, nsrc - 1))
    vol = rng.uniform(0.0, 1.0, shape[1:]).astype(dtype)
    vol[0, 0, :4] = 0
    dst = _zeros3(shape[1:], dtype)
    gpu.RK45DPStage(dst, [_vec(s, dtype) for s in srcs], facs, vol=_t(vol, dtype), normalize=bool(normalize))
    got = _h3(dst).astype(np.float64)
    want = np.stack([_oracle_madd(oracle, np.float64, [s[c].astype(np.float64) for s in srcs],
                                  [float(dtype(f)) for f in facs]) for c in range(3)]).reshape(shape)
    if normalize:
        w = np.ascontiguousarray(want.reshape(3, -1))
        v = np.ascontiguousarray(vol.astype(np.float64).reshape(-1))
        oracle.fn("normalize2", np.float64)(oracle.ptr(w[0]), oracle.ptr(w[1]), oracle.ptr(w[2]), oracle.ptr(v),
                                             w.shape[1])
        want = w.reshape(shape)
    tol = 1e-6 if prec == 0 else 1e-14
    assert np.abs(got - want).max() <= tol * max(np.abs(want).max(), 1.0)
    if normalize:
        assert np.all(got[:, vol == 0] == 0)


E_DP = [35 / 384 - 5179 / 57600, 500 / 1113 - 7571 / 16695, 125 / 192 - 393 / 640,
        -2187 / 6784 - -92097 / 339200, 11 / 84 - 187 / 2100, 0 - 1 / 40]
E_BS = [7 / 24 - 2 / 9, 1 / 4 - 1 / 3, 1 / 3 - 4 / 9, 1 / 8]


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("e", [E_DP, E_BS], ids=["rk45dp", "rk23"])
def test_rk45dp_error_vs_oracle(gpu, oracle, prec, e):
    """err = MaxVecNorm(Err) h, rlerr = MaxAbs(VecNorm(Err) / VecNorm(delM)) (0
    where |delM| = 0, as div.cl), maxdm = MaxVecNorm(delM), against the oracle's
    Madd and MaxVecNorm and the same norms in FP64."""
    dtype = np.float32 if prec == 0 else np.float64
    shape = (3, 4, 24, 40)
    rng = np.random.default_rng(len(e))
    ks = [rng.normal(0, 1.0, shape).astype(dtype) for _ in e]
    ks[-1][:, 0, 0, :3] = 0                  # |delM| = 0 cells
    h = float(np.float32(1e-14 * GAMMA))
    got = gpu.RK45DPError([_vec(k, dtype) for k in ks], e, h)
    ef = [float(dtype(x)) for x in e]
    Err = np.stack([_oracle_madd(oracle, np.float64, [k[c].astype(np.float64) for k in ks], ef) for c in range(3)])
    dM = ks[-1].reshape(3, -1).astype(np.float64)
    en, dn = np.linalg.norm(Err, axis=0), np.linalg.norm(dM, axis=0)
    r = np.where(dn == 0, 0.0, en / np.where(dn == 0, 1.0, dn))
    Ec = np.ascontiguousarray(Err)
    err_o = oracle.fn("maxvecnorm", np.float64)(oracle.ptr(Ec[0]), oracle.ptr(Ec[1]), oracle.ptr(Ec[2]),
                                                Ec.shape[1]) * h
    want = (err_o, float(r.max()), float(dn.max()))
    tol = 1e-6 if prec == 0 else 1e-13
    for g, w in zip(got, want):
        assert abs(g - w) <= tol * abs(w), (got, want)


# ---- an RK45DP integrator over the boundary, as the engine drives it ---------------------------------
class EngineRK45DP:
    """RK45DP.Step (engine/rk45dp.go:12-98) with integralController
    (engine/stepcontroller.go:35-97) and adaptDt (engine/run.go:150-187),
    written over the C ABI's per-evaluation and per-stage entry points the
    way the engine calls the opencl package: SetTorque = demag Exec + one
    um_field_torque; every stage one um_rk45dp_stage; the controller one
    um_rk45dp_error."""

    def __init__(self, gpu, N, cs, dtype, params, m, dt=1e-15, fixdt=0.0, maxerr=1e-5, relerr=1e-4, minslope=1e8,
                 headroom=0.8):
        from umagnus_amd.params import exchange_lut
        import torch
        self.gpu, self.N, self.cs, self.dtype = gpu, N, cs, dtype
        self.shape = (N[2], N[1], N[0])
        self.prec = 0 if dtype == np.float32 else 1
        self.conv = gpu.DemagConvolution(N, cs, precision=self.prec)
        self.m = _vec(m, dtype)
        self.gpu.Normalize(self.m)
        self.m0, self.bd = _zeros3(self.shape, dtype), _zeros3(self.shape, dtype)
        self.k = [_zeros3(self.shape, dtype) for _ in range(6)]
        self.k1_valid = False
        self.p = params
        Dind = params.get("Dind", 0.0)
        self.aex = _t(exchange_lut(params["Aex"], dtype=dtype), dtype)
        self.dmi = _t(exchange_lut(Dind, dtype=dtype), dtype) if Dind else None
        self.dt, self.fixdt, self.maxerr, self.relerr, self.minslope = dt, fixdt, maxerr, relerr, minslope
        self.headroom, self.mindt, self.maxdt = headroom, 0.0, 0.0
        self.Time, self.nsteps, self.nundone, self.lastErr, self.peakErr, self.lastTorque = 0.0, 0, 0, 0.0, 0.0, 0.0
        self.nevals = 0
        self.torch = torch

    def torque_fn(self, dst):
        p = self.p
        self.conv.Exec(self.bd, self.m, None, p["Ms"])
        self.gpu.FieldTorque(self.m, self.cs, Msat=p["Ms"], alpha=p["alpha"], bdemag=self.bd, torque=dst,
                             aex_lut=self.aex, dmi_lut=self.dmi, Ku1=p.get("Ku1", 0.0), anisU=p.get("anisU", (0, 0, 1)),
                             Bext=p.get("Bext", (0, 0, 0)))
        self.nevals += 1

    def adapt(self, corr):
        if self.fixdt != 0:
            self.dt = self.fixdt
            return
        if math.isnan(corr):
            corr = 1.0
        corr *= self.headroom
        corr = min(max(corr, 0.5), 2.0)
        self.dt *= corr

    def copy(self, dst, src):
        for c in range(3):
            dst[c].copy_(src[c])

    def step(self):
        if self.fixdt != 0:
            self.dt = self.fixdt
        k1, k2, k3, k4, k5, k6 = self.k
        if not self.k1_valid:
            self.torque_fn(k1)
            self.k1_valid = True
        t0 = self.Time
        self.copy(self.m0, self.m)
        f32 = np.float32 if self.dtype == np.float32 else np.float64
        h = f32(self.dt * GAMMA)
        m, m0 = self.m, self.m0

        def stage(srcs, coef):
            self.gpu.RK45DPStage(m, srcs, [1.0] + [float(f32(c) * h) for c in coef])
        stage([m0, k1], [1 / 5])
        self.torque_fn(k2)
        stage([m0, k1, k2], [3 / 40, 9 / 40])
        self.torque_fn(k3)
        stage([m0, k1, k2, k3], [44 / 45, -56 / 15, 32 / 9])
        self.torque_fn(k4)
        stage([m0, k1, k2, k3, k4], [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729])
        self.torque_fn(k5)
        stage([m0, k1, k2, k3, k4, k5], [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656])
        self.torque_fn(k6)
        stage([m0, k1, k3, k4, k5, k6], [35 / 384, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84])
        k7 = k2
        self.torque_fn(k7)
        err, rlerr_v, maxdm = self.gpu.RK45DPError([k1, k3, k4, k5, k6, k7], E_DP, float(h))
        if err < self.maxerr or self.dt <= self.mindt or self.fixdt != 0:
            fail, rlerr = (0, 0.0) if maxdm < self.minslope else (1, rlerr_v)
            if fail == 0 or self.relerr <= 0 or rlerr < self.relerr or self.dt <= self.mindt or self.fixdt != 0:
                self.lastErr = err
                self.peakErr = max(self.peakErr, err)
                self.lastTorque = maxdm
                self.nsteps += 1
                self.Time = t0 + self.dt
                self.adapt(math.pow(self.maxerr / err, 1 / 5) if fail == 0 else math.pow(self.relerr / rlerr, 1 / 5))
                self.copy(k1, k7)
            else:
                self.Time = t0
                self.copy(m, m0)
                self.nundone += 1
                self.adapt(math.pow(self.relerr / rlerr, 1 / 6))
        else:
            self.Time = t0
            self.copy(m, m0)
            self.nundone += 1
            self.adapt(math.pow(self.maxerr / err, 1 / 6))

    def get_m(self):
        return _h3(self.m)


def _two_kernel_sim(gpu, N, cs, dtype, p, m):
    """A simulation handle whose torque runs the X inverse and the field kernel
    apart, the kernels um_demag_exec + um_field_torque launch: the engine
    sequence over the boundary is then the same arithmetic, bit for bit (the
    fused X-inverse + field kernel computes the same terms, with the compiler
    free to contract them into FMAs differently).  The demag plan reads the
    switch when it is built, at the first evaluation."""
    import os
    os.environ["UMHIP_FUSED_FIELD"] = "0"
    try:
        sim = gpu.Sim(N, cs, dtype=dtype)
        setup(sim, gpu, dtype=dtype, **p)
        sim.set_m(m)
        sim.demag_field()   # builds the plan without counting an evaluation
        return sim
    finally:
        os.environ.pop("UMHIP_FUSED_FIELD", None)


@pytest.mark.parametrize("prec", [1, 0])
def test_engine_rk45dp_adaptive_matches_sim(gpu, prec):
    """The 128x32x1 strip, adaptive RK45DP (MaxErr 1e-5), 40 steps: the engine's
    sequence over the boundary and the simulation handle agree bit for bit."""
    dtype = np.float32 if prec == 0 else np.float64
    N, cs = (128, 32, 1), (3.90625e-9, 3.90625e-9, 3e-9)
    p = dict(Ms=8e5, Aex=1.3e-11, alpha=0.02, Bext=(-24.6e-3, 4.3e-3, 0.0))
    m = strip_m(1)
    eng = EngineRK45DP(gpu, N, cs, dtype, p, m)
    sim = _two_kernel_sim(gpu, N, cs, dtype, p, m)
    sim.set_solver_params(dt=1e-15)
    while eng.nsteps < 40:
        eng.step()
    sim.steps(40)
    st = sim.state()
    _say("engine", eng.nsteps, eng.nundone, eng.Time, eng.dt, "sim", st)
    assert (st["nsteps"], st["nundone"], st["nevals"]) == (eng.nsteps, eng.nundone, eng.nevals)
    assert st["t"] == eng.Time and st["dt"] == eng.dt
    assert st["lasterr"] == eng.lastErr and st["peakerr"] == eng.peakErr and st["lasttorque"] == eng.lastTorque
    assert np.array_equal(sim.get_m(), eng.get_m())


@pytest.mark.parametrize("prec", [1, 0])
def test_engine_rk45dp_fixed_vs_oracle(gpu, oracle, prec):
    """Config-2 physics on a small film, FixDt 1e-14, 5 steps: the engine's
    sequence over the boundary vs the simulation handle (bit for bit) and the
    FP64 oracle (per cell |dm| <= 1e-5 in FP32, 1e-11 in FP64)."""
    dtype = np.float32 if prec == 0 else np.float64
    N, cs = (32, 24, 6), (2e-9,) * 3
    p = dict(Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1)
    m = random_unit((N[2], N[1], N[0]), 23)
    eng = EngineRK45DP(gpu, N, cs, dtype, p, m, dt=1e-14, fixdt=1e-14)
    for _ in range(5):
        eng.step()
    sim = _two_kernel_sim(gpu, N, cs, dtype, p, m)
    sim.set_solver_params(dt=1e-14, fixdt=1e-14)
    sim.steps(5)
    assert np.array_equal(sim.get_m(), eng.get_m())
    so = oracle.Sim(N, cs, dtype=np.float64)
    setup(so, oracle, dtype=np.float64, **p)
    so.set_m(m)
    so.set_solver_params(dt=1e-14, fixdt=1e-14)
    so.steps(5)
    d = float(np.linalg.norm(eng.get_m().astype(np.float64) - so.get_m(), axis=0).max())
    _say("engine vs oracle", d)
    assert d <= (1e-5 if prec == 0 else 1e-11)
