"""ctypes binding of libumhip.so (include/umhip.h).

Two layers, as in the reference:
  * the exported API of package opencl (opencl/*.go) -- one function per
    reference function (AddExchange, AddDMI, LLTorque, Madd, MaxVecNorm,
    DemagConvolution.Exec, ...) operating on device arrays; device arrays are
    torch tensors on cuda:0 (torch is only the allocator/stream here);
  * the engine-level simulation handle (Sim): parameters, effective field,
    torque and the RK45DP / Heun / RK4 steppers on host numpy arrays.

There is no CPU fallback: importing works anywhere, but every call needs the
built library and a GPU, and raises otherwise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libumhip.so")
_lib = None

F32, F64 = 0, 1
(P_MSAT, P_KU1, P_KU2, P_KC1, P_KC2, P_KC3, P_ALPHA, P_TEMP, P_ANISU, P_ANISC1, P_ANISC2, P_BEXT,
 P_XI, P_POL, P_LAMBDA, P_EPSPRIME, P_FREELAYER, P_J, P_FIXEDLAYER) = range(19)
SOLVER_HEUN, SOLVER_RK23, SOLVER_RK4, SOLVER_RK45DP = 2, 3, 4, 5
NREGION = 256
SYMLUT = NREGION * (NREGION + 1) // 2

# every symbol declared in include/umhip.h
EXPORTS = [
    "um_init", "um_last_error", "um_device_info", "um_set_stream", "um_sync", "um_malloc", "um_free",
    "um_memcpy_h2d", "um_memcpy_d2h", "um_memcpy_d2d", "um_zero", "um_memset",
    "um_add_exchange", "um_add_dmi", "um_add_uniaxial2", "um_add_cubic2", "um_region_addv", "um_zeromask",
    "um_lltorque", "um_llnoprecess", "um_thermal_field", "um_madd", "um_normalize", "um_vecnorm", "um_div",
    "um_mul", "um_reduce_sum", "um_reduce_dot", "um_reduce_maxabs", "um_reduce_maxvecnorm",
    "um_reduce_maxvecdiff", "um_demag_create", "um_demag_create_from_kernel", "um_demag_exec",
    "um_demag_destroy", "um_demag_info", "um_demag_kspec", "um_demag_kernel_bf",
    "um_demag_create_newell", "um_demag_kernel_newell", "um_sim_set_demag_newell",
    "um_sim_tape_begin", "um_sim_tape_end", "um_sim_tape_steps", "um_sim_backward",
    "um_add_dmi_bulk", "um_add_zhangli", "um_add_slonczewski", "um_sim_set_dmi_bulk", "um_sim_set_stt_flags",
    "um_sim_minimize", "um_set_kernel_cache", "um_kernel_cache_basename", "um_load_kernel", "um_save_kernel",
    "um_set_autotune", "um_demag_schedule", "um_sim_set_nodemag_lut", "um_sim_set_time_function",
    "um_sim_create", "um_sim_destroy", "um_sim_set_scalar", "um_sim_set_vector", "um_sim_set_regions",
    "um_sim_set_exchange_luts", "um_sim_set_bext_lut", "um_sim_set_frozen_lut", "um_sim_set_geometry",
    "um_sim_set_flags", "um_sim_set_demag_accuracy", "um_sim_set_thermal_seed", "um_sim_set_m",
    "um_sim_get_m", "um_sim_m_device", "um_sim_beff", "um_sim_torque", "um_sim_demag_field",
    "um_sim_set_solver", "um_sim_set_solver_params", "um_sim_get_state", "um_sim_step", "um_sim_steps",
    "um_sim_run", "um_sim_average_m", "um_sim_energy", "um_sim_stream", "um_sim_set_profiling",
    "um_sim_kernel_times", "um_sim_set_graph", "um_sims_steps", "um_sims_steps_batched", "um_graph_begin",
    "um_graph_end",
    "um_graph_launch", "um_graph_destroy", "um_sim_relax", "um_trace_kernels", "um_traced_kernels",
    "um_field_torque", "um_rk45dp_stage", "um_rk45dp_error", "um_rk45dp_error_noalloc",
    "um_sim_add_field_term", "um_sim_clear_field_terms",
]


class UmError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load libumhip.so (no GPU needed to load; calls need one)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise UmError(f"{path} is not built: run `python -m umagnus_amd.build` "
                          "(there is no CPU fallback)")
        lib = C.CDLL(path)
        lib.um_last_error.restype = C.c_char_p
        for name in EXPORTS:
            f = getattr(lib, name)
            if name != "um_last_error":
                f.restype = C.c_int
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise UmError(f"libumhip error {rc}: {load().um_last_error().decode()}")


def _call(name, *args):
    _check(getattr(load(), name)(*args))


_initialised = False


def init(device=0):
    """Init(gpu) and bind the library's unfused ops to torch's current stream."""
    global _initialised
    if _initialised:
        return
    import torch
    if not torch.cuda.is_available():
        raise UmError("no GPU visible: libumhip has no CPU fallback")
    torch.cuda.set_device(device)
    _call("um_init", C.c_int(device))
    _call("um_set_stream", C.c_void_p(torch.cuda.current_stream().cuda_stream))
    _initialised = True


def device_info():
    init()
    name = C.create_string_buffer(256)
    cu = C.c_int()
    mem = C.c_size_t()
    _call("um_device_info", name, 256, C.byref(cu), C.byref(mem))
    return name.value.decode(), cu.value, mem.value


def sync():
    _call("um_sync")


# ---- device-array helpers (torch tensors) ---------------------------------------------
def _prec(t):
    import torch
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    raise UmError(f"unsupported dtype {t.dtype}")


def _dp(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device arrays must be contiguous cuda tensors"
    return C.c_void_p(t.data_ptr())


def _vec(ts):
    return (C.c_void_p * 3)(*[t.data_ptr() for t in ts])


def _par(p):
    """MSlice convention: (tensor or None, multiplier)."""
    if p is None:
        return None, C.c_double(0.0)
    if isinstance(p, (int, float)):
        return None, C.c_double(float(p))
    return _dp(p), C.c_double(1.0)


# ---- opencl package mirror (opencl/*.go) ----------------------------------------------
def AddExchange(B, m, aLUT, Msat, regions, cellsize, pbc=0):
    """opencl.AddExchange (opencl/exchange.go:16); wx = float32(2/c^2)."""
    init()
    Nz, Ny, Nx = m[0].shape
    ms, msm = _par(Msat)
    w = [np.float32(2 / (c * c)) if _prec(m[0]) == F32 else 2 / (c * c) for c in cellsize]
    _call("um_add_exchange", _prec(m[0]), _vec(B), _vec(m), ms, msm, _dp(aLUT), _dp(regions),
          C.c_double(float(w[0])), C.c_double(float(w[1])), C.c_double(float(w[2])), Nx, Ny, Nz, pbc)


def AddDMI(B, m, aLUT, dLUT, Msat, regions, cellsize, pbc=0, OpenBC=False):
    """opencl.AddDMI (opencl/dmi.go:15)."""
    init()
    Nz, Ny, Nx = m[0].shape
    ms, msm = _par(Msat)
    _call("um_add_dmi", _prec(m[0]), _vec(B), _vec(m), ms, msm, _dp(aLUT), _dp(dLUT), _dp(regions),
          *[C.c_double(c) for c in cellsize], Nx, Ny, Nz, pbc, int(OpenBC))


def AddDMIBulk(B, m, aLUT, dLUT, Msat, regions, cellsize, pbc=0, OpenBC=False):
    """opencl.AddDMIBulk (opencl/dmibulk.go:14)."""
    init()
    Nz, Ny, Nx = m[0].shape
    ms, msm = _par(Msat)
    _call("um_add_dmi_bulk", _prec(m[0]), _vec(B), _vec(m), ms, msm, _dp(aLUT), _dp(dLUT), _dp(regions),
          *[C.c_double(c) for c in cellsize], Nx, Ny, Nz, pbc, int(OpenBC))


def AddZhangLiTorque(torque, m, Msat, J, alpha, xi, pol, cellsize, pbc=0):
    """opencl.AddZhangLiTorque (opencl/zhangli.go:12); J: three parameters."""
    init()
    Nz, Ny, Nx = m[0].shape
    js = [_par(j) for j in J]
    _call("um_add_zhangli", _prec(m[0]), _vec(torque), _vec(m), *_par(Msat),
          (C.c_void_p * 3)(*[j[0] for j in js]), (C.c_double * 3)(*[j[1].value for j in js]),
          *_par(alpha), *_par(xi), *_par(pol), *[C.c_double(c) for c in cellsize], Nx, Ny, Nz, pbc)


def AddSlonczewskiTorque2(torque, m, Msat, Jz, fixedP, alpha, pol, lam, eps_prime, thickness, mesh_thickness,
                          flp=1.0):
    """opencl.AddSlonczewskiTorque2 (opencl/slonczewski.go:12); fixedP: three parameters."""
    init()
    ps = [_par(q) for q in fixedP]
    _call("um_add_slonczewski", _prec(m[0]), _vec(torque), _vec(m), *_par(Msat), *_par(Jz),
          (C.c_void_p * 3)(*[q[0] for q in ps]), (C.c_double * 3)(*[q[1].value for q in ps]),
          *_par(alpha), *_par(pol), *_par(lam), *_par(eps_prime), *_par(thickness),
          C.c_double(mesh_thickness), C.c_double(flp), m[0].numel())


def AddUniaxialAnisotropy2(B, m, Msat, k1, k2, u):
    """opencl.AddUniaxialAnisotropy2 (opencl/anisotropy.go:153)."""
    init()
    ms, msm = _par(Msat)
    a1, m1 = _par(k1)
    a2, m2 = _par(k2)
    up = [_par(x) for x in u]
    uarr = (C.c_void_p * 3)(*[p[0].value if p[0] is not None else None for p in up])
    umul = (C.c_double * 3)(*[p[1].value for p in up])
    _call("um_add_uniaxial2", _prec(m[0]), _vec(B), _vec(m), ms, msm, a1, m1, a2, m2, uarr, umul, m[0].numel())


def AddCubicAnisotropy2(B, m, Msat, k1, k2, k3, c1, c2):
    """opencl.AddCubicAnisotropy2 (opencl/anisotropy.go:12)."""
    init()
    ms, msm = _par(Msat)
    ks = [_par(k) for k in (k1, k2, k3)]
    cs = []
    for c in (c1, c2):
        pp = [_par(x) for x in c]
        cs.append(((C.c_void_p * 3)(*[p[0].value if p[0] is not None else None for p in pp]),
                   (C.c_double * 3)(*[p[1].value for p in pp])))
    _call("um_add_cubic2", _prec(m[0]), _vec(B), _vec(m), ms, msm, ks[0][0], ks[0][1], ks[1][0], ks[1][1],
          ks[2][0], ks[2][1], cs[0][0], cs[0][1], cs[1][0], cs[1][1], m[0].numel())


def RegionAddV(dst, lut, regions):
    init()
    _call("um_region_addv", _prec(dst[0]), _vec(dst), _vec(lut), _dp(regions), dst[0].numel())


def ZeroMask(dst, maskLUT, regions):
    init()
    _call("um_zeromask", _prec(dst), _dp(dst), _dp(maskLUT), _dp(regions), dst.numel())


def LLTorque(torque, m, B, alpha):
    init()
    a, am = _par(alpha)
    _call("um_lltorque", _prec(m[0]), _vec(torque), _vec(m), _vec(B), a, am, m[0].numel())


def LLNoPrecess(torque, m, B):
    init()
    _call("um_llnoprecess", _prec(m[0]), _vec(torque), _vec(m), _vec(B), m[0].numel())


def ThermalField(B, seed, step, stream, k, Msat, Temp, Alpha):
    """SetTemperature (opencl/temperature.go:14) with in-place Philox normals."""
    init()
    ms, msm = _par(Msat)
    t, tm = _par(Temp)
    a, am = _par(Alpha)
    _call("um_thermal_field", _prec(B[0]), _vec(B), C.c_uint64(seed), C.c_uint64(step), C.c_uint32(stream),
          C.c_double(k), ms, msm, t, tm, a, am, B[0].numel())


def Madd(dst, srcs, facs):
    """Madd2..Madd7 (opencl/madd.go:111-489), one component array."""
    init()
    arr = (C.c_void_p * len(srcs))(*[s.data_ptr() for s in srcs])
    f = (C.c_double * len(facs))(*facs)
    _call("um_madd", _prec(dst), _dp(dst), len(srcs), arr, f, dst.numel())


def Normalize(v, vol=None):
    init()
    _call("um_normalize", _prec(v[0]), _vec(v), _dp(vol), v[0].numel())


def VecNorm(dst, a):
    init()
    _call("um_vecnorm", _prec(dst), _dp(dst), _vec(a), dst.numel())


def Div(dst, a, b):
    init()
    _call("um_div", _prec(dst), _dp(dst), _dp(a), _dp(b), dst.numel())


def Mul(dst, a, b):
    init()
    _call("um_mul", _prec(dst), _dp(dst), _dp(a), _dp(b), dst.numel())


def _red(name, prec, *args):
    out = C.c_double()
    _call(name, prec, *args, C.byref(out))
    return out.value


def Sum(x):
    init()
    return _red("um_reduce_sum", _prec(x), _dp(x), x.numel())


def Dot(a, b):
    init()
    return _red("um_reduce_dot", _prec(a), _dp(a), _dp(b), a.numel())


def MaxAbs(x):
    init()
    return _red("um_reduce_maxabs", _prec(x), _dp(x), x.numel())


def MaxVecNorm(v):
    init()
    return _red("um_reduce_maxvecnorm", _prec(v[0]), _vec(v), v[0].numel())


def MaxVecDiff(a, b, sqrt3_quirk=False):
    init()
    return _red("um_reduce_maxvecdiff", _prec(a[0]), _vec(a), _vec(b), a[0].numel(), int(sqrt3_quirk))


# ---- fused evaluation / RK stage / controller norms (um_field_torque, um_rk45dp_*) ----------
class um_param(C.Structure):
    _fields_ = [("arr", C.c_void_p), ("mul", C.c_double)]


class um_field_args(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("pbc", C.c_int),
                ("cellsize", C.c_double * 3), ("m", C.c_void_p * 3), ("bdemag", C.c_void_p * 3),
                ("torque", C.c_void_p * 3), ("beff", C.c_void_p * 3), ("regions", C.c_void_p),
                ("Msat", um_param), ("aex_lut", C.c_void_p), ("dmi_lut", C.c_void_p), ("dmi_bulk", C.c_int),
                ("openbc", C.c_int), ("Ku1", um_param), ("Ku2", um_param), ("anisU", um_param * 3),
                ("Kc1", um_param), ("Kc2", um_param), ("Kc3", um_param), ("anisC1", um_param * 3),
                ("anisC2", um_param * 3), ("bext_lut", C.c_void_p * 3), ("Bext", um_param * 3),
                ("thermal", C.c_int), ("seed", C.c_uint64), ("step", C.c_uint64), ("stream", C.c_uint32),
                ("Temp", um_param), ("thermal_k", C.c_double), ("badd", C.c_void_p * 3), ("alpha", um_param),
                ("precess", C.c_int), ("zhangli", C.c_int), ("slonczewski", C.c_int), ("J", um_param * 3),
                ("xi", um_param), ("pol", um_param), ("lambda_", um_param), ("eps_prime", um_param),
                ("free_layer_thickness", um_param), ("fixed_layer", um_param * 3),
                ("fixed_layer_top", C.c_double), ("tadd", C.c_void_p * 3), ("frozen_lut", C.c_void_p)]


def _up(v):
    """um_param from a float, None (zero) or a device tensor (per cell)."""
    arr, mul = _par(v)
    return um_param(arr.value if arr is not None else None, mul.value)


def _v3(ts):
    if ts is None:
        return (C.c_void_p * 3)()
    return (C.c_void_p * 3)(*[t.data_ptr() for t in ts])


def FieldTorque(m, cellsize, *, Msat, alpha, torque=None, beff=None, bdemag=None, regions=None, pbc=0,
                aex_lut=None, dmi_lut=None, dmi_bulk=False, openbc=False, Ku1=0.0, Ku2=0.0, anisU=(0, 0, 0),
                Kc1=0.0, Kc2=0.0, Kc3=0.0, anisC1=(0, 0, 0), anisC2=(0, 0, 0), bext_lut=None, Bext=(0, 0, 0),
                thermal=None, badd=None, precess=True, stt=None, tadd=None, frozen_lut=None):
    """SetTorque / SetEffectiveField in one pass (um_field_torque).  Device
    arrays are torch tensors; parameters are floats or per-cell tensors.
    thermal: (seed, step, stream, Temp, k) or None; stt: dict with J (3),
    xi, pol, lambda_, eps_prime, free_layer_thickness, fixed_layer (3),
    fixed_layer_top, zhangli, slonczewski."""
    init()
    Nz, Ny, Nx = m[0].shape
    a = um_field_args()
    a.Nx, a.Ny, a.Nz, a.pbc = Nx, Ny, Nz, int(pbc)
    a.cellsize = (C.c_double * 3)(*cellsize)
    a.m, a.bdemag, a.torque, a.beff = _v3(m), _v3(bdemag), _v3(torque), _v3(beff)
    a.regions = regions.data_ptr() if regions is not None else None
    a.Msat, a.alpha = _up(Msat), _up(alpha)
    a.aex_lut = aex_lut.data_ptr() if aex_lut is not None else None
    a.dmi_lut = dmi_lut.data_ptr() if dmi_lut is not None else None
    a.dmi_bulk, a.openbc = int(dmi_bulk), int(openbc)
    a.Ku1, a.Ku2, a.Kc1, a.Kc2, a.Kc3 = _up(Ku1), _up(Ku2), _up(Kc1), _up(Kc2), _up(Kc3)
    a.anisU = (um_param * 3)(*[_up(x) for x in anisU])
    a.anisC1 = (um_param * 3)(*[_up(x) for x in anisC1])
    a.anisC2 = (um_param * 3)(*[_up(x) for x in anisC2])
    a.bext_lut = _v3(bext_lut)
    a.Bext = (um_param * 3)(*[_up(x) for x in Bext])
    if thermal is not None:
        seed, step, stream, temp, k = thermal
        a.thermal, a.seed, a.step, a.stream, a.Temp, a.thermal_k = 1, seed, step, stream, _up(temp), float(k)
    a.badd, a.tadd = _v3(badd), _v3(tadd)
    a.precess = int(precess)
    if stt is not None:
        a.zhangli, a.slonczewski = int(stt.get("zhangli", 1)), int(stt.get("slonczewski", 1))
        a.J = (um_param * 3)(*[_up(x) for x in stt["J"]])
        a.fixed_layer = (um_param * 3)(*[_up(x) for x in stt.get("fixed_layer", (0, 0, 0))])
        for key in ("xi", "pol", "lambda_", "eps_prime", "free_layer_thickness"):
            setattr(a, key, _up(stt.get(key, 0.0)))
        a.fixed_layer_top = float(stt.get("fixed_layer_top", 1.0))
    a.frozen_lut = frozen_lut.data_ptr() if frozen_lut is not None else None
    _call("um_field_torque", _prec(m[0]), C.byref(a))


def RK45DPStage(dst, srcs, facs, vol=None, normalize=True):
    """MaddN(dst, srcs...) then Normalize(dst, vol) in one pass (um_rk45dp_stage)."""
    init()
    flat = (C.c_void_p * (3 * len(srcs)))(*[t.data_ptr() for v in srcs for t in v])
    _call("um_rk45dp_stage", _prec(dst[0]), _vec(dst), len(srcs), flat, (C.c_double * len(facs))(*facs),
          _dp(vol), int(normalize), dst[0].numel())


def RK45DPError(ks, e, h):
    """integralController's (err, rlerr, maxdm) for Err = sum e_j k_j, delM = ks[-1] (um_rk45dp_error)."""
    init()
    flat = (C.c_void_p * (3 * len(ks)))(*[t.data_ptr() for v in ks for t in v])
    out = (C.c_double * 3)()
    _call("um_rk45dp_error", _prec(ks[0][0]), len(ks), flat, (C.c_double * len(e))(*e), C.c_double(h),
          ks[0][0].numel(), out)
    return tuple(out)


class DemagConvolution:
    """opencl.DemagConvolution (opencl/conv_demag.go:14-23): NewDemag + Exec."""

    def __init__(self, N, cellsize, pbc=(0, 0, 0), accuracy=6.0, precision=F32, K=None, newell=False,
                 asymptotic_radius=32, zero_self_demag=0):
        init()
        self.h = C.c_void_p()
        self.N = tuple(N)
        self.precision = precision
        if K is None and newell:
            _call("um_demag_create_newell", precision, N[0], N[1], N[2], (C.c_int * 3)(*pbc),
                  (C.c_double * 3)(*cellsize), int(asymptotic_radius), int(zero_self_demag), C.byref(self.h))
        elif K is None:
            _call("um_demag_create", precision, N[0], N[1], N[2], (C.c_int * 3)(*pbc),
                  (C.c_double * 3)(*cellsize), C.c_double(accuracy), C.byref(self.h))
        else:
            P = (K[0].shape[2], K[0].shape[1], K[0].shape[0])
            ks = [np.ascontiguousarray(k, np.float64) for k in K]
            arr = (C.c_void_p * 6)(*[k.ctypes.data for k in ks])
            _call("um_demag_create_from_kernel", precision, N[0], N[1], N[2], (C.c_int * 3)(*P), arr,
                  C.byref(self.h))

    def Exec(self, B, m, vol, Msat):
        ms, msm = _par(Msat)
        _call("um_demag_exec", self.h, _vec(B), _vec(m), _dp(vol), ms, msm)

    def padded_size(self):
        P = (C.c_int * 3)()
        _call("um_demag_info", self.h, P)
        return tuple(P)

    def kspec(self, c):
        P = self.padded_size()
        shape = (P[2] // 2 + 1, P[1] // 2 + 1, P[0] // 2 + 1)
        out = np.zeros(shape, np.float32 if self.precision == F32 else np.float64)
        _call("um_demag_kspec", self.h, c, out.ctypes.data_as(C.c_void_p))
        return out

    def schedule(self):
        """Launch schedule: nl_x, nl_xi, four_step, nl_y, B_conv, sched_x, sched_y, sched_z."""
        out = (C.c_int * 8)()
        _call("um_demag_schedule", self.h, out)
        return list(out)

    def Free(self):
        if self.h:
            _call("um_demag_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.Free()
        except Exception:
            pass


def _padded(N, pbc):
    P = []
    for i in range(3):
        if pbc[i]:
            P.append(N[i])
        elif i != 2 or N[i] > 5:
            P.append(2 * N[i])
        else:
            P.append(2 * N[i] - 1)
    K = [np.zeros((P[2], P[1], P[0]), np.float64) for _ in range(6)]
    return K, (C.c_void_p * 6)(*[k.ctypes.data for k in K])


def demag_kernel_bf(N, cellsize, pbc=(0, 0, 0), accuracy=6.0):
    """CalcDemagKernel (mag/demagkernel.go:115) on the GPU in FP64 -> 6 host arrays."""
    init()
    K, arr = _padded(N, pbc)
    _call("um_demag_kernel_bf", N[0], N[1], N[2], (C.c_int * 3)(*pbc), (C.c_double * 3)(*cellsize),
          C.c_double(accuracy), arr)
    return K


def demag_kernel_newell(N, cellsize, pbc=(0, 0, 0), asymptotic_radius=32, zero_self_demag=0):
    """CalcNewellDemagKernel (mag/newellDemagKernel.go:1108) on the GPU in FP64 -> 6 host arrays."""
    init()
    K, arr = _padded(N, pbc)
    _call("um_demag_kernel_newell", N[0], N[1], N[2], (C.c_int * 3)(*pbc), (C.c_double * 3)(*cellsize),
          int(asymptotic_radius), int(zero_self_demag), arr)
    return K


class KernelTrace:
    """`with KernelTrace() as t: ...` records the kernels the library launches
    inside the block (um_trace_kernels); t.names lists their demangled names."""

    def __enter__(self):
        _call("um_trace_kernels", 1)
        self.names = []
        return self

    def __exit__(self, et, ev, tb):
        need = C.c_size_t()
        _call("um_traced_kernels", None, C.c_size_t(0), C.byref(need))
        buf = C.create_string_buffer(need.value)
        _call("um_traced_kernels", buf, C.c_size_t(need.value), None)
        _call("um_trace_kernels", 0)
        self.names = [l for l in buf.value.decode().split("\n") if l]
        return False


def set_autotune(mode, path=None):
    """Launch-schedule autotuning of new demag plans: 0 static tables, 1 time
    unknown shapes, 2 known results only; `path`: per-shape results file."""
    _call("um_set_autotune", int(mode), path.encode() if path else None)


def set_kernel_cache(directory):
    """The -cache flag (engine/gofiles.go:17): demag plans read and write their
    tensor as OVF files in `directory` (None or "" disables caching)."""
    _call("um_set_kernel_cache", (directory or "").encode())


def kernel_cache_basename(directory, N, cellsize, pbc=(0, 0, 0), newell=False, accuracy=6.0, zero_self_demag=0):
    """File prefix of a cache entry (mag/demagkernel.go:40, newellDemagKernel.go:1023)."""
    buf = C.create_string_buffer(4096)
    _call("um_kernel_cache_basename", directory.encode(), (C.c_int * 3)(*N), (C.c_int * 3)(*pbc),
          (C.c_double * 3)(*cellsize), int(newell), C.c_double(accuracy), int(zero_self_demag), buf, 4096)
    return buf.value.decode()


def load_kernel(fname, P):
    """LoadKernel (mag/demagkernel.go:97): scalar OVF field of size P (x, y, z) -> (Pz, Py, Px) float64."""
    K = np.zeros((P[2], P[1], P[0]), np.float64)
    _call("um_load_kernel", str(fname).encode(), (C.c_int * 3)(*P), K.ctypes.data_as(C.c_void_p))
    return K


def save_kernel(fname, K, cellsize, name):
    """SaveKernel (mag/demagkernel.go:102): (Pz, Py, Px) array as an OVF2 binary 4 scalar field."""
    K = np.ascontiguousarray(K, np.float64)
    P = (K.shape[2], K.shape[1], K.shape[0])
    _call("um_save_kernel", str(fname).encode(), (C.c_int * 3)(*P), (C.c_double * 3)(*cellsize),
          name.encode(), K.ctypes.data_as(C.c_void_p))


# ---- simulation handle (engine/) -------------------------------------------------------
class Sim:
    """Engine-level simulation on one GPU stream; host numpy arrays in and out."""

    def __init__(self, N, cellsize, pbc=(0, 0, 0), dtype=np.float32):
        load()
        init()
        self.dtype = np.dtype(dtype)
        self.prec = F32 if self.dtype == np.float32 else F64
        self.N = tuple(N)
        self.n = N[0] * N[1] * N[2]
        self.shape = (N[2], N[1], N[0])
        self.h = C.c_void_p()
        _call("um_sim_create", self.prec, N[0], N[1], N[2], C.c_double(cellsize[0]), C.c_double(cellsize[1]),
              C.c_double(cellsize[2]), pbc[0], pbc[1], pbc[2], C.byref(self.h))

    def __del__(self):
        try:
            if self.h:
                _call("um_sim_destroy", self.h)
                self.h = None
        except Exception:
            pass

    def _a(self, v):
        return None if v is None else np.ascontiguousarray(v, self.dtype)

    @staticmethod
    def _p(a):
        return None if a is None else a.ctypes.data_as(C.c_void_p)

    def set_scalar(self, pid, value):
        if value is None or np.isscalar(value):
            _call("um_sim_set_scalar", self.h, pid, None, C.c_double(0.0 if value is None else float(value)))
        else:
            a = self._a(value)
            _call("um_sim_set_scalar", self.h, pid, self._p(a), C.c_double(1.0))

    def set_vector(self, pid, value):
        if value is None:
            value = (0.0, 0.0, 0.0)
        if all(np.isscalar(v) for v in value):
            _call("um_sim_set_vector", self.h, pid, None, None, None, *[C.c_double(float(v)) for v in value])
        else:
            a = [self._a(v) for v in value]
            _call("um_sim_set_vector", self.h, pid, self._p(a[0]), self._p(a[1]), self._p(a[2]),
                  C.c_double(1), C.c_double(1), C.c_double(1))

    def set_regions(self, regions):
        r = None if regions is None else np.ascontiguousarray(regions, np.uint8)
        _call("um_sim_set_regions", self.h, self._p(r))

    def set_exchange_luts(self, aLUT, dLUT=None):
        a, d = self._a(aLUT), self._a(dLUT)
        _call("um_sim_set_exchange_luts", self.h, self._p(a), self._p(d))

    def set_bext_lut(self, lut3):
        if lut3 is None:
            _call("um_sim_set_bext_lut", self.h, None, None, None)
            return
        ls = [self._a(l) for l in lut3]
        _call("um_sim_set_bext_lut", self.h, self._p(ls[0]), self._p(ls[1]), self._p(ls[2]))

    def set_frozen_lut(self, lut):
        _call("um_sim_set_frozen_lut", self.h, self._p(self._a(lut)))

    _TFN = C.CFUNCTYPE(None, C.c_double, C.POINTER(C.c_double), C.c_void_p)

    def set_time_function(self, pid, fn, ncomp=1):
        """Parameter pid as a function of time: fn(t) -> value (ncomp 1) or 3 values,
        evaluated at every stage; None removes it."""
        if not hasattr(self, "_tfn"):
            self._tfn = {}
        if fn is None:
            self._tfn.pop(pid, None)
            _call("um_sim_set_time_function", self.h, int(pid), int(ncomp), None, None)
            return

        def cb(t, out, ctx):
            v = fn(t)
            if ncomp == 1:
                out[0] = float(v)
            else:
                for c in range(3):
                    out[c] = float(v[c])
        self._tfn[pid] = Sim._TFN(cb)
        _call("um_sim_set_time_function", self.h, int(pid), int(ncomp), self._tfn[pid], None)

    FIELD_TERM_BEXT, FIELD_TERM_CUSTOM, TORQUE_TERM_CUSTOM = 0, 1, 2

    def add_field_term(self, kind, arr3, mul=1.0, fn=None):
        """Extra per-cell term (um_sim_add_field_term): kind 0 B_ext mask term
        (Excitation.Add), 1 custom field (AddFieldTerm), 2 custom torque
        (AddTorqueTerm); arr3 = three (Nz, Ny, Nx) arrays; fn(t) -> multiplier."""
        a = [self._a(x) for x in arr3]
        if not hasattr(self, "_xfn"):
            self._xfn = []
        cb = None
        if fn is not None:
            def _cb(t, out, ctx):
                out[0] = float(fn(t))
            cb = Sim._TFN(_cb)
            self._xfn.append(cb)
        _call("um_sim_add_field_term", self.h, int(kind), self._p(a[0]), self._p(a[1]), self._p(a[2]),
              C.c_double(mul), cb, None)

    def clear_field_terms(self, kind=-1):
        _call("um_sim_clear_field_terms", self.h, int(kind))

    def set_nodemag_lut(self, lut):
        """NoDemagSpins per region (256 values); None clears."""
        _call("um_sim_set_nodemag_lut", self.h, self._p(self._a(lut)) if lut is not None else None)

    def set_geometry(self, vol):
        _call("um_sim_set_geometry", self.h, self._p(self._a(vol)))

    def set_flags(self, enable_demag=1, openbc=0, precess=1, heun_sqrt3=1):
        _call("um_sim_set_flags", self.h, int(enable_demag), int(openbc), int(precess), int(heun_sqrt3))

    def set_demag_accuracy(self, acc):
        _call("um_sim_set_demag_accuracy", self.h, C.c_double(acc))

    def set_dmi_bulk(self, lut):
        """Dbulk LUT (engine/exchange.go:49-66); None: none."""
        if lut is None:
            _call("um_sim_set_dmi_bulk", self.h, None)
        else:
            lut = np.ascontiguousarray(lut, self.dtype)
            _call("um_sim_set_dmi_bulk", self.h, lut.ctypes.data_as(C.c_void_p))

    def set_stt_flags(self, disable_zhangli=0, disable_slonczewski=0, fixed_layer_top=True):
        _call("um_sim_set_stt_flags", self.h, int(disable_zhangli), int(disable_slonczewski),
              C.c_double(1.0 if fixed_layer_top else -1.0))

    def set_demag_newell(self, enable=1, asymptotic_radius=32, zero_self_demag=0):
        """EnableNewellDemag = enable (engine/demag.go:26-56)."""
        _call("um_sim_set_demag_newell", self.h, int(enable), int(asymptotic_radius), int(zero_self_demag))

    def set_thermal_seed(self, seed, stream=0):
        _call("um_sim_set_thermal_seed", self.h, C.c_uint64(seed), C.c_uint32(stream))

    def set_m(self, m):
        m = np.ascontiguousarray(m, self.dtype).reshape(3, -1)
        _call("um_sim_set_m", self.h, self._p(m[0]), self._p(m[1]), self._p(m[2]))

    def _get3(self, name):
        out = np.zeros((3,) + self.shape, self.dtype)
        _call(name, self.h, self._p(out[0]), self._p(out[1]), self._p(out[2]))
        return out

    def get_m(self):
        return self._get3("um_sim_get_m")

    def beff(self):
        return self._get3("um_sim_beff")

    def torque(self):
        return self._get3("um_sim_torque")

    def demag_field(self):
        return self._get3("um_sim_demag_field")

    def set_solver(self, solver):
        _call("um_sim_set_solver", self.h, int(solver))

    def set_solver_params(self, dt=1e-15, fixdt=0.0, maxerr=1e-5, relerr=1e-4, minslope=1e8, headroom=0.8,
                          mindt=0.0, maxdt=0.0, gammaLL=1.7595e11):
        k = (C.c_double * 9)(dt, fixdt, maxerr, relerr, minslope, headroom, mindt, maxdt, gammaLL)
        _call("um_sim_set_solver_params", self.h, k)

    def state(self):
        o = (C.c_double * 8)()
        _call("um_sim_get_state", self.h, o)
        keys = ("t", "dt", "nsteps", "nundone", "nevals", "lasterr", "peakerr", "lasttorque")
        return dict(zip(keys, list(o)))

    def step(self):
        _call("um_sim_step", self.h)

    def steps(self, n):
        _call("um_sim_steps", self.h, int(n))

    def run(self, seconds):
        _call("um_sim_run", self.h, C.c_double(seconds))

    def average_m(self):
        o = (C.c_double * 3)()
        _call("um_sim_average_m", self.h, o)
        return np.array(list(o))

    def energy(self, term):
        o = C.c_double()
        _call("um_sim_energy", self.h, int(term), C.byref(o))
        return o.value

    def stream(self):
        s = C.c_void_p()
        _call("um_sim_stream", self.h, C.byref(s))
        return s.value

    def m_device(self):
        p = (C.c_void_p * 3)()
        _call("um_sim_m_device", self.h, p)
        return [p[i] for i in range(3)]

    def set_profiling(self, on):
        _call("um_sim_set_profiling", self.h, int(on))

    def kernel_times(self):
        ms = (C.c_double * 3)()
        ne = C.c_int()
        _call("um_sim_kernel_times", self.h, ms, C.byref(ne))
        return list(ms), ne.value

    def minimize(self, stop=1e-6, samples=10):
        """Minimize(), engine/minimizer.go:120-160 (um_sim_minimize)."""
        _call("um_sim_minimize", self.h, C.c_double(stop), int(samples))

    def relax(self, torque_threshold=-1.0):
        """Relax(), engine/relax.go:24-100 (um_sim_relax)."""
        _call("um_sim_relax", self.h, C.c_double(torque_threshold))

    # ---- discrete RK4 adjoint ------------------------------------------------------
    def tape_begin(self, max_steps, checkpoint_every=0):
        """Record the following fixed-step RK4 steps for backward() (um_sim_tape_begin)."""
        _call("um_sim_tape_begin", self.h, int(max_steps), int(checkpoint_every))

    def tape_end(self):
        _call("um_sim_tape_end", self.h)

    def tape_steps(self):
        o = C.c_longlong()
        _call("um_sim_tape_steps", self.h, C.byref(o))
        return o.value

    def backward(self, dm):
        """Reverse sweep from dL/dm_final (host array, shape (3, Nz, Ny, Nx)).
        Returns (dL/dm0, dL/dMsat per cell, dL/dB_ext uniform) as host arrays."""
        n = int(np.prod(self.N))
        it = self.dtype.itemsize
        dm = np.ascontiguousarray(dm, self.dtype).reshape(3, n)
        bufs = []
        try:
            for _ in range(7):
                p = C.c_void_p()
                _call("um_malloc", C.byref(p), C.c_size_t(n * it))
                bufs.append(p)
            for c in range(3):
                _call("um_memcpy_h2d", bufs[c], dm[c].ctypes.data_as(C.c_void_p), C.c_size_t(n * it))
            gb = (C.c_double * 3)()
            _call("um_sim_backward", self.h, (C.c_void_p * 3)(*bufs[:3]), (C.c_void_p * 3)(*bufs[3:6]),
                  bufs[6], gb)
            out = np.zeros((4, n), self.dtype)
            for c in range(4):
                _call("um_memcpy_d2h", out[c].ctypes.data_as(C.c_void_p), bufs[3 + c], C.c_size_t(n * it))
        finally:
            for p in bufs:
                _call("um_free", p)
        shape = (self.N[2], self.N[1], self.N[0])
        return out[:3].reshape((3,) + shape), out[3].reshape(shape), np.array(list(gb))

    def set_graph(self, on):
        """Replay fixed-dt steps as a captured HIP graph (um_sim_set_graph)."""
        _call("um_sim_set_graph", self.h, int(on))


def sims_steps(sims, n):
    """n steps of every simulation in `sims`, graph replays interleaved across
    their streams (um_sims_steps)."""
    arr = (C.c_void_p * len(sims))(*[s.h for s in sims])
    _call("um_sims_steps", len(sims), arr, int(n))


def sims_steps_batched(sims, n):
    """n steps of every simulation in `sims` in batched launches (one launch per
    kernel variant and phase for the whole ensemble, um_sims_steps_batched);
    returns the launches that replay two steps of the ensemble (0: the members
    were replayed one graph each)."""
    arr = (C.c_void_p * len(sims))(*[s.h for s in sims])
    nl = C.c_int(0)
    _call("um_sims_steps_batched", len(sims), arr, int(n), C.byref(nl))
    return nl.value


class Graph:
    """Capture of the library stream: `with Graph() as g: <um_* calls>`, then
    g.launch() replays them (um_graph_begin/end/launch/destroy)."""

    def __init__(self):
        self.h = None

    def __enter__(self):
        _call("um_graph_begin")
        return self

    def __exit__(self, et, ev, tb):
        h = C.c_void_p()
        rc = load().um_graph_end(C.byref(h))
        if et is None and rc != 0:
            raise UmError(f"libumhip error {rc}: {load().um_last_error().decode()}")
        self.h = h.value
        return False

    def launch(self):
        _call("um_graph_launch", C.c_void_p(self.h))

    def __del__(self):
        if self.h is not None and _lib is not None:
            _lib.um_graph_destroy(C.c_void_p(self.h))
            self.h = None
