"""ctypes wrapper of the CPU oracle (oracle/libumoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libumoracle.so")
_lib = None

NREGION = 256
SYMLUT = NREGION * (NREGION + 1) // 2

(P_MSAT, P_KU1, P_KU2, P_KC1, P_KC2, P_KC3, P_ALPHA, P_TEMP, P_ANISU, P_ANISC1, P_ANISC2, P_BEXT,
 P_XI, P_POL, P_LAMBDA, P_EPSPRIME, P_FREELAYER, P_J, P_FIXEDLAYER) = range(19)
VECTOR_PARAMS = (P_ANISU, P_ANISC1, P_ANISC2, P_BEXT, P_J, P_FIXEDLAYER)
SOLVER_HEUN, SOLVER_RK23, SOLVER_RK4, SOLVER_RK45DP = 2, 3, 4, 5


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.um_or_maxvecnorm_f32.restype = C.c_double
        for sfx in ("_f32", "_f64"):
            for fn in ("maxvecnorm", "maxvecdiff", "maxabs", "sum", "dot", "sim_energy"):
                getattr(_lib, "um_or_" + fn + sfx).restype = C.c_double
            for fn in ("demag_new", "demag_new_from_kernel", "sim_new"):
                getattr(_lib, "um_or_" + fn + sfx).restype = C.c_void_p
        _lib.um_or_exch_average.restype = C.c_float
        _lib.um_or_exch_average.argtypes = [C.c_float, C.c_float]
        _lib.um_or_adapt_dt.restype = C.c_double
        _lib.um_or_adapt_dt.argtypes = [C.c_double] * 8
    return _lib


def _sfx(dtype):
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def _rt(dtype):
    return C.c_float if np.dtype(dtype) == np.float32 else C.c_double


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def fn(name, dtype):
    return getattr(lib(), "um_or_" + name + _sfx(dtype))


# ---- precision independent --------------------------------------------------
def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().um_or_philox4x32_10(c, k, o)
    return list(o)


def pad_size(N, pbc=(0, 0, 0)):
    n = (C.c_int * 3)(*N)
    p = (C.c_int * 3)(*pbc)
    P = (C.c_int * 3)()
    lib().um_or_pad_size(n, p, P)
    return tuple(P)


def demag_kernel_bf(N, cellsize, pbc=(0, 0, 0), accuracy=6.0):
    """Six real-space FP64 tensor components (xx,yy,zz,yz,xz,xy), shape (Pz,Py,Px)."""
    P = pad_size(N, pbc)
    K = [np.zeros((P[2], P[1], P[0]), np.float64) for _ in range(6)]
    arr = (C.c_void_p * 6)(*[ptr(k) for k in K])
    rc = lib().um_or_demag_kernel_bf((C.c_int * 3)(*N), (C.c_int * 3)(*pbc),
                                     (C.c_double * 3)(*cellsize), C.c_double(accuracy), arr)
    assert rc == 0
    return K


def demag_kernel_newell(N, cellsize, pbc=(0, 0, 0), asymptotic_radius=32, zero_self_demag=0):
    """Newell tensor (mag/newellDemagKernel.go:1108), same layout as demag_kernel_bf."""
    P = pad_size(N, pbc)
    K = [np.zeros((P[2], P[1], P[0]), np.float64) for _ in range(6)]
    arr = (C.c_void_p * 6)(*[ptr(k) for k in K])
    rc = lib().um_or_demag_kernel_newell((C.c_int * 3)(*N), (C.c_int * 3)(*pbc),
                                         (C.c_double * 3)(*cellsize), int(asymptotic_radius),
                                         int(zero_self_demag), arr)
    assert rc == 0, rc
    return K


def _d(name, *args):
    f = getattr(lib(), name)
    f.restype = C.c_double
    return f(*[C.c_double(a) for a in args])


def newell_f(x, y, z):
    return _d("um_or_newell_f", x, y, z)


def newell_g(x, y, z):
    return _d("um_or_newell_g", x, y, z)


def self_demag_nx(x, y, z):
    return _d("um_or_self_demag_nx", x, y, z)


def newell_cell(cellsize):
    d = (C.c_double * 3)()
    lib().um_or_newell_cell((C.c_double * 3)(*cellsize), d)
    return tuple(d)


def newell_element(comp, d, R, far=False):
    """Kernel element -N (comp 0..5 = xx yy zz yz xz xy) at offset R (normalised
    cell units d): exact 27-point form, or the far-field expansion."""
    f = lib().um_or_newell_asymptotic if far else lib().um_or_newell_exact
    f.restype = C.c_double
    v = f(int(comp), (C.c_double * 3)(*d), *[C.c_double(r) for r in R])
    return -v if far else v


def symmidx(i, j):
    return lib().um_or_symmidx(i, j)


def exch_average(a, b):
    return lib().um_or_exch_average(a, b)


def adapt_dt(dt, corr, fixdt, headroom, mindt, maxdt, time, alarm):
    return lib().um_or_adapt_dt(dt, corr, fixdt, headroom, mindt, maxdt, time, alarm)


# ---- FFT ----------------------------------------------------------------------
def fft(x, sign=-1):
    """In-place-style 1D complex FFT (unnormalised) of a complex array."""
    dt = np.float32 if x.dtype == np.complex64 else np.float64
    buf = np.ascontiguousarray(x).view(dt).copy()
    fn("fft", dt)(ptr(buf), C.c_int(x.size), C.c_int(sign))
    return buf.view(x.dtype)


def fft3d(x, sign=-1):
    dt = np.float32 if x.dtype == np.complex64 else np.float64
    buf = np.ascontiguousarray(x).view(dt).copy()
    P = (C.c_int * 3)(x.shape[2], x.shape[1], x.shape[0])
    fn("fft3d", dt)(ptr(buf), P, C.c_int(sign))
    return buf.view(x.dtype).reshape(x.shape)


# ---- demag ----------------------------------------------------------------------
class Demag:
    def __init__(self, N, cellsize, pbc=(0, 0, 0), accuracy=6.0, dtype=np.float64, K=None):
        self.dtype = np.dtype(dtype)
        self.N = tuple(N)
        if K is None:
            self.h = fn("demag_new", dtype)((C.c_int * 3)(*N), (C.c_int * 3)(*pbc),
                                            (C.c_double * 3)(*cellsize), C.c_double(accuracy), 0)
        else:
            P = (K[0].shape[2], K[0].shape[1], K[0].shape[0])
            arr = (C.c_void_p * 6)(*[ptr(np.ascontiguousarray(k, np.float64)) for k in K])
            self.h = fn("demag_new_from_kernel", dtype)((C.c_int * 3)(*N), (C.c_int * 3)(*P), arr)
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            fn("demag_free", self.dtype)(C.c_void_p(self.h))
            self.h = None

    def exec(self, m, Ms, vol=None):
        """m: (3,Nz,Ny,Nx). Ms: float or array (Nz,Ny,Nx). Returns B (3,Nz,Ny,Nx)."""
        m = np.ascontiguousarray(m, self.dtype)
        B = np.zeros_like(m)
        ms_arr, ms_mul = _param(Ms, self.dtype)
        volp = None if vol is None else np.ascontiguousarray(vol, self.dtype)
        fn("demag_exec", self.dtype)(C.c_void_p(self.h), ptr(B[0]), ptr(B[1]), ptr(B[2]),
                                     ptr(m[0]), ptr(m[1]), ptr(m[2]), ptr(volp), ptr(ms_arr),
                                     _rt(self.dtype)(ms_mul))
        return B


def demag_brute(N, K, m, Ms, vol=None, dtype=np.float64):
    P = (K[0].shape[2], K[0].shape[1], K[0].shape[0])
    m = np.ascontiguousarray(m, dtype)
    B = np.zeros_like(m)
    ms_arr, ms_mul = _param(Ms, dtype)
    arr = (C.c_void_p * 6)(*[ptr(np.ascontiguousarray(k, np.float64)) for k in K])
    volp = None if vol is None else np.ascontiguousarray(vol, dtype)
    fn("demag_brute", dtype)((C.c_int * 3)(*N), (C.c_int * 3)(*P), arr, ptr(B[0]), ptr(B[1]),
                             ptr(B[2]), ptr(m[0]), ptr(m[1]), ptr(m[2]), ptr(volp), ptr(ms_arr),
                             _rt(dtype)(ms_mul))
    return B


def _param(v, dtype):
    """(array or None, multiplier) in the reference's MSlice convention."""
    if v is None:
        return None, 0.0
    if np.isscalar(v):
        return None, float(v)
    return np.ascontiguousarray(v, dtype), 1.0


def thermal_normals(N, seed, step, stream=0, dtype=np.float64):
    out = np.zeros((3, N), dtype)
    fn("thermal_normals", dtype)(ptr(out[0]), ptr(out[1]), ptr(out[2]), C.c_int(N),
                                 C.c_uint64(seed), C.c_uint64(step), C.c_uint32(stream))
    return out


# ---- simulation ---------------------------------------------------------------------
class Sim:
    """Engine-level restatement: parameters, effective field, torque, steppers."""

    def __init__(self, N, cellsize, pbc=(0, 0, 0), dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.N = tuple(N)
        self.n = N[0] * N[1] * N[2]
        self.shape = (N[2], N[1], N[0])
        self.h = fn("sim_new", dtype)(N[0], N[1], N[2], C.c_double(cellsize[0]),
                                      C.c_double(cellsize[1]), C.c_double(cellsize[2]),
                                      pbc[0], pbc[1], pbc[2])
        self._f = lambda name: fn(name, self.dtype)

    def __del__(self):
        if getattr(self, "h", None):
            fn("sim_free", self.dtype)(C.c_void_p(self.h))
            self.h = None

    def _h(self):
        return C.c_void_p(self.h)

    def set_scalar(self, pid, value):
        arr, mul = _param(value, self.dtype)
        self._f("sim_set_scalar")(self._h(), pid, ptr(arr), C.c_double(mul))

    def set_vector(self, pid, value):
        if value is None:
            value = (0.0, 0.0, 0.0)
        if isinstance(value, (tuple, list)) and all(np.isscalar(v) for v in value):
            self._f("sim_set_vector")(self._h(), pid, None, None, None, *[C.c_double(v) for v in value])
        else:
            a = [np.ascontiguousarray(v, self.dtype) for v in value]
            self._f("sim_set_vector")(self._h(), pid, ptr(a[0]), ptr(a[1]), ptr(a[2]),
                                      C.c_double(1), C.c_double(1), C.c_double(1))

    def set_regions(self, regions):
        r = None if regions is None else np.ascontiguousarray(regions, np.uint8)
        self._f("sim_set_regions")(self._h(), ptr(r))

    def set_exchange_luts(self, aLUT, dLUT=None):
        a = np.ascontiguousarray(aLUT, self.dtype)
        d = None if dLUT is None else np.ascontiguousarray(dLUT, self.dtype)
        self._f("sim_set_exchange_luts")(self._h(), ptr(a), ptr(d))

    def set_bext_lut(self, lut3):
        ls = [np.ascontiguousarray(l, self.dtype) for l in lut3]
        self._f("sim_set_bext_lut")(self._h(), ptr(ls[0]), ptr(ls[1]), ptr(ls[2]))

    def set_nodemag_lut(self, lut):
        """NoDemagSpins per region (engine/demag.go:22-23, 67-98); None clears."""
        self._f("sim_set_nodemag_lut")(self._h(), ptr(np.ascontiguousarray(lut, self.dtype)) if lut is not None
                                       else None)

    def set_frozen_lut(self, lut):
        self._f("sim_set_frozen_lut")(self._h(), ptr(np.ascontiguousarray(lut, self.dtype)))

    def set_geometry(self, vol):
        v = None if vol is None else np.ascontiguousarray(vol, self.dtype)
        self._f("sim_set_geometry")(self._h(), ptr(v))

    def set_flags(self, enable_demag=1, openbc=0, precess=1, heun_sqrt3=1):
        self._f("sim_set_flags")(self._h(), int(enable_demag), int(openbc), int(precess), int(heun_sqrt3))

    def set_demag_accuracy(self, acc):
        self._f("sim_set_demag_accuracy")(self._h(), C.c_double(acc))

    def set_dmi_bulk(self, lut):
        """Dbulk LUT (engine/exchange.go:49-66); None: none."""
        v = None if lut is None else np.ascontiguousarray(lut, self.dtype)
        self._f("sim_set_dmi_bulk")(self._h(), ptr(v))

    def set_stt_flags(self, disable_zhangli=0, disable_slonczewski=0, fixed_layer_top=True):
        self._f("sim_set_stt_flags")(self._h(), int(disable_zhangli), int(disable_slonczewski),
                                     C.c_double(1.0 if fixed_layer_top else -1.0))

    def set_demag_newell(self, enable=1, asymptotic_radius=32, zero_self_demag=0):
        self._f("sim_set_demag_newell")(self._h(), int(enable), int(asymptotic_radius), int(zero_self_demag))

    def set_thermal_seed(self, seed, stream=0):
        self._f("sim_set_thermal_seed")(self._h(), C.c_uint64(seed), C.c_uint32(stream))

    def set_m(self, m):
        m = np.ascontiguousarray(m, self.dtype).reshape(3, -1)
        self._f("sim_set_m")(self._h(), ptr(m[0]), ptr(m[1]), ptr(m[2]))

    def _get3(self, name):
        out = np.zeros((3,) + self.shape, self.dtype)
        self._f(name)(self._h(), ptr(out[0]), ptr(out[1]), ptr(out[2]))
        return out

    def get_m(self):
        return self._get3("sim_get_m")

    def beff(self):
        return self._get3("sim_beff")

    def torque(self):
        return self._get3("sim_torque")

    def demag_field(self):
        return self._get3("sim_demag_field")

    def set_solver(self, solver):
        self._f("sim_set_solver")(self._h(), int(solver))

    def set_solver_params(self, dt=1e-15, fixdt=0.0, maxerr=1e-5, relerr=1e-4, minslope=1e8,
                          headroom=0.8, mindt=0.0, maxdt=0.0, gammaLL=1.7595e11):
        k = (C.c_double * 9)(dt, fixdt, maxerr, relerr, minslope, headroom, mindt, maxdt, gammaLL)
        self._f("sim_set_solver_params")(self._h(), k)

    def state(self):
        o = (C.c_double * 8)()
        self._f("sim_get_state")(self._h(), o)
        keys = ("t", "dt", "nsteps", "nundone", "nevals", "lasterr", "peakerr", "lasttorque")
        return dict(zip(keys, list(o)))

    def step(self):
        self._f("sim_step")(self._h())

    def steps(self, n):
        self._f("sim_steps")(self._h(), int(n))

    def run(self, seconds):
        self._f("sim_run")(self._h(), C.c_double(seconds))

    def minimize(self, stop=1e-6, samples=10):
        self._f("sim_minimize")(self._h(), C.c_double(stop), int(samples))

    def relax(self, torque_threshold=-1.0):
        """Relax(), engine/relax.go:24-100."""
        self._f("sim_relax")(self._h(), C.c_double(torque_threshold))

    def average_m(self):
        o = (C.c_double * 3)()
        self._f("sim_average_m")(self._h(), o)
        return np.array(list(o))

    def energy(self, term):
        return self._f("sim_energy")(self._h(), int(term))


def exchange_lut(values_by_region, dtype=np.float64):
    """Symmetric inter-region LUT: harmonic mean (engine/exchange.go:153-167)."""
    ex = np.zeros(NREGION, np.float32)
    for r, v in values_by_region.items():
        ex[r] = v
    lut = np.zeros(SYMLUT, dtype)
    for i in range(NREGION):
        for j in range(i, NREGION):
            lut[symmidx(i, j)] = exch_average(float(ex[i]), float(ex[j]))
    return lut
