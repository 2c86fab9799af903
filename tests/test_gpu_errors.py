"""Argument errors through the C ABI on the GPU.  The reference panics or
calls util.Fatal on bad sizes (opencl/fft3dr2c.go:35-41, util.Argument) and
rejects unknown solvers (engine/solver.go SetSolver); the library returns a
status code with a message instead, before any kernel is launched, and stays
usable afterwards."""
import ctypes as C
import sys

import numpy as np
import pytest

from cases import random_unit, setup

pytestmark = pytest.mark.gpu


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


def _raises(gpu, fn, *words):
    with pytest.raises(gpu.UmError) as e:
        fn()
    for w in words:
        assert w in str(e.value), (w, str(e.value))


def test_sim_creation_rejects(gpu):
    _say("sim creation errors")
    _raises(gpu, lambda: gpu.Sim((0, 4, 1), (1e-9,) * 3), "-1", "grid size")
    _raises(gpu, lambda: gpu.Sim((4, -2, 1), (1e-9,) * 3), "grid size")
    _raises(gpu, lambda: gpu.Sim((4, 4, 1), (1e-9, 0.0, 1e-9)), "cell size")
    _raises(gpu, lambda: gpu.Sim((4, 4, 1), (1e-9, float("nan"), 1e-9)), "cell size")
    _raises(gpu, lambda: gpu.Sim((4, 4, 1), (1e-9,) * 3, pbc=(-1, 0, 0)), "PBC")
    _raises(gpu, lambda: gpu.Sim((4, 4, 1), (1e-9,) * 3, pbc=(0, 1 << 20, 0)), "PBC")


def test_sim_setters_reject(gpu):
    _say("sim setter errors")
    N = (8, 6, 2)
    s = gpu.Sim(N, (2e-9,) * 3, dtype=np.float32)
    _raises(gpu, lambda: s.set_solver(7), "-4", "solver 7")
    _raises(gpu, lambda: s.set_solver(1), "solver 1")
    _raises(gpu, lambda: gpu._call("um_sim_set_m", s.h, None, None, None), "-1", "NULL m")
    _raises(gpu, lambda: gpu._call("um_sim_set_solver_params", s.h, None), "NULL knobs")
    _raises(gpu, lambda: s.set_time_function(999, lambda t: 1.0), "time function")
    _raises(gpu, lambda: gpu._call("um_sim_step", None), "-1")
    # a non-positive DemagAccuracy is refused when the tensor is built
    setup(s, gpu, Ms=8e5, Aex=1.3e-11, alpha=0.1, dtype=np.float32)
    s.set_m(random_unit((N[2], N[1], N[0]), 1))
    s.set_demag_accuracy(-1.0)
    _raises(gpu, lambda: s.demag_field(), "accuracy")
    # the handle and the library remain usable
    s.set_demag_accuracy(6.0)
    s.set_solver(gpu.SOLVER_RK4)
    s.set_solver_params(dt=1e-14, fixdt=1e-14)
    s.steps(3)
    st = s.state()
    assert st["nsteps"] == 3 and np.all(np.isfinite(s.get_m()))


def test_tensor_and_ops_reject(gpu):
    _say("tensor / op errors")
    _raises(gpu, lambda: gpu.demag_kernel_bf((4, 4, 1), (1e-9,) * 3, accuracy=0.0), "accuracy")
    _raises(gpu, lambda: gpu.demag_kernel_bf((4, 4, 1), (1e-9,) * 3, pbc=(-2, 0, 0)), "PBC")
    _raises(gpu, lambda: gpu.demag_kernel_bf((0, 4, 1), (1e-9,) * 3), "bad arguments")
    import torch
    d = torch.zeros(16, device="cuda")
    _raises(gpu, lambda: gpu.Madd(d, [], []), "um_madd")
    _raises(gpu, lambda: gpu.Madd(d, [d] * 9, [1.0] * 9), "um_madd")
    arr = (C.c_void_p * 2)(d.data_ptr(), None)
    f = (C.c_double * 2)(1.0, 1.0)
    _raises(gpu, lambda: gpu._call("um_madd", gpu.F32, C.c_void_p(d.data_ptr()), 2, arr, f, 16), "NULL source")
    # valid calls still work
    a = torch.arange(16, dtype=torch.float32, device="cuda")
    gpu.Madd(d, [a, a], [1.0, 2.0])
    assert torch.equal(d.cpu(), 3 * a.cpu())
