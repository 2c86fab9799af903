"""Discrete adjoint of fixed-step RK4 through the C ABI (um_sim_tape_begin /
um_sim_backward): gradients of L = sum |m_N - target|^2 with respect to the
per-cell Msat map, a uniform B_ext and the initial m.

What pins it:
- the symmetry the adjoint relies on, on the forward operators themselves:
  <v, K*u> = <u, K*v> for the demag convolution and <v, Lu> = <u, Lv> for the
  exchange + interfacial-DMI operator (with regions, open and closed DMI
  boundaries, periodic edges), in FP64 to 1e-12;
- central finite differences of the oracle's FP64 forward run (the CPU
  restatement of the reference's RK4 path), on per-cell Msat, each B_ext
  component and tangent perturbations of m0: relative error <= 1e-4;
- the checkpointed sweep (segments recomputed from checkpoints) reproduces the
  full-tape gradients exactly; FP32 agrees with FP64 to 2e-3 normwise.
"""
import math
import sys

import numpy as np
import pytest

from cases import MU0, random_unit, setup

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


def _t(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype)).cuda()


def _h(t):
    return t.detach().cpu().numpy()


# ---- symmetry of the forward operators ------------------------------------------------------
@pytest.mark.parametrize("N,pbc,openbc", [((12, 10, 1), 0, 0), ((12, 10, 3), 0, 1), ((8, 6, 2), 3, 0),
                                          ((9, 7, 4), 7, 0)])
def test_exchange_dmi_operator_symmetric(gpu, N, pbc, openbc):
    """<v, Ms B_dmi(u)> = <u, Ms B_dmi(v)> for arbitrary (non-unit) fields."""
    from umagnus_amd.params import exchange_lut
    _say("L symmetry", N, pbc, openbc)
    rng = np.random.default_rng(sum(N) + pbc)
    shape = (N[2], N[1], N[0])
    regions = rng.integers(0, 3, shape).astype(np.uint8)
    A = exchange_lut({0: 1.1e-11, 1: 1.4e-11, 2: 0.9e-11}, dtype=np.float64)
    D = exchange_lut({0: 2e-3, 1: -1e-3, 2: 3e-3}, dtype=np.float64)
    cs = (2e-9, 3e-9, 2.5e-9)
    Ms = rng.uniform(5e5, 1e6, shape)
    u = rng.standard_normal((3,) + shape)
    v = rng.standard_normal((3,) + shape)

    def apply(x):
        B = [_t(np.zeros(shape), np.float64) for _ in range(3)]
        gpu.AddDMI(B, [_t(x[c], np.float64) for c in range(3)], _t(A, np.float64), _t(D, np.float64),
                   _t(Ms, np.float64), _t(regions, np.uint8), cs, pbc=pbc, OpenBC=bool(openbc))
        return np.stack([_h(b) for b in B]) * Ms

    Lu, Lv = apply(u), apply(v)
    a, b = float((v * Lu).sum()), float((u * Lv).sum())
    assert abs(a - b) <= 1e-12 * max(abs(a), np.abs(v).sum() * np.abs(Lu).max()), (a, b)


@pytest.mark.parametrize("N", [(16, 12, 1), (10, 8, 5)])
def test_demag_operator_self_adjoint(gpu, N):
    _say("K symmetry", N)
    rng = np.random.default_rng(sum(N))
    shape = (N[2], N[1], N[0])
    u = rng.standard_normal((3,) + shape)
    v = rng.standard_normal((3,) + shape)
    conv = gpu.DemagConvolution(N, (2e-9, 3e-9, 2.5e-9), precision=1)

    def apply(x):
        B = [_t(np.zeros(shape), np.float64) for _ in range(3)]
        conv.Exec(B, [_t(x[c], np.float64) for c in range(3)], None, 1 / MU0)
        return np.stack([_h(b) for b in B])

    a, b = float((v * apply(u)).sum()), float((u * apply(v)).sum())
    assert abs(a - b) <= 1e-12 * abs(a), (a, b)


# ---- gradients vs finite differences of the oracle ------------------------------------------
DT = 1e-13
PARAMS = dict(Aex=1.3e-11, Dind=2e-3, alpha=0.3, Ku1=2e5, anisU=(0.2, 0.1, 1.0))


def _sim(mod, N, cs, Ms, bext, m0, dtype, steps=0):
    s = mod.Sim(N, cs, dtype=dtype)
    setup(s, mod, Ms=1.0, Bext=tuple(bext), dtype=dtype, **PARAMS)
    s.set_scalar(mod.P_MSAT, Ms)
    s.set_solver(mod.SOLVER_RK4)
    s.set_solver_params(dt=DT, fixdt=DT)
    s.set_m(m0)
    if steps:
        s.steps(steps)
    return s


def _loss(m, target):
    return float(((m - target) ** 2).sum())


def _gpu_grads(gpu, N, cs, Ms, bext, m0, target, nsteps, dtype, every=0):
    s = _sim(gpu, N, cs, Ms, bext, m0, dtype)
    s.tape_begin(nsteps, every)
    s.steps(nsteps)
    mN = s.get_m()
    gm0, gms, gb = s.backward(2 * (mN - target))
    s.tape_end()
    return _loss(mN, target), gm0, gms, gb


def _problem(N, seed):
    shape = (N[2], N[1], N[0])
    rng = np.random.default_rng(seed)
    Ms = rng.uniform(6e5, 1e6, shape)
    m0 = random_unit(shape, seed + 1)
    target = random_unit(shape, seed + 2)
    bext = (0.02, -0.01, 0.03)
    return Ms, m0, target, bext


@pytest.mark.parametrize("N", [(12, 10, 1), (8, 6, 3)])
def test_adjoint_vs_oracle_fd(gpu, oracle, N):
    _say("adjoint vs fd", N)
    cs = (2e-9, 2e-9, 1.5e-9)
    nsteps = 20
    Ms, m0, target, bext = _problem(N, 11)
    L, gm0, gms, gb = _gpu_grads(gpu, N, cs, Ms, bext, m0, target, nsteps, np.float64)
    # the forward run agrees with the oracle's
    so = _sim(oracle, N, cs, Ms, bext, m0, np.float64, nsteps)
    assert abs(_loss(so.get_m(), target) - L) <= 1e-10 * L

    def fd_loss(Ms_=Ms, bext_=bext, m0_=m0):
        s = _sim(oracle, N, cs, Ms_, bext_, m0_, np.float64, nsteps)
        return _loss(s.get_m(), target)

    rng = np.random.default_rng(5)
    shape = Ms.shape
    cells = [tuple(rng.integers(0, d) for d in shape) for _ in range(6)]
    for c in cells:
        h = 1e-4 * Ms[c]
        Mp, Mm = Ms.copy(), Ms.copy()
        Mp[c] += h
        Mm[c] -= h
        fd = (fd_loss(Ms_=Mp) - fd_loss(Ms_=Mm)) / (2 * h)
        assert abs(gms[c] - fd) <= 1e-4 * abs(fd) + 1e-6 * np.abs(gms).max(), (c, gms[c], fd)
    for k in range(3):
        h = 1e-4
        bp, bm = list(bext), list(bext)
        bp[k] += h
        bm[k] -= h
        fd = (fd_loss(bext_=bp) - fd_loss(bext_=bm)) / (2 * h)
        assert abs(gb[k] - fd) <= 1e-4 * abs(fd), (k, gb[k], fd)
    # m0 along a tangent direction of one cell (set_m renormalises)
    for c in cells[:2]:
        e = rng.standard_normal(3)
        mc = m0[(slice(None),) + c]
        e -= mc * (e @ mc)
        e /= np.linalg.norm(e)
        h = 1e-5
        mp, mm = m0.copy(), m0.copy()
        mp[(slice(None),) + c] += h * e
        mm[(slice(None),) + c] -= h * e
        fd = (fd_loss(m0_=mp) - fd_loss(m0_=mm)) / (2 * h)
        an = float(gm0[(slice(None),) + c] @ e)
        assert abs(an - fd) <= 1e-4 * abs(fd) + 1e-8 * np.abs(gm0).max(), (c, an, fd)


def test_adjoint_checkpointing_and_fp32(gpu):
    """Segments of 5 steps recomputed from checkpoints give the full-tape result;
    the FP32 adjoint follows the FP64 one."""
    N, cs, nsteps = (16, 12, 1), (2e-9, 2e-9, 1e-9), 23
    Ms, m0, target, bext = _problem(N, 3)
    _, g0a, gma, gba = _gpu_grads(gpu, N, cs, Ms, bext, m0, target, nsteps, np.float64, every=0)
    _, g0b, gmb, gbb = _gpu_grads(gpu, N, cs, Ms, bext, m0, target, nsteps, np.float64, every=5)
    assert np.array_equal(gma, gmb) and np.array_equal(g0a, g0b) and np.array_equal(gba, gbb)
    _, g0f, gmf, gbf = _gpu_grads(gpu, N, cs, Ms, bext, m0, target, nsteps, np.float32, every=0)
    assert np.abs(gmf - gma).max() <= 2e-3 * np.abs(gma).max()
    assert np.abs(g0f - g0a).max() <= 2e-3 * np.abs(g0a).max()
    assert np.abs(gbf - gba).max() <= 2e-3 * np.abs(gba).max()


def test_tape_guards(gpu):
    N, cs = (8, 8, 1), (2e-9, 2e-9, 1e-9)
    Ms, m0, target, bext = _problem(N, 4)
    s = _sim(gpu, N, cs, Ms, bext, m0, np.float64)
    s.tape_begin(3)
    s.steps(3)
    with pytest.raises(gpu.UmError):
        s.steps(1)                        # tape full
    with pytest.raises(gpu.UmError):
        s.set_m(m0)                       # m is part of the recorded trajectory
    s.set_scalar(gpu.P_ALPHA, 0.1)
    with pytest.raises(gpu.UmError):
        s.backward(np.zeros((3,) + m0.shape[1:]))   # parameters changed
    s.tape_end()
    s.set_solver(gpu.SOLVER_RK45DP)
    with pytest.raises(gpu.UmError):
        s.tape_begin(3)                   # not RK4


def test_autograd_function(gpu):
    """The adjoint behind a torch.autograd.Function (test harness only)."""
    import torch
    N, cs, nsteps = (12, 8, 1), (2e-9, 2e-9, 1e-9), 10
    Ms0, m0, target, bext0 = _problem(N, 8)
    sim = _sim(gpu, N, cs, Ms0, bext0, m0, np.float64)

    class LLGRun(torch.autograd.Function):
        @staticmethod
        def forward(ctx, Ms, bext):
            sim.set_scalar(gpu.P_MSAT, Ms.detach().cpu().numpy())
            sim.set_vector(gpu.P_BEXT, tuple(float(b) for b in bext.detach().cpu()))
            sim.set_m(m0)
            sim.tape_begin(nsteps)
            sim.steps(nsteps)
            return torch.as_tensor(sim.get_m())

        @staticmethod
        def backward(ctx, gout):
            _, gms, gb = sim.backward(gout.detach().cpu().numpy())
            sim.tape_end()
            return torch.as_tensor(gms), torch.as_tensor(gb)

    Ms = torch.tensor(Ms0, dtype=torch.float64, requires_grad=True)
    bext = torch.tensor(bext0, dtype=torch.float64, requires_grad=True)
    loss = ((LLGRun.apply(Ms, bext) - torch.as_tensor(target)) ** 2).sum()
    loss.backward()
    L, _, gms, gb = _gpu_grads(gpu, N, cs, Ms0, bext0, m0, target, nsteps, np.float64)
    assert abs(float(loss.detach()) - L) <= 1e-12 * L
    assert np.array_equal(Ms.grad.numpy(), gms) and np.array_equal(bext.grad.numpy(), gb)


def test_differentiation_workload_fd(gpu):
    """The 256 x 256 x 1 film, 2000 RK4 steps: FP64 adjoint (sqrt(N)
    checkpoints) vs central differences on 64 random cells, the perturbed runs
    stepped together as one ensemble.  Central differences with h = 1e-4 Ms
    carry their own O(h^2) truncation error (measured: 1.6e-4 normwise over the
    64 cells, up to 3e-3 on a cell with a small gradient), so the per-cell
    check is against the Richardson combination of h and h/2 (error O(h^4)):
    the adjoint matches it to ~5e-6.  The FP32 adjoint follows the FP64 one."""
    import bench
    N, m0, tgt, Ms, make = bench.adjoint_workload()
    steps = 2000
    grads = {}
    for dt in (np.float64, np.float32):
        s = make(gpu, dt)
        s.tape_begin(steps, 45)
        s.steps(steps)
        _, gms, gb = s.backward(2 * (s.get_m() - tgt))
        s.tape_end()
        grads[dt] = (gms, gb)
        del s
        _say("adjoint done", dt)
    gms, gb = grads[np.float64]
    rng = np.random.default_rng(64)
    flat = rng.choice(gms.size, 64, replace=False)
    sims = []
    for i in flat:
        for d in (1e-4, -1e-4, 5e-5, -5e-5):
            Mp = Ms.copy().reshape(-1)
            Mp[i] += d * Ms.reshape(-1)[i]
            s = make(gpu, np.float64, Ms_map=Mp.reshape(Ms.shape))
            s.set_graph(1)
            sims.append(s)
    gpu.sims_steps(sims, steps)
    _say("fd runs done")
    g = gms.reshape(-1)[flat]
    L = np.array([_loss(s.get_m(), tgt) for s in sims]).reshape(64, 4)
    h = 1e-4 * Ms.reshape(-1)[flat]
    fd1 = (L[:, 0] - L[:, 1]) / (2 * h)
    fd2 = (L[:, 2] - L[:, 3]) / h
    rich = (4 * fd2 - fd1) / 3
    e_norm = np.linalg.norm(g - fd1) / np.linalg.norm(fd1)
    e_cell = np.abs(g - rich) / np.abs(rich)
    _say("normwise vs FD(h=1e-4): %.2e  max per-cell vs Richardson: %.2e  max per-cell vs FD(h): %.2e"
         % (e_norm, e_cell.max(), (np.abs(g - fd1) / np.abs(fd1)).max()))
    assert e_norm <= 5e-4, e_norm
    assert (np.abs(g - rich) <= 2e-5 * np.abs(rich) + 1e-7 * np.abs(gms).max()).all(), e_cell.max()
    g32, gb32 = grads[np.float32]
    e32 = np.abs(g32 - gms).max() / np.abs(gms).max()
    eb32 = np.abs(gb32 - gb).max() / np.abs(gb).max()
    _say("fp32 vs fp64: Ms %.2e  Bext %.2e" % (e32, eb32))
    assert e32 <= 2e-2 and eb32 <= 2e-2
