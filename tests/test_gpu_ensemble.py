"""The thermal Heun ensemble (um_sims_steps) through the C ABI.

Members of uneven shapes (odd padded Pz = 3, 5, 7 among them) step together
as in the 256-member ensemble workload: Heun at FixDt 5e-15 s, T = 300 K, one
Philox stream per member (engine/heun.go:15-59, engine/temperature.go:55-127).
Checks: member trajectories against the FP64 oracle stepping the same member
alone with the same thermal stream (per cell |dm| <= 1e-5 after 20 FP32
steps; the Philox normals are bit-exact, so the difference is FP32 rounding),
the full 256-member ensemble under |m| = 1, finite state and batched ==
stepped alone (bit for bit) on sampled members, and the step controller's
FSAL torque dropped at the start of every ensemble call as Steps drops it.
"""
import sys

import numpy as np
import pytest

from cases import random_unit, setup

pytestmark = pytest.mark.gpu

DT = 5e-15


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


def _member(mod, N, i, dtype, graph=None):
    s = mod.Sim(N, (2e-9,) * 3, dtype=dtype)
    setup(s, mod, Ms=8e5, Aex=1.3e-11, alpha=0.02, temp=300.0, seed=5, dtype=dtype)
    s.set_thermal_seed(5, i)
    s.set_solver(mod.SOLVER_HEUN)
    s.set_solver_params(dt=DT, fixdt=DT)
    s.set_m(random_unit((N[2], N[1], N[0]), 100 + i))
    if graph is not None:
        s.set_graph(graph)
    return s


ENSEMBLES = {
    "odd_pz": [(32, 48, 2), (64, 32, 3), (48, 48, 4)],
    "mixed": [(96, 32, 1), (32, 32, 1), (64, 64, 2), (48, 96, 3), (32, 64, 4), (192, 32, 2), (32, 256, 1)],
}


@pytest.mark.parametrize("name", sorted(ENSEMBLES))
def test_ensemble_members_vs_oracle(gpu, oracle, name):
    sizes = ENSEMBLES[name]
    _say("ensemble", name, sizes)
    batch = [_member(gpu, N, i, np.float32, graph=1) for i, N in enumerate(sizes)]
    gpu.sims_steps(batch, 20)
    for i, N in enumerate(sizes):
        so = _member(oracle, N, i, np.float64)
        so.steps(20)
        mo, mg = so.get_m(), batch[i].get_m().astype(np.float64)
        d = float(np.linalg.norm(mg - mo, axis=0).max())
        moved = float(np.linalg.norm(mo - random_unit((N[2], N[1], N[0]), 100 + i), axis=0).max())
        _say(N, "max |dm|", d, "moved", moved)
        assert moved > 1e-2          # the thermal field acts
        assert d <= 1e-5, (N, d)
        st = batch[i].state()
        assert st["nsteps"] == 20 and abs(st["t"] - 20 * DT) < 1e-25


def ensemble_members(count=256, seed=4):
    sizes = (32, 48, 64, 96, 128, 192, 256)
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(sizes)), int(rng.choice(sizes)), int(rng.integers(1, 5))) for _ in range(count)]


def test_ensemble_256_properties(gpu):
    """The 256-member ensemble of the benchmark, 20 steps through um_sims_steps:
    unit |m| and finite state for every member; members 0, 1 odd-Pz member,
    a middle one and the last stepped alone give the same bits."""
    import torch
    sizes = ensemble_members()
    batch = [_member(gpu, N, i, np.float32, graph=1) for i, N in enumerate(sizes)]
    _say("256 members built", sum(a * b * c for a, b, c in sizes), "cells")
    gpu.sims_steps(batch, 20)
    for i, s in enumerate(batch):
        m = s.get_m()
        assert np.isfinite(m).all(), i
        assert np.abs(np.linalg.norm(m, axis=0) - 1).max() < 1e-5, i
        assert s.state()["nsteps"] == 20
    odd = next(i for i, N in enumerate(sizes) if N[2] in (2, 3, 4))
    for i in sorted({0, odd, len(sizes) // 2, len(sizes) - 1}):
        alone = _member(gpu, sizes[i], i, np.float32, graph=1)
        alone.steps(20)
        assert np.array_equal(alone.get_m(), batch[i].get_m()), i
        del alone
    del batch
    torch.cuda.empty_cache()


@pytest.mark.parametrize("solver", [5, 3])
def test_ensemble_drops_fsal_like_steps(gpu, solver):
    """An FSAL stepper (RK45DP, RK23) stepped through um_sims_steps after a
    parameter change starts from a fresh torque, as Steps does (engine/run.go
    RunWhile -> stepper.Free()): the ensemble member equals the member stepped
    alone, statistics and evaluation count included."""
    def make():
        s = gpu.Sim((32, 16, 2), (2e-9,) * 3, dtype=np.float32)
        setup(s, gpu, Ms=8e5, Aex=1.3e-11, Dind=1e-3, Ku1=2e5, alpha=0.05, dtype=np.float32)
        s.set_solver(solver)
        s.set_solver_params(dt=1e-14, fixdt=2e-14)
        s.set_m(random_unit((2, 16, 32), 9))
        s.set_graph(1)
        return s
    a, b = make(), make()
    a.steps(5)
    gpu.sims_steps([b], 5)
    for s in (a, b):
        s.set_vector(gpu.P_BEXT, (0.02, -0.01, 0.03))
    a.steps(6)
    gpu.sims_steps([b], 6)
    sa, sb = a.state(), b.state()
    for key in ("t", "nsteps", "nevals", "lasterr", "peakerr", "lasttorque"):
        assert sa[key] == sb[key], (key, sa[key], sb[key])
    assert np.array_equal(a.get_m(), b.get_m())


# ---- batched launches (um_sims_steps_batched) -----------------------------------------------
BATCH_MIXES = {
    # 2D and 3D, odd Pz, X lengths with and without the fused / four-step kernels
    "heun_f32": [((32, 48, 2), "heun", np.float32), ((64, 32, 3), "heun", np.float32),
                 ((48, 48, 4), "heun", np.float32), ((96, 32, 1), "heun", np.float32),
                 ((128, 64, 1), "heun", np.float32), ((256, 32, 2), "heun", np.float32),
                 ((64, 64, 2), "heun", np.float32), ((192, 96, 3), "heun", np.float32)],
    # different solvers and precisions in one ensemble
    "mixed": [((32, 32, 1), "heun", np.float32), ((64, 48, 2), "rk45dp", np.float32),
              ((48, 32, 3), "rk4", np.float64), ((64, 64, 1), "rk23", np.float32),
              ((32, 48, 2), "heun", np.float64), ((128, 32, 4), "rk45dp", np.float32)],
}
SOLVERS = {"heun": 2, "rk23": 3, "rk4": 4, "rk45dp": 5}


def _bmember(mod, N, i, solver, dtype):
    s = mod.Sim(N, (2e-9,) * 3, dtype=dtype)
    if solver == "heun":
        setup(s, mod, Ms=8e5, Aex=1.3e-11, alpha=0.02, temp=300.0, seed=5, dtype=dtype)
        s.set_thermal_seed(5, i)
        dt = DT
    else:
        setup(s, mod, Ms=8e5, Aex=1.3e-11, Dind=1e-3, Ku1=2e5, alpha=0.05, dtype=dtype)
        dt = 2e-14
    s.set_solver(SOLVERS[solver])
    s.set_solver_params(dt=dt, fixdt=dt)
    s.set_m(random_unit((N[2], N[1], N[0]), 200 + i))
    s.set_graph(1)
    return s


@pytest.mark.parametrize("name", sorted(BATCH_MIXES))
def test_batched_equals_alone(gpu, name):
    """Members stepped in batched launches give the bits of the same members
    stepped alone (graph replays), statistics and evaluation counts included,
    over two calls (the second reuses the merged launch plan) and across a
    parameter change (which rebuilds it); the ensemble replays two steps in
    far fewer launches than its members would one by one."""
    mix = BATCH_MIXES[name]
    _say("batched", name)
    batch = [_bmember(gpu, N, i, sol, dt) for i, (N, sol, dt) in enumerate(mix)]
    alone = [_bmember(gpu, N, i, sol, dt) for i, (N, sol, dt) in enumerate(mix)]
    nl = gpu.sims_steps_batched(batch, 9)
    nl2 = gpu.sims_steps_batched(batch, 7)
    assert nl2 == nl
    for s in alone:
        s.steps(9)
        s.steps(7)
    for s in batch + alone:
        s.set_scalar(gpu.P_ALPHA, 0.03)
    gpu.sims_steps_batched(batch, 6)
    for s in alone:
        s.steps(6)
    for i, (a, b) in enumerate(zip(alone, batch)):
        sa, sb = a.state(), b.state()
        for key in ("t", "nsteps", "nevals", "lasterr", "peakerr", "lasttorque"):
            assert sa[key] == sb[key], (i, mix[i], key, sa[key], sb[key])
        assert np.array_equal(a.get_m(), b.get_m()), (i, mix[i])
        assert sb["nsteps"] == 22
    # launches of two steps of each member alone (eager steps: an upper bound
    # of its captured steps by the few fills a captured step folds away)
    alone_launches = 0
    for i, (N, sol, dt) in enumerate(mix):
        t = _bmember(gpu, N, i, sol, dt)
        t.steps(1)
        with gpu.KernelTrace() as tr:
            t.steps(2)
        alone_launches += len(tr.names)
    _say("launches per two steps: batched", nl, "members one by one", alone_launches)
    # one solver and precision: the members share most kernel variants; mixed
    # solvers and precisions share few, and never take more launches
    assert 0 < nl < (0.75 if name == "heun_f32" else 1.0) * alone_launches


def test_batched_members_vs_oracle(gpu, oracle):
    """Batched thermal Heun members against the FP64 oracle stepping each
    member alone with the same Philox stream: per cell |dm| <= 1e-5 after 20
    FP32 steps (2D, odd Pz 3, 5, 7)."""
    sizes = [(32, 48, 2), (64, 32, 3), (48, 48, 4), (96, 32, 1)]
    batch = [_member(gpu, N, i, np.float32, graph=1) for i, N in enumerate(sizes)]
    assert gpu.sims_steps_batched(batch, 20) > 0
    for i, N in enumerate(sizes):
        so = _member(oracle, N, i, np.float64)
        so.steps(20)
        d = float(np.linalg.norm(batch[i].get_m().astype(np.float64) - so.get_m(), axis=0).max())
        _say(N, "batched vs oracle max |dm|", d)
        assert d <= 1e-5, (N, d)


def test_batched_256_properties(gpu):
    """The 256-member ensemble of the benchmark in batched launches: 20 steps,
    unit |m| and finite state for every member, sampled members (first, an odd
    Pz one, middle, last) bit-identical to stepping alone."""
    import torch
    sizes = ensemble_members()
    batch = [_member(gpu, N, i, np.float32, graph=1) for i, N in enumerate(sizes)]
    nl = gpu.sims_steps_batched(batch, 20)
    _say("256 members: launches per two steps", nl)
    assert 0 < nl < 1000
    for i, s in enumerate(batch):
        m = s.get_m()
        assert np.isfinite(m).all(), i
        assert np.abs(np.linalg.norm(m, axis=0) - 1).max() < 1e-5, i
        assert s.state()["nsteps"] == 20
    odd = next(i for i, N in enumerate(sizes) if N[2] in (2, 3, 4))
    for i in sorted({0, odd, len(sizes) // 2, len(sizes) - 1}):
        alone = _member(gpu, sizes[i], i, np.float32, graph=1)
        alone.steps(20)
        assert np.array_equal(alone.get_m(), batch[i].get_m()), i
        del alone
    del batch
    torch.cuda.empty_cache()


def test_batched_short_runs_and_eager_members(gpu):
    """n = 0, 1, 2 (no pair to replay after the eager step) and a member not in
    graph mode (stepped eagerly) next to batched ones: every member matches its
    twin stepped alone."""
    mix = [((32, 32, 2), "heun", np.float32), ((64, 32, 1), "rk45dp", np.float32)]
    batch = [_bmember(gpu, N, i, sol, dt) for i, (N, sol, dt) in enumerate(mix)]
    alone = [_bmember(gpu, N, i, sol, dt) for i, (N, sol, dt) in enumerate(mix)]
    eager_b = _bmember(gpu, (32, 32, 1), 9, "heun", np.float32)
    eager_a = _bmember(gpu, (32, 32, 1), 9, "heun", np.float32)
    eager_b.set_graph(0)
    eager_a.set_graph(0)
    for n in (0, 1, 2, 5):
        gpu.sims_steps_batched(batch + [eager_b], n)
        for s in alone + [eager_a]:
            s.steps(n)
    for a, b in zip(alone + [eager_a], batch + [eager_b]):
        assert a.state()["nsteps"] == b.state()["nsteps"] == 8
        assert np.array_equal(a.get_m(), b.get_m())
