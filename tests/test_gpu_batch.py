"""Parameter-set batch (cmg_batch_*): S Guccione parameter sets on one mesh, solved
together, against independent single-set solves and the CPU FP64 oracle.

Tolerances:
  batched residual          bitwise equal to the single-set residual (same kernels)
  batched static steps      ||x_batch - x_single|| <= 5e-8 ||u_single|| per set and step (two
                            Newton runs that stop at |r| <= 1e-11 on matrices with K/C up to
                            500 and differ in their Krylov method and iteration counts)
                            ||x_batch - x_oracle|| <= 1e-8 ||u_oracle||
  graph replay vs eager     bitwise equal node coordinates
"""
import numpy as np
import pytest

from oracle.drivers import CpuMechanics
from cardiomechanics_amd import materials as MT
from cardiomechanics_amd import mesh as M
from cardiomechanics_amd.context import MechanicsContext, ParameterBatch

pytestmark = pytest.mark.gpu

TIGHT = dict(newton_atol=1e-11, newton_rtol=1e-12, newton_stol=1e-15, krylov_maxit=20000, newton_maxit=200)
NO_PTAN = dict(TIGHT, pressure_tangent=0)   # open pressure face (cube): keep the batched matrices symmetric


def deformed(mesh, amp, seed):
    L = np.ptp(mesh.coords, axis=0).max()
    return mesh.coords + M.smooth_displacement(mesh.coords, L, amp, seed)


def test_batch_residual_is_bitwise_the_single_set_residual():
    mesh = M.cube_mesh(2)
    sets = MT.guccione_sweep(3, seed=7)
    xs = np.stack([deformed(mesh, 0.02, 100 + i).ravel() for i in range(3)])
    with MechanicsContext(mesh, [sets[0]]) as ctx:
        ctx.set_pressures([1500.0])
        with ParameterBatch(ctx, [[m] for m in sets]) as batch:
            fb = batch.residual(xs, 0.0)
    for i, m in enumerate(sets):
        with MechanicsContext(mesh, [m]) as one:
            one.set_pressures([1500.0])
            f1 = one.residual(xs[i], 0.0)
        assert fb[i].tobytes() == f1.tobytes(), i


def run_batch(mesh, sets, pressure, steps, use_graph=True, lanes=0, variant=0, **opts):
    hist = []
    with MechanicsContext(mesh, [sets[0]], **opts) as ctx:
        with ParameterBatch(ctx, [[m] for m in sets], use_graph=use_graph, spmv_lanes=lanes,
                            element_variant=variant) as batch:
            for k in range(steps + 1):
                t = k / steps
                ctx.set_pressures([pressure * t])
                rc, st = batch.static_step(t, 1.0 / steps)
                assert rc == 2, [s.as_dict() for s in st]
                hist.append((batch.nodes(), [(s.newton_iterations, s.krylov_iterations) for s in st]))
            info = batch.info()
            vols = batch.cavity_volumes()
    return hist, info, vols


def test_batch_static_steps_match_independent_solves():
    """4 sweep sets on a small LV, passive inflation to 1 kPa in 8 steps (closed T3
    cavity: the pressure tangent is symmetric and included)."""
    mesh = M.lv_mesh(dims=(2, 4, 8))
    sets = MT.guccione_sweep(4, seed=0)
    hist, info, vols = run_batch(mesh, sets, -1000.0, 8, **TIGHT)
    print("autotuner:", info)
    assert info["tuned"] == 1 and info["element_variant"] in range(1, 9) and info["spmv_lanes"] in (4, 8, 16)
    assert info["preconditioner"] == "vertex-patch Schwarz"   # T10: the context built its patches
    for i, m in enumerate(sets):
        with MechanicsContext(mesh, [m], krylov_method=1, **TIGHT) as one:
            for k in range(9):
                t = k / 8
                one.set_pressures([-1000.0 * t])
                rc, st = one.static_step(t, 0.125)
                assert rc == 2
                g = one.get_state()["nodes"]
                xb = hist[k][0][i]
                u = g - mesh.coords.ravel()
                err = np.linalg.norm(xb - g) / max(np.linalg.norm(u), 1e-300)
                print(f"set {i} t={t:.2f} batch newton/cg {hist[k][1][i]} single newton {st.newton_iterations} "
                      f"rel err {err:.2e}")
                assert err <= 5e-8
            v1 = one.cavity_volumes()[0]
        assert abs(vols[i][0] - v1) <= 5e-8 * abs(v1)


def test_batch_matches_oracle_on_the_cube():
    mesh = M.cube_mesh(2)
    sets = MT.guccione_sweep(2, seed=3)
    hist, _, _ = run_batch(mesh, sets, 2000.0, 3, **NO_PTAN)
    for i, m in enumerate(sets):
        cpu = CpuMechanics(mesh, [m])
        for k in range(4):
            t = k / 3
            cpu.pressures[:] = 2000.0 * t
            cpu.static_step(t, 1.0 / 3)
            u = cpu.nodes - mesh.coords.ravel()
            err = np.linalg.norm(hist[k][0][i] - cpu.nodes) / max(np.linalg.norm(u), 1e-300)
            print(f"set {i} t={t:.2f} rel err vs oracle {err:.2e}")
            assert err <= 1e-8


def test_graph_replay_equals_eager_launches():
    mesh = M.cube_mesh(2)
    sets = MT.guccione_sweep(3, seed=5)
    h_graph, info_g, _ = run_batch(mesh, sets, 2000.0, 2, use_graph=True, lanes=8, variant=2, **NO_PTAN)
    h_eager, info_e, _ = run_batch(mesh, sets, 2000.0, 2, use_graph=False, lanes=8, variant=2, **NO_PTAN)
    assert info_g["use_graph"] == 1 and info_e["use_graph"] == 0
    for (a, _), (b, _) in zip(h_graph, h_eager):
        assert a.tobytes() == b.tobytes()


def test_batch_set_nodes_starts_each_set_from_its_configuration():
    """cmg_batch_set_nodes: each set's configuration is replaced (read back bit for bit),
    and a static step from it reaches the root of the single-set step from the same
    nodes (cmg_set_state)."""
    mesh = M.cube_mesh(2)
    sets = MT.guccione_sweep(2, seed=11)
    starts = np.stack([deformed(mesh, 0.002, 200 + i).ravel() for i in range(2)])
    with MechanicsContext(mesh, [sets[0]], **NO_PTAN) as ctx:
        with ParameterBatch(ctx, [[m] for m in sets]) as batch:
            batch.set_nodes(starts)
            assert batch.nodes().tobytes() == starts.tobytes()
            ctx.set_pressures([500.0])
            rc, st = batch.static_step(1.0, 1.0)
            assert rc == 2, [s.as_dict() for s in st]
            xb = batch.nodes()
    for i, m in enumerate(sets):
        with MechanicsContext(mesh, [m], **NO_PTAN) as one:
            one.set_state(nodes=starts[i])
            one.set_pressures([500.0])
            rc, st = one.static_step(1.0, 1.0)
            assert rc == 2
            g = one.get_state()["nodes"]
        err = np.linalg.norm(xb[i] - g) / np.linalg.norm(g - mesh.coords.ravel())
        print(f"set {i}: batch vs single from the same start, rel diff {err:.2e}")
        assert err <= 5e-8
