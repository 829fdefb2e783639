"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle.

Tolerances (FP64 path):
  element residuals        max_e |f_gpu - f_orc| <= 1e-12 * max|f_orc| of that element
  element tangents         ||K_gpu - K_cs||_F <= 1e-12 ||K_cs||_F per element (complex-step oracle)
  assembled vector/matrix  <= 1e-12 relative to the largest element contribution / ||K||_F
  solver steps             ||x_gpu - x_orc|| <= 1e-8 ||u_orc||, cavity volume rel. 1e-8
Mixed precision (FP32 tangent values): element tangents 2e-6 relative; converged
displacements still 1e-8 (the residual and the convergence checks stay FP64).
"""
import math

import numpy as np
import pytest

import oracle as O
from oracle.drivers import CpuMechanics
from cardiomechanics_amd import materials as MT
from cardiomechanics_amd import mesh as M
from cardiomechanics_amd import solver as SV
from cardiomechanics_amd.context import CmgError, MechanicsContext

pytestmark = pytest.mark.gpu


def deformed(mesh, amp=0.02, seed=1234):
    L = np.ptp(mesh.coords, axis=0).max()
    return mesh.coords + M.smooth_displacement(mesh.coords, L, amp, seed)


def oracle_elements(mesh):
    ed = O.ElementData(mesh)
    ed.bcmask = np.zeros_like(ed.bcmask)       # the GPU per-element outputs are unmasked
    return ed


def lv_small(dims=(2, 4, 8), surface_type=3):
    mesh = M.lv_mesh(dims=dims, surface_type=surface_type)
    return mesh, [MT.lv_holzapfel()]


def cube_small(n=2, surface_type=3):
    return M.cube_mesh(n, surface_type=surface_type), [MT.cube_guccione()]


def ellipsoid_fixture():
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ellipsoid_subset.npz"))
    E = z["elements"].shape[0]
    mesh = M.Mesh(coords=z["coords"], elements=z["elements"], bases=z["bases"], bc=z["bc"],
                  material=np.zeros(E, np.int32), activation=np.linspace(0, 0.05, E),
                  surfaces=z["surfaces"], surface_id=np.ones(z["surfaces"].shape[0], np.int32), name="ellipsoid")
    return mesh


def lv_with_tension_fn(kind, **kw):
    m = MT.lv_holzapfel()
    m.tension_fn, m.fn = MT.function_params(kind, **kw)
    return M.lv_mesh(dims=(2, 4, 8)), [m]


CASES = {
    "cube_guccione": lambda: cube_small(2),
    "lv_holzapfel_active": lambda: lv_small(),
    "ellipsoid_fixture_holzapfel": lambda: (ellipsoid_fixture(), [MT.lv_holzapfel()]),
    "ellipsoid_fixture_guccione": lambda: (ellipsoid_fixture(), [MT.cube_guccione()]),
    "cube_t4_guccione": lambda: (M.cube_mesh(2, order=4), [MT.cube_guccione()]),
    "lv_holzapfel_tanh": lambda: (M.lv_mesh(dims=(2, 4, 8)), [MT.with_tanh(MT.lv_holzapfel(), PeakTension=60.0)]),
    "cube_t4_guccione_tanh": lambda: (M.cube_mesh(2, order=4), [MT.with_tanh(MT.cube_guccione())]),
    "lv_holzapfel_sinusdriver": lambda: lv_with_tension_fn("SinusDriver", start=0.0, stop=0.3),
    "lv_holzapfel_doublehill": lambda: lv_with_tension_fn("DoubleHill", start=0.0, stop=10.0, onset=0.02),
}


@pytest.mark.parametrize("case", list(CASES))
def test_element_residuals_match_oracle(case):
    mesh, mats = CASES[case]()
    x = deformed(mesh)
    t = 0.12
    with MechanicsContext(mesh, mats) as ctx:
        fg = ctx.element_residuals(x, t)
    rc, fo, _ = O.residual_batch(oracle_elements(mesh), x, mats, t)
    assert rc == O.SUCCESS
    scale = np.abs(fo).max(axis=1, keepdims=True)
    err = np.abs(fg - fo) / scale
    print(f"{case}: max element residual rel err {err.max():.2e}")
    assert err.max() <= 1e-12


@pytest.mark.parametrize("case", list(CASES))
def test_element_tangents_match_complex_step(case):
    mesh, mats = CASES[case]()
    x = deformed(mesh)
    t = 0.12
    with MechanicsContext(mesh, mats) as ctx:
        Kg = ctx.element_tangents(x, t)
    rc, Ko, _ = O.tangent_batch(oracle_elements(mesh), x, mats, t, kind="cs")
    assert rc == O.SUCCESS
    rel = np.linalg.norm((Kg - Ko).reshape(len(Ko), -1), axis=1) / np.linalg.norm(Ko.reshape(len(Ko), -1), axis=1)
    print(f"{case}: max element tangent rel err {rel.max():.2e}")
    assert rel.max() <= 1e-12


@pytest.mark.parametrize("tension", ["sinus", "tanh"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("variant", [1, 3, 4, 5, 6, 7, 8])
def test_element_kernel_variants_agree_with_the_oracle(variant, precision, tension):
    """every residual+tangent kernel variant against the complex-step oracle (and the
    residual against the oracle residual), T10 LV with Holzapfel and active tension
    (FromFunction sin^2 or the stretch-dependent TanH model)"""
    mesh, mats = lv_small()
    if tension == "tanh":
        mats = [MT.with_tanh(mats[0], PeakTension=60.0)]
    x = deformed(mesh)
    with MechanicsContext(mesh, mats, element_variant=variant, precision=precision) as ctx:
        Kg = ctx.element_tangents(x, 0.12).astype(np.float64)
        fg = ctx.element_residuals(x, 0.12)
    rc, Ko, _ = O.tangent_batch(oracle_elements(mesh), x, mats, 0.12, kind="cs")
    rc2, fo, _ = O.residual_batch(oracle_elements(mesh), x, mats, 0.12)
    rel = np.linalg.norm((Kg - Ko).reshape(len(Ko), -1), axis=1) / np.linalg.norm(Ko.reshape(len(Ko), -1), axis=1)
    print(f"variant {variant} precision {precision}: element tangent rel err {rel.max():.2e}")
    assert rel.max() <= (1e-12 if precision == 0 else 2e-6)
    assert (np.abs(fg - fo) / np.abs(fo).max(axis=1, keepdims=True)).max() <= 1e-12


def test_mixed_precision_element_tangents():
    mesh, mats = lv_small()
    x = deformed(mesh)
    with MechanicsContext(mesh, mats, precision=1) as ctx:
        Kg = ctx.element_tangents(x, 0.12).astype(np.float64)
        fg = ctx.element_residuals(x, 0.12)
    rc, Ko, _ = O.tangent_batch(oracle_elements(mesh), x, mats, 0.12, kind="cs")
    rc, fo, _ = O.residual_batch(oracle_elements(mesh), x, mats, 0.12)
    rel = np.linalg.norm((Kg - Ko).reshape(len(Ko), -1), axis=1) / np.linalg.norm(Ko.reshape(len(Ko), -1), axis=1)
    print(f"mixed precision: element tangent rel err {rel.max():.2e}")
    assert rel.max() <= 2e-6
    assert (np.abs(fg - fo) / np.abs(fo).max(axis=1, keepdims=True)).max() <= 1e-12   # residual stays FP64


ASSEMBLY_CASES = dict(CASES, lv_holzapfel_t6=lambda: lv_small(surface_type=6))


@pytest.mark.parametrize("case", ["cube_guccione", "lv_holzapfel_active", "ellipsoid_fixture_holzapfel",
                                  "lv_holzapfel_t6"])
def test_assembled_residual_and_tangent(case):
    mesh, mats = ASSEMBLY_CASES[case]()
    x = deformed(mesh)
    t = 0.12
    p = -4000.0
    cpu = CpuMechanics(mesh, mats)
    cpu.pressures[:] = p
    with MechanicsContext(mesh, mats) as ctx:
        ctx.set_pressures([p])
        rg = ctx.residual(x.ravel(), t)
        Kg = ctx.bsr(ctx.tangent(x.ravel(), t)).toarray()
    ro = cpu.residual(x.ravel(), t)
    Ko = cpu.tangent(x.ravel(), t).toarray()
    rc, fo, _ = O.residual_batch(oracle_elements(mesh), x, mats, t)
    scale = np.abs(fo).max()
    print(f"{case}: residual err {np.abs(rg - ro).max() / scale:.2e}  tangent err "
          f"{np.linalg.norm(Kg - Ko) / np.linalg.norm(Ko):.2e}")
    assert np.abs(rg - ro).max() <= 1e-12 * scale
    assert np.linalg.norm(Kg - Ko) <= 1e-12 * np.linalg.norm(Ko)


def test_assembled_tangent_without_the_pressure_tangent():
    """pressure_tangent = 0: the surface pressure adds its forces but no tangent blocks
    (the reference's T3 surface sets no matrix entries); the residual keeps the load."""
    mesh, mats = lv_small()
    x = deformed(mesh)
    t, p = 0.12, -4000.0
    cpu = CpuMechanics(mesh, mats, pressure_tangent=False)
    cpu.pressures[:] = p
    with MechanicsContext(mesh, mats, pressure_tangent=0) as ctx:
        ctx.set_pressures([p])
        rg = ctx.residual(x.ravel(), t)
        Kg = ctx.bsr(ctx.tangent(x.ravel(), t)).toarray()
    ro = cpu.residual(x.ravel(), t)
    Ko = cpu.tangent(x.ravel(), t).toarray()
    cpu_p = CpuMechanics(mesh, mats)
    cpu_p.pressures[:] = p
    dK = np.linalg.norm(cpu_p.tangent(x.ravel(), t).toarray() - Ko) / np.linalg.norm(Ko)
    assert dK > 1e-8          # the pressure tangent is large enough for the check to see it
    print(f"pressure tangent {dK:.2e} of ||K||; without it: residual err "
          f"{np.abs(rg - ro).max() / np.abs(ro).max():.2e}, tangent err {np.linalg.norm(Kg - Ko) / np.linalg.norm(Ko):.2e}")
    assert np.abs(rg - ro).max() <= 1e-12 * np.abs(ro).max()
    assert np.linalg.norm(Kg - Ko) <= 1e-12 * np.linalg.norm(Ko)


def test_transposed_assembly_convention():
    """transpose_tangent=1 assembles each solid element block transposed, the T10
    reference convention (CBElementSolidT10.cpp:559-565)."""
    mesh, mats = lv_small()
    x = deformed(mesh)
    with MechanicsContext(mesh, mats) as a, MechanicsContext(mesh, mats, transpose_tangent=1) as b:
        Ka = a.bsr(a.tangent(x.ravel(), 0.12, include_pressure=False)).toarray()
        Kb = b.bsr(b.tangent(x.ravel(), 0.12, include_pressure=False)).toarray()
        # the surface pressure blocks enter untransposed in either convention
        for ctx in (a, b):
            ctx.set_pressures([-4000.0])
        Pa = a.bsr(a.tangent(x.ravel(), 0.12)).toarray() - Ka
        Pb = b.bsr(b.tangent(x.ravel(), 0.12)).toarray() - Kb
    assert np.linalg.norm(Ka - Ka.T) > 1e-3 * np.linalg.norm(Ka)       # Holzapfel as coded is unsymmetric
    np.testing.assert_array_equal(Kb, Ka.T)
    assert np.linalg.norm(Pa) > 1e-6 * np.linalg.norm(Ka)
    assert np.linalg.norm(Pb - Pa) <= 1e-12 * np.linalg.norm(Ka)


@pytest.mark.parametrize("case", ["lv_t10", "cube_t4"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_matrix_gather_is_the_in_order_sum_of_the_element_tangents(precision, case):
    """Each assembled block is its element blocks summed in FP64 in element order, then
    its surface blocks (zeros without the pressure tangent), rounded once to the value
    type; constrained rows and columns hold the identity.  Formed on the host from the
    device's own element tangents, that sum equals the device's values bit for bit."""
    if case == "lv_t10":
        mesh, mats = lv_small()
    else:
        mesh, mats = M.cube_mesh(3, order=4), [MT.cube_guccione()]
    x = deformed(mesh).ravel()
    with MechanicsContext(mesh, mats, precision=precision) as ctx:
        vals = ctx.tangent(x, 0.12, include_pressure=False).reshape(-1, 3, 3)
        Ke = ctx.element_tangents(x, 0.12)
        rowptr, colind = ctx.rowptr.copy(), ctx.colind.copy()
    N = mesh.n_nodes
    E, npe = mesh.elements.shape
    keys = np.repeat(np.arange(N, dtype=np.int64), np.diff(rowptr)) * N + colind
    assert np.all(np.diff(keys) > 0)

    def slots(conn):
        k = conn.shape[1]
        r = np.repeat(conn.astype(np.int64), k, axis=1).ravel()
        c = np.tile(conn.astype(np.int64), (1, k)).ravel()
        s = np.searchsorted(keys, r * N + c)
        assert np.array_equal(keys[s], r * N + c)
        return s

    blocks = Ke.astype(np.float64).reshape(E, npe, 3, npe, 3).transpose(0, 1, 3, 2, 4).reshape(-1, 3, 3)
    acc = np.zeros((colind.size, 3, 3))
    np.add.at(acc, slots(mesh.elements), blocks)          # sequential: element order per block
    np.add.at(acc, slots(mesh.surfaces), 0.0)
    blkrow = np.repeat(np.arange(N), np.diff(rowptr))
    bc = mesh.bc.astype(np.int64)
    for i in range(3):
        for j in range(3):
            m = ((bc[blkrow] >> i) & 1).astype(bool) | ((bc[colind] >> j) & 1).astype(bool)
            acc[m, i, j] = np.where(blkrow[m] == colind[m], 1.0 if i == j else 0.0, 0.0)
    want = acc.astype(vals.dtype)
    assert (bc != 0).any()
    print(f"precision {precision}: {colind.size} blocks, {np.count_nonzero(want != vals)} differ")
    assert want.tobytes() == vals.tobytes()


@pytest.mark.parametrize("surface_type", [3, 6])
def test_surface_forces_and_cavity_volume(surface_type):
    mesh, mats = lv_small(surface_type=surface_type)
    x = deformed(mesh, 0.01)
    cpu = CpuMechanics(mesh, mats)
    cpu.pressures[:] = -16000.0
    with MechanicsContext(mesh, mats) as ctx:
        fg = ctx.surface_forces(x.ravel(), [-16000.0])
        vg = ctx.cavity_volumes(x.ravel())
    fo = cpu.surface_forces(x.ravel())
    fo[cpu.bcdof] = 0
    vo = cpu.cavity_volumes(x.ravel())
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    assert vg[0] == pytest.approx(vo[0], rel=1e-12)
    assert vg[0] < 0   # endocardial normals point into the cavity, as in the reference's ellipsoid.sur


@pytest.mark.parametrize("case", ["lv_t10", "cube_t4"])
@pytest.mark.parametrize("consistent", [1, 0], ids=["consistent", "lumped"])
def test_mass_matrix(consistent, case):
    """consistent (CBElementSolidT10.cpp:352-394, CBElementSolidT4.cpp:365-410) or lumped
    (CBElementSolidT10.cpp:396-422, CBElementSolidT4.cpp:432-456) mass"""
    mesh, mats = lv_small() if case == "lv_t10" else (M.cube_mesh(2, order=4), [MT.cube_guccione()])
    cpu = CpuMechanics(mesh, mats, consistent_mass=bool(consistent))
    with MechanicsContext(mesh, mats, consistent_mass=consistent) as ctx:
        Mg = ctx.bsr(ctx.mass_matrix()).toarray()
    Mo = cpu.M.toarray()
    if not consistent:
        assert np.count_nonzero(Mo - np.diag(np.diag(Mo))) == 0
    assert np.abs(Mg - Mo).max() <= 1e-14 * np.abs(Mo).max()


def test_t4_static_load_steps_match_oracle():
    """Linear tetrahedra (T4, one quadrature point, CBElementSolidT4.cpp:130-210) in the
    quasi-static pressure ramp on the cube."""
    mesh, mats = M.cube_mesh(2, order=4), [MT.cube_guccione()]
    res = run_static_pair(mesh, mats, 2000.0, 3, newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15)
    for t, err, nit, kit in res:
        print(f"T4 t={t:.2f} newton={nit} krylov={kit} rel err={err:.2e}")
        assert err <= 1e-8


def test_newmark_steps_with_the_beat_parameters_match_oracle():
    """Newmark with beta = 0.3025, gamma = 0.6 (the LV beat's settings): predictor and
    corrector (CBSolverNewmarkBeta.cpp:845-852, 984-1015) depend on gamma only away from
    the trapezoidal rule, so nodes, velocities and accelerations are all compared."""
    mesh, mats = lv_small()
    beta, gamma, dt = 0.3025, 0.6, 0.005
    cpu = CpuMechanics(mesh, mats)
    plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
    with MechanicsContext(mesh, mats, newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15,
                          newmark_beta=beta, newmark_gamma=gamma) as ctx:
        for k in range(1, 7):
            t = k * dt
            p = plug.apply(t, dt)[0]
            ctx.set_pressures([p])
            cpu.pressures[:] = p
            rc, st = ctx.newmark_step(t, dt)
            assert rc == 2, st.as_dict()
            cpu.newmark_step(t, dt, beta=beta, gamma=gamma)
            g = ctx.get_state()
            errs = [np.linalg.norm(g["nodes"] - cpu.nodes) / np.linalg.norm(cpu.nodes - mesh.coords.ravel()),
                    np.linalg.norm(g["velocity"] - cpu.vel) / np.linalg.norm(cpu.vel),
                    np.linalg.norm(g["acceleration"] - cpu.acc) / np.linalg.norm(cpu.acc)]
            print(f"t={t:.3f} newton={st.newton_iterations} rel err disp {errs[0]:.2e} vel {errs[1]:.2e} "
                  f"acc {errs[2]:.2e}")
            assert max(errs) <= 1e-8


def test_t4_newmark_steps_with_lumped_mass_match_oracle():
    """Linear tetrahedra in Newmark steps (beta=1/4, gamma=1/2) with the lumped mass
    (Solver.NewmarkBeta.ConsistentMassMatrix = false, CBElementSolidT4.cpp:432-456) under
    a pressure ramp on the cube."""
    mesh, mats = M.cube_mesh(2, order=4), [MT.cube_guccione()]
    cpu = CpuMechanics(mesh, mats, consistent_mass=False)
    dt = 0.005
    with MechanicsContext(mesh, mats, consistent_mass=0, newton_atol=1e-12, newton_rtol=1e-12,
                          newton_stol=1e-15) as ctx:
        for k in range(1, 5):
            t, p = k * dt, 500.0 * k
            ctx.set_pressures([p])
            cpu.pressures[:] = p
            rc, st = ctx.newmark_step(t, dt)
            assert rc == 2, st.as_dict()
            cpu.newmark_step(t, dt)
            g = ctx.get_state()["nodes"]
            u = cpu.nodes - mesh.coords.ravel()
            err = np.linalg.norm(g - cpu.nodes) / np.linalg.norm(u)
            vg, vo = ctx.cavity_volumes()[0], cpu.cavity_volumes()[0]
            print(f"T4 lumped t={t:.3f} newton={st.newton_iterations} disp err={err:.2e} "
                  f"vol err={abs(vg - vo) / abs(vo):.2e}")
            assert err <= 1e-8
            assert abs(vg - vo) <= 1e-8 * abs(vo)


@pytest.mark.parametrize("pc", [1, 2], ids=["block_jacobi", "schwarz"])
@pytest.mark.parametrize("method", [0, 1, 2], ids=["gmres", "bicgstab_graph", "bicgstab"])
def test_spmv_and_krylov_solve(method, pc):
    """The Newmark system matrix K + M/(beta dt^2) at the unloaded state, dt = 1 ms;
    block Jacobi or vertex-patch Schwarz preconditioning."""
    mesh, mats = lv_small()
    rng = np.random.default_rng(3)
    with MechanicsContext(mesh, mats, krylov_restart=60, krylov_maxit=20000, krylov_method=method,
                          preconditioner=pc) as ctx:
        Kt = ctx.bsr(ctx.tangent(mesh.coords.ravel(), 0.0))
        Mm = ctx.bsr(ctx.mass_matrix())
        vals = np.ascontiguousarray((Kt.data + Mm.data / 0.25e-6).reshape(-1))   # same block pattern
        K = ctx.bsr(vals).tocsr()
        v = rng.normal(size=ctx.n_dofs)
        np.testing.assert_allclose(ctx.spmv(vals, v), K @ v, rtol=1e-13, atol=1e-13 * np.abs(K @ v).max())
        for layout, lane_set in ((0, (4, 8, 16)), (1, (4, 8, 16)), (2, (16, 32, 64))):
            for lanes in lane_set:
                np.testing.assert_allclose(ctx.spmv(vals, v, lanes, layout), K @ v, rtol=1e-13,
                                           atol=1e-13 * np.abs(K @ v).max())
        b = K @ rng.normal(size=ctx.n_dofs)
        xs, st = ctx.linear_solve(vals, b, rtol=1e-12)
    import scipy.sparse.linalg as spla
    xr = spla.spsolve(K.tocsc(), b)
    print(f"method {method} preconditioner {pc} iterations {st.iterations}, "
          f"rel err {np.linalg.norm(xs - xr) / np.linalg.norm(xr):.2e}")
    assert st.converged
    assert np.linalg.norm(xs - xr) <= 1e-7 * np.linalg.norm(xr)


def test_bitwise_reproducible_assembly():
    mesh, mats = lv_small()
    x = deformed(mesh)
    with MechanicsContext(mesh, mats) as ctx:
        ctx.set_pressures([-8000.0])
        r1, k1 = ctx.residual(x.ravel(), 0.1), ctx.tangent(x.ravel(), 0.1)
        r2, k2 = ctx.residual(x.ravel(), 0.1), ctx.tangent(x.ravel(), 0.1)
    assert r1.tobytes() == r2.tobytes() and k1.tobytes() == k2.tobytes()


def test_corrupt_element_is_reported():
    mesh, mats = cube_small(2)
    x = mesh.coords.copy()
    el = mesh.elements[5]
    # push corner 0 of element 5 through its opposite face
    x[el[0]] = x[el[1:4]].mean(axis=0) * 2 - x[el[0]]
    with MechanicsContext(mesh, mats) as ctx:
        with pytest.raises(CmgError) as ei:
            ctx.residual(x.ravel(), 0.0)
    assert ei.value.code == -6
    assert "corrupt" in str(ei.value)


def run_static_pair(mesh, mats, pressure, steps, **opts):
    cpu = CpuMechanics(mesh, mats)
    out = []
    with MechanicsContext(mesh, mats, **opts) as ctx:
        dt = 1.0 / steps
        for k in range(steps + 1):
            t = k * dt
            p = pressure * t
            cpu.pressures[:] = p
            ctx.set_pressures([p])
            rc, st = ctx.static_step(t, dt)
            assert rc == 2, (k, st.as_dict())
            cpu.static_step(t, dt)
            g = ctx.get_state()["nodes"]
            u = cpu.nodes - mesh.coords.ravel()
            err = np.linalg.norm(g - cpu.nodes) / max(np.linalg.norm(u), 1e-300)
            out.append((t, err, st.newton_iterations, st.krylov_iterations))
    return out


def test_static_load_steps_match_oracle():
    """Quasi-static pressure ramp on the cube (the unit-cube benchmark at n=2)."""
    mesh, mats = cube_small(2)
    res = run_static_pair(mesh, mats, 2000.0, 5, newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15)
    for t, err, nit, kit in res:
        print(f"t={t:.2f} newton={nit} krylov={kit} rel err={err:.2e}")
        assert err <= 1e-8


def test_static_mixed_precision_converges_to_fp64_root():
    mesh, mats = cube_small(2)
    res = run_static_pair(mesh, mats, 2000.0, 3, precision=1, newton_atol=1e-12, newton_rtol=1e-12,
                          newton_stol=1e-15)
    for t, err, nit, kit in res:
        print(f"mixed t={t:.2f} newton={nit} krylov={kit} rel err vs FP64 oracle={err:.2e}")
        assert err <= 1e-8


@pytest.mark.parametrize("method,pc,reuse", [(0, 1, 0), (1, 1, 0), (2, 1, 0), (0, 2, 0), (1, 2, 0), (2, 2, 0),
                                             (2, 2, 1)],
                         ids=["gmres", "bicgstab_graph", "bicgstab", "gmres_schwarz", "bicgstab_graph_schwarz",
                              "bicgstab_schwarz", "bicgstab_schwarz_reused"])
def test_newmark_steps_match_oracle(method, pc, reuse):
    """Newmark (beta=1/4, gamma=1/2) with Holzapfel + active tension + cavity pressure;
    reuse: the preconditioner of each step's first tangent serves all its linear solves."""
    mesh, mats = lv_small()
    cpu = CpuMechanics(mesh, mats)
    plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
    dt = 0.005
    with MechanicsContext(mesh, mats, newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15,
                          krylov_method=method, preconditioner=pc, precond_reuse=reuse) as ctx:
        for k in range(1, 9):
            t = k * dt
            p = plug.apply(t, dt)[0]
            ctx.set_pressures([p])
            cpu.pressures[:] = p
            rc, st = ctx.newmark_step(t, dt)
            assert rc == 2, st.as_dict()
            cpu.newmark_step(t, dt)
            g = ctx.get_state()["nodes"]
            u = cpu.nodes - mesh.coords.ravel()
            err = np.linalg.norm(g - cpu.nodes) / np.linalg.norm(u)
            vg, vo = ctx.cavity_volumes()[0], cpu.cavity_volumes()[0]
            print(f"t={t:.3f} p={p:.0f} newton={st.newton_iterations} krylov={st.krylov_iterations} "
                  f"disp err={err:.2e} vol={vg * 1e6:.6f} ml err={abs(vg - vo) / vo:.2e}")
            assert err <= 1e-8
            assert abs(vg - vo) <= 1e-8 * abs(vo)


def test_full_size_lv_properties():
    """200k-element LV at its real size: size-independent checks plus sampled parity."""
    mesh = M.lv_mesh(200_000)
    mats = [MT.lv_holzapfel()]
    with MechanicsContext(mesh, mats) as ctx:
        # unloaded reference configuration before activation: no force
        r0 = ctx.residual(mesh.coords.ravel(), -1.0, include_pressure=False)
        x = deformed(mesh, 0.01)
        fe = ctx.element_residuals(x, 0.1)
        vals = ctx.tangent(x.ravel(), 0.1, include_pressure=False)
        K = ctx.bsr(vals).tocsr()
        rng = np.random.default_rng(5)
        v = rng.normal(size=ctx.n_dofs) * 1e-6
        v[np.repeat(mesh.bc, 3) != 0] = 0
        h = 1e-3
        rp = ctx.residual((x.ravel() + h * v), 0.1, include_pressure=False)
        rm = ctx.residual((x.ravel() - h * v), 0.1, include_pressure=False)
    # rounding of the volumetric penalty (kappa = 1e6 Pa) is all that remains at the reference state
    assert np.abs(r0).max() <= 1e-10 * np.abs(fe).max()
    Kv = K @ v
    fd = (rp - rm) / (2 * h)
    free = np.repeat(mesh.bc, 3) == 0
    print(f"200k LV: |K v - FD|/|K v| = {np.linalg.norm((Kv - fd)[free]) / np.linalg.norm(Kv[free]):.2e}")
    assert np.linalg.norm((Kv - fd)[free]) <= 1e-5 * np.linalg.norm(Kv[free])
    sample = np.sort(rng.choice(mesh.n_elements, 300, replace=False))
    sub = M.Mesh(coords=mesh.coords, elements=mesh.elements[sample], bases=mesh.bases[sample], bc=mesh.bc,
                 material=mesh.material[sample], activation=mesh.activation[sample])
    rc, fo, _ = O.residual_batch(oracle_elements(sub), x, mats, 0.1)
    err = np.abs(fe[sample] - fo) / np.abs(fo).max(axis=1, keepdims=True)
    assert err.max() <= 1e-12


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_newmark_steps_with_rayleigh_damping_match_oracle(precision):
    """Rayleigh damping (the reference LV settings' alpha = 500 1/s, beta_R = 0.01 s):
    every step against the oracle's damped Newmark step; the damping is active (the
    undamped oracle run is far away)."""
    mesh, mats = lv_small()
    damp = (500.0, 0.01)
    cpu = CpuMechanics(mesh, mats, rayleigh=damp)
    free = CpuMechanics(mesh, mats)
    plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
    dt = 0.005
    # mixed precision builds the damping matrix from FP64 element tangents and keeps the
    # residual in FP64: it converges to the same root
    tol = 1e-8
    with MechanicsContext(mesh, mats, newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15,
                          precision=precision, rayleigh_alpha=damp[0], rayleigh_beta=damp[1]) as ctx:
        for k in range(1, 9):
            t = k * dt
            p = plug.apply(t, dt)[0]
            ctx.set_pressures([p])
            cpu.pressures[:] = p
            free.pressures[:] = p
            rc, st = ctx.newmark_step(t, dt)
            assert rc == 2, st.as_dict()
            cpu.newmark_step(t, dt)
            free.newmark_step(t, dt)
            s = ctx.get_state()
            u = cpu.nodes - mesh.coords.ravel()
            err = np.linalg.norm(s["nodes"] - cpu.nodes) / np.linalg.norm(u)
            verr = np.linalg.norm(s["velocity"] - cpu.vel) / np.linalg.norm(cpu.vel)
            gap = np.linalg.norm(free.nodes - cpu.nodes) / np.linalg.norm(u)
            ke, ed = ctx.newmark_energies()
            print(f"t={t:.3f} newton={st.newton_iterations} krylov={st.krylov_iterations} disp err={err:.2e} "
                  f"vel err={verr:.2e} damped vs undamped={gap:.2e} kinetic={ke:.3e} J dissipated={ed:.3e} J")
            assert err <= tol and verr <= 100 * tol
            assert gap > 1e3 * tol
            assert abs(ke - cpu.kinetic_energy) <= 100 * tol * cpu.kinetic_energy
            assert abs(ed - cpu.damping_energy) <= 100 * tol * abs(cpu.damping_energy)


def test_spmv_layout_tuning_and_fixed_choice():
    """The context times the nine SpMV layouts on its first solve and keeps the
    fastest (cmg_spmv_tuning); spmv_choice fixes one without timing.  Both solve the
    same static step to the Newton tolerance."""
    mesh, mats = lv_small()
    tol = dict(newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15)
    res = {}
    for choice in (0, 3):
        with MechanicsContext(mesh, mats, spmv_choice=choice, **tol) as ctx:
            ctx.set_pressures([-300.0])
            rc, st = ctx.static_step(1.0, 1.0)
            assert rc == 2
            ms, layout, lanes = ctx.spmv_tuning()
            res[choice] = ctx.get_state()["nodes"]
        print(f"spmv_choice {choice}: layout {layout}, lanes {lanes}, ms {np.round(ms, 4)}")
        if choice == 0:
            assert np.all(ms > 0)
            k = int(np.argmin(ms))
            assert layout == k // 3 and lanes == ((4, 8, 16) if k < 6 else (16, 32, 64))[k % 3]
        else:
            assert layout == 0 and lanes == 16 and np.all(ms == 0)
    u = res[0] - mesh.coords.ravel()
    assert np.linalg.norm(res[3] - res[0]) <= 1e-8 * np.linalg.norm(u)


def test_initial_guess_overrides_the_predictor_once():
    """cmg_set_initial_guess: the next step's Newton starts from the given increment
    (guess_source 1) -- here the converged increment of the same step, so Newton needs
    at most one iteration and lands on the same root -- and the step after that uses the
    predictor again (guess_source 0)."""
    mesh, mats = lv_small()
    tol = dict(newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15)
    dt = 0.005
    plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
    p1, p2 = plug.apply(dt, dt)[0], plug.apply(2 * dt, dt)[0]
    with MechanicsContext(mesh, mats, **tol) as a:
        a.set_pressures([p1])
        rc, sa = a.newmark_step(dt, dt)
        assert rc == 2
        st_a = a.get_state()
    with MechanicsContext(mesh, mats, **tol) as b:
        b.set_pressures([p1])
        b.set_initial_guess(st_a["increment"])
        rc, sb = b.newmark_step(dt, dt)
        assert rc == 2 and sb.guess_source == 1, sb.as_dict()
        st_b = b.get_state()
        u = st_a["nodes"] - mesh.coords.ravel()
        err = np.linalg.norm(st_b["nodes"] - st_a["nodes"]) / np.linalg.norm(u)
        print(f"Newton iterations: predictor {sa.newton_iterations}, given guess {sb.newton_iterations}; "
              f"rel diff {err:.2e}")
        assert sb.newton_iterations <= 1 and sb.newton_iterations < sa.newton_iterations
        assert err <= 1e-10
        b.set_pressures([p2])
        rc, sc = b.newmark_step(2 * dt, dt)
        assert rc == 2 and sc.guess_source == 0, sc.as_dict()


def test_newton_reports_a_krylov_solve_without_update_as_failure():
    """A Krylov solve that returns no update (here: krylov_maxit = 0) ends the step as
    SNES_DIVERGED_LINEAR_SOLVE (-3); it must not read as a short step (step-norm
    convergence) with the residual still large."""
    mesh, mats = lv_small()
    plug = SV.sinus_pulse(-16000.0, 0.0, 0.8)
    dt = 0.005
    with MechanicsContext(mesh, mats) as ctx:
        for k in (1, 2):
            ctx.set_pressures([plug.apply(k * dt, dt)[0]])
            rc, st = ctx.newmark_step(k * dt, dt)
            assert rc == 2, st.as_dict()
        s = ctx.get_state()
    with MechanicsContext(mesh, mats, krylov_maxit=0) as ctx:
        ctx.set_state(nodes=s["nodes"], velocity=s["velocity"], acceleration=s["acceleration"])
        ctx.set_pressures([plug.apply(3 * dt, dt)[0]])
        rc, st = ctx.newmark_step(3 * dt, dt)       # non-zero predictor: d != 0
        print(st.as_dict())
        assert rc == -3
        assert st.converged_reason == -3
        assert st.initial_residual_norm > 1e-8


@pytest.mark.parametrize("precision,stall", [(0, 0), (1, 0), (0, 5)], ids=["fp64", "mixed", "fp64_stall5"])
def test_lv_beat_passes_activation_onset(precision, stall):
    """The 20k-element LV beat through activation onset (0-40 ms, Newmark 1/4, 1/2,
    dt = 1 ms with CBTiming halving): every step is accepted, on the absolute test
    (||r|| < Solver.Precision = 1e-8 N), on the relative one (||r|| < 1e-8 ||r_0||), or
    on the step-norm test with ||r|| below 1e-6 N (such steps are counted and printed).
    On the first LV generator's apex needles Newton failed here until dt reached its
    minimum, and a mixed-precision run once took a broken-down Krylov solve for
    convergence.  stall: stalled Newton solves end early (newton_stall) and are halved;
    the accepted steps meet the same tests."""
    mesh = M.lv_mesh(20_000)
    recs = []
    with MechanicsContext(mesh, [MT.lv_holzapfel()], precision=precision, newton_stall=stall) as ctx:
        SV.run(ctx, "newmark", SV.Timing(0.0, 0.04, 1e-3), SV.sinus_pulse(-16000.0, 0.0, 0.8),
               on_step=lambda rec, st: recs.append((rec.time, rec.dt, st.residual_norm, st.converged_reason,
                                                    st.initial_residual_norm)))
    halved = sum(1 for r in recs if r[1] < 0.999e-3)
    by_norm = [r for r in recs if r[3] == 4]
    print(f"{len(recs)} accepted steps, {halved} shorter than 1 ms, {len(by_norm)} on the step-norm test "
          f"(residuals {[f'{r[2]:.1e}' for r in by_norm]})")
    assert recs[-1][0] >= 0.04 - 1e-9
    for t, dt, rn, reason, r0 in recs:
        assert reason in (2, 3, 4), (t, dt, rn, reason)
        if reason == 2:
            assert rn < 1e-8, (t, dt, rn, reason)
        elif reason == 3:
            assert rn < 1e-8 * r0, (t, dt, rn, reason, r0)
        else:
            assert rn < 1e-6, (t, dt, rn, reason)


def land17_case(rate="ON"):
    mesh = M.lv_mesh(dims=(2, 4, 8))
    m = MT.with_land17(MT.lv_holzapfel(), rateDependancy=rate, CalciumTransientType="Brixius")
    rng = np.random.default_rng(17)
    E = mesh.n_elements
    st = np.empty((E, 8))
    st[:, 0] = rng.uniform(0.0, 0.1, E)       # XS
    st[:, 1] = rng.uniform(0.0, 0.1, E)       # XW
    st[:, 2] = rng.uniform(0.1, 0.6, E)       # TRPN
    st[:, 3] = rng.uniform(0.3, 0.9, E)       # TmBlocked
    st[:, 4] = rng.uniform(-0.05, 0.05, E)    # ZetaS
    st[:, 5] = rng.uniform(-0.05, 0.05, E)    # ZetaW
    st[:, 6] = rng.uniform(0.0, 0.02, E)      # Cd
    st[:, 7] = rng.uniform(0.98, 1.05, E)     # lambda at the commit
    return mesh, [m], st


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_land17_element_kernels_match_oracle(precision):
    """Land17 active tension (CBTensionModelLand17.cpp:298-372) from a committed cross-bridge
    state per element, rate dependence on: element residuals (<= 1e-12) and tangents (the
    analytic dT/dlambda against the complex-step oracle, <= 1e-12 FP64, 2e-6 mixed).  A
    context with a Land17 material runs one kernel shape whatever element_variant says."""
    mesh, mats, st = land17_case()
    x = deformed(mesh)
    t_commit, t = 0.05, 0.0515
    with MechanicsContext(mesh, mats, precision=precision) as ctx:
        ctx.set_land17_state(st.T.copy(), t_commit)
        fg = ctx.element_residuals(x, t)
        Kg = ctx.element_tangents(x, t).astype(np.float64)
        got, tc = ctx.land17_state()
    assert tc == t_commit and np.array_equal(got, st.T)
    l17 = O.Land17State(mesh, mats)
    l17.state[:] = st
    l17.t_prev = t_commit
    l17.configure(t)
    rc, fo, _ = O.residual_batch(oracle_elements(mesh), x, mats, t)
    rc2, Ko, _ = O.tangent_batch(oracle_elements(mesh), x, mats, t, kind="cs")
    assert rc == O.SUCCESS and rc2 == O.SUCCESS
    ef = (np.abs(fg - fo) / np.abs(fo).max(axis=1, keepdims=True)).max()
    rel = np.linalg.norm((Kg - Ko).reshape(len(Ko), -1), axis=1) / np.linalg.norm(Ko.reshape(len(Ko), -1), axis=1)
    print(f"Land17 precision {precision}: residual {ef:.2e}, tangent {rel.max():.2e}")
    assert ef <= 1e-12
    assert rel.max() <= (1e-12 if precision == 0 else 2e-6)


def test_land17_newmark_steps_and_state_commits_match_oracle():
    """Newmark (1/4, 1/2, dt = 1 ms) with Land17 tension (Coppini calcium) and cavity
    pressure: nodes (<= 1e-8) and the committed cross-bridge state (<= 1e-9) after every
    step against the oracle driver, which commits the state of each element's last
    quadrature point like the reference's SaveStateVariablesAsNeeded.  The cavity
    pressure ramps by 100 Pa per step, so every step moves the wall by far more than the
    Newton tolerances resolve (a sin^2 pulse moves it by micrometres in the first
    milliseconds, where the two roots at 1e-12 N differ by ~2e-8 of it)."""
    mesh = M.lv_mesh(dims=(2, 4, 8))
    mats = [MT.with_land17(MT.lv_holzapfel())]
    cpu = CpuMechanics(mesh, mats)
    dt = 1e-3
    with MechanicsContext(mesh, mats, newton_atol=1e-12, newton_rtol=1e-12, newton_stol=1e-15) as ctx:
        for k in range(1, 31):
            t = k * dt
            ctx.set_pressures([-100.0 * k])
            cpu.pressures[:] = -100.0 * k
            rc, st = ctx.newmark_step(t, dt)
            assert rc == 2, st.as_dict()
            cpu.newmark_step(t, dt)
            g = ctx.get_state()["nodes"]
            u = cpu.nodes - mesh.coords.ravel()
            err = np.linalg.norm(g - cpu.nodes) / np.linalg.norm(u)
            sg, tc = ctx.land17_state()
            so = cpu.land17.state.T
            serr = np.abs(sg - so).max() / max(np.abs(so[:7]).max(), 1e-300)
            if k % 5 == 0 or err > 1e-8 or serr > 1e-9:
                print(f"t={t:.3f} newton={st.newton_iterations} disp err={err:.2e} state err={serr:.2e} "
                      f"XS max={so[0].max():.2e}")
            assert tc == pytest.approx(t, abs=1e-15)
            assert err <= 1e-8
            assert serr <= 1e-9
        assert so[0].max() > 0       # cross-bridges formed
