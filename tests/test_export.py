"""VTK/PVD export (host side): the file format round trip, the reference's field
formulas on states with closed-form answers, and the collection bookkeeping, with a
stand-in for the context (no GPU)."""
import os

import numpy as np
import pytest

from cardiomechanics_amd import export as X
from cardiomechanics_amd import mesh as M


def rot(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.diag(r))


def test_vtu_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pts = rng.standard_normal((7, 3))
    cells = [np.arange(4), np.array([1, 2, 3, 4, 5, 6, 0, 1, 2, 3]), np.array([4, 5, 6])]
    types = [X.VTK_TETRA, X.VTK_QUADRATIC_TETRA, X.VTK_TRIANGLE]
    pdata = {"Displacement": rng.standard_normal((7, 3)), "Fixation": np.arange(7, dtype=np.int32)}
    cdata = {"J": rng.standard_normal(3), "Lambda": rng.standard_normal((3, 3))}
    path = tmp_path / "a.vtu"
    X.write_vtu(str(path), pts, cells, types, pdata, cdata)
    r = X.read_vtu(str(path))
    assert np.array_equal(r["points"], pts)
    assert np.array_equal(r["connectivity"], np.concatenate(cells))
    assert np.array_equal(r["offsets"], [4, 14, 17])
    assert np.array_equal(r["types"], types)
    for k, v in pdata.items():
        assert np.array_equal(r["point_data"][k], v)
    for k, v in cdata.items():
        assert np.array_equal(r["cell_data"][k], v)


def test_element_fields_closed_forms():
    """A pure stretch D in the fibre frame R: F = R D R^T.  Then F_loc = D, J = det D,
    Lambda = diag D, E = (D^2 - I)/2, C = D^2, and the exported fibre F_loc R e_f."""
    rng = np.random.default_rng(1)
    E = 5
    R = np.stack([rot(rng) for _ in range(E)])
    D = np.stack([np.diag(rng.uniform(0.8, 1.3, 3)) for _ in range(E)])
    F = np.einsum("eij,ejk,elk->eil", R, D, R)
    S = rng.standard_normal((E, 3, 3))
    f = X.element_fields(F, S, R)
    d = np.einsum("eii->ei", D)
    np.testing.assert_allclose(f["J"], d.prod(axis=1), rtol=1e-13)
    np.testing.assert_allclose(f["Lambda"], d, rtol=1e-13)
    np.testing.assert_allclose(f["E_ii"], 0.5 * (d * d - 1), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(f["E_ij"], 0.0, atol=1e-13)
    np.testing.assert_allclose(f["C_ii"], d * d, rtol=1e-13)
    np.testing.assert_allclose(f["f"], np.einsum("eij,ej->ei", D, R[:, :, 0]), rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(f["S_ii"], np.einsum("eii->ei", S))
    np.testing.assert_array_equal(f["S_ij"], S[:, [0, 0, 1], [1, 2, 2]])


class FakeCtx:
    def __init__(self, mesh):
        self.mesh = mesh
        self.t = 0.0

    def get_state(self):
        n = self.mesh.coords.size
        x = self.mesh.coords.ravel() * (1.0 + self.t)
        return {"nodes": x, "velocity": np.full(n, self.t), "acceleration": np.zeros(n),
                "increment": np.full(n, 0.5 * self.t)}

    def element_export(self):
        E = self.mesh.n_elements
        F = np.tile(np.eye(3) * (1.0 + self.t), (E, 1, 1))
        return F, np.zeros((E, 3, 3))


def test_exporter_collection_and_fields(tmp_path):
    mesh = M.cube_mesh(1)
    ctx = FakeCtx(mesh)
    params = {"Export.Prefix": str(tmp_path / "out" / "cube"), "Export.TimeStep": "0.1",
              "Export.Options.Jacobian": "true", "Export.Options.Velocity": "false"}
    ex = X.VTKExporter(ctx, mesh, params)
    files = []
    for k, t in enumerate([0.0, 0.05, 0.1, 0.2]):
        ctx.t = t
        if ex.due(t):
            files.append(ex.export(t))
    assert len(files) == 3                             # 0.05 s is not due (Export.TimeStep = 0.1)
    pvd = open(tmp_path / "out" / "cube.pvd").read()
    assert pvd.count("<DataSet") == 3 and 'file="cube_vtu/cube.2.vtu"' in pvd
    assert pvd.index('timestep="0.1"') < pvd.index('timestep="0.2"')
    r = X.read_vtu(files[-1])
    E, S = mesh.n_elements, mesh.surfaces.shape[0]
    assert len(r["types"]) == E + S
    assert list(np.unique(r["types"])) == [X.VTK_TRIANGLE, X.VTK_QUADRATIC_TETRA]
    assert "Velocity" not in r["point_data"] and "Acceleration" in r["point_data"]
    np.testing.assert_allclose(r["cell_data"]["J"][:E], 1.2 ** 3)
    assert np.all(r["cell_data"]["J"][E:] == 0.0)     # surface cells carry zeros
    np.testing.assert_allclose(r["point_data"]["AbsDisplacement"], 0.2 * mesh.coords, rtol=1e-12)
    assert os.path.exists(tmp_path / "out" / "cube_vtu" / "cube.0.vtu")


def test_missing_prefix_raises():
    with pytest.raises(ValueError):
        X.VTKExporter(FakeCtx(M.cube_mesh(1)), M.cube_mesh(1), {})
