"""VTK export of simulation results (host side): a ParaView collection (.pvd) and one
unstructured-grid file (.vtu) per exported time, as CBModelExporterVTK writes them
(mechanics/src/CBModel/CBModelExporterVTK.cpp:29-499), with the per-element fields of
CBSolver::Export (mechanics/src/CBSolver/CBSolver.cpp:538-1164) and the Newmark node
fields (CBSolverNewmarkBeta.cpp:800-827).

Layout: Export.Prefix = "<dir>/<prefix>" gives <dir>/<prefix>.pvd and
<dir>/<prefix>_vtu/<prefix>.<counter>.vtu; a DataSet entry is inserted before
</Collection> at every export.  Cells are the solid elements (T4 -> VTK_TETRA, T10 ->
VTK_QUADRATIC_TETRA with the element's own node order, as the reference passes it)
followed by the pressure-surface elements (T3 -> VTK_TRIANGLE, T6 ->
VTK_QUADRATIC_TRIANGLE); element fields are zero on the surface cells.  Arrays are
base64-encoded little-endian binary with a UInt64 byte-count header.

The per-element state (deformation gradient at the centroid, passive PK2 stress
averaged over the quadrature points) comes from the GPU (cmg_element_export); the
derived fields follow the reference's formulas:
  F_loc = B0^T F B0 (B0 the centroid fibre frame, CBElementSolidT10.cpp:307-322)
  J = det F_loc;  Lambda = column norms of F_loc;  C = F_loc^T F_loc;  E = (C - I)/2
  f, s, n = F_loc B0[:, k] (the reference multiplies the fibre-frame tensor with the
            global-frame basis vector, CBSolver.cpp:713)
  S_ii / S_ij, E_ii / E_ij, C_ii / C_ij = (00, 11, 22) / (01, 02, 12)
Export.Options.<Name> switches each field with the reference's defaults.
"""
from __future__ import annotations

import base64
import os

import numpy as np

from .settings import as_bool

VTK_TETRA, VTK_QUADRATIC_TETRA, VTK_TRIANGLE, VTK_QUADRATIC_TRIANGLE = 10, 24, 5, 22

# Export.Options defaults of the reference (CBSolver.cpp, CBSolverNewmarkBeta.cpp, CBModelExporterVTK.cpp)
DEFAULT_OPTIONS = {
    "Material": True, "Fixation": True, "Fiber": True, "Sheet": True, "SheetNormal": True, "Lambda": True,
    "Jacobian": False, "Deformation": False, "PK2Stress": True, "GreenLagrangeStrain": True,
    "LocalActivationTime": False, "Velocity": True, "Acceleration": True, "Displacement": True,
    "AbsDisplacement": True,
}


def _b64(arr: np.ndarray) -> str:
    raw = np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
    return base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")


def _data_array(name, arr, ncomp=1):
    types = {np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
             np.dtype(np.uint8): "UInt8"}
    arr = np.asarray(arr)
    return (f'        <DataArray type="{types[arr.dtype]}" Name="{name}" NumberOfComponents="{ncomp}" '
            f'format="binary">\n          {_b64(arr.ravel())}\n        </DataArray>\n')


def cell_arrays(blocks):
    """flat connectivity and offsets of cell blocks, each an (M_k, n_k) array of node ids"""
    blocks = [np.asarray(b, np.int64) for b in blocks if len(b)]
    if not blocks:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    conn = np.concatenate([b.ravel() for b in blocks])
    offsets = np.cumsum(np.concatenate([np.full(b.shape[0], b.shape[1], np.int64) for b in blocks]))
    return conn, offsets


def write_vtu(path, points, cells, cell_types, point_data=None, cell_data=None):
    """points (N,3); cells: list of connectivity rows (ragged allowed) or a prepared
    (connectivity, offsets) pair from cell_arrays; cell_types (M,); point_data /
    cell_data: {name: array (N,) / (N,k)}"""
    if isinstance(cells, tuple):
        conn, offsets = cells
    else:
        conn = np.concatenate([np.asarray(c, np.int64).ravel() for c in cells]) if cells else np.zeros(0, np.int64)
        offsets = np.cumsum([len(c) for c in cells]).astype(np.int64)
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n',
           '  <UnstructuredGrid>\n',
           f'    <Piece NumberOfPoints="{len(points)}" NumberOfCells="{len(cell_types)}">\n']
    for tag, data in (("PointData", point_data or {}), ("CellData", cell_data or {})):
        out.append(f"      <{tag}>\n")
        for name, arr in data.items():
            arr = np.asarray(arr)
            out.append(_data_array(name, arr, 1 if arr.ndim == 1 else arr.shape[1]))
        out.append(f"      </{tag}>\n")
    out.append("      <Points>\n")
    out.append(_data_array("Points", np.asarray(points, np.float64), 3))
    out.append("      </Points>\n      <Cells>\n")
    out.append(_data_array("connectivity", conn))
    out.append(_data_array("offsets", offsets))
    out.append(_data_array("types", np.asarray(cell_types, np.uint8)))
    out.append("      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as fh:
        fh.write("".join(out))


def read_vtu(path):
    """Reader of the files write_vtu produces: {"points", "connectivity", "offsets",
    "types", "point_data": {...}, "cell_data": {...}}."""
    import xml.etree.ElementTree as ET
    types = {"Float64": np.float64, "Int32": np.int32, "Int64": np.int64, "UInt8": np.uint8}
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")

    def arr(da):
        raw = base64.b64decode(da.text.strip())
        n = int(np.frombuffer(raw[:8], "<u8")[0])
        a = np.frombuffer(raw[8:8 + n], np.dtype(types[da.get("type")]).newbyteorder("<"))
        k = int(da.get("NumberOfComponents", "1"))
        return a.reshape(-1, k) if k > 1 else a

    out = {"point_data": {}, "cell_data": {}}
    for tag, key in (("PointData", "point_data"), ("CellData", "cell_data")):
        for da in piece.find(tag):
            out[key][da.get("Name")] = arr(da)
    out["points"] = arr(piece.find("Points/DataArray"))
    for da in piece.find("Cells"):
        out[da.get("Name")] = arr(da)
    return out


def element_fields(F_glob, S_pk2, bases0):
    """reference field formulas from the centroid F (global frame), the averaged passive
    PK2 (fibre frame) and the centroid fibre frame B0 (columns f, s, n)"""
    B = bases0
    F = np.einsum("eji,ejk,ekl->eil", B, F_glob, B)          # B^T F B
    C = np.einsum("eji,ejk->eik", F, F)
    E = 0.5 * (C - np.eye(3))
    iu = ([0, 1, 2], [0, 1, 2]), ([0, 0, 1], [1, 2, 2])
    out = {
        "J": np.linalg.det(F),
        "Lambda": np.linalg.norm(F, axis=1),                  # norms of the columns
        "f": np.einsum("eij,ej->ei", F, B[:, :, 0]),
        "s": np.einsum("eij,ej->ei", F, B[:, :, 1]),
        "n": np.einsum("eij,ej->ei", F, B[:, :, 2]),
    }
    for key, M in (("S", S_pk2), ("E", E), ("C", C)):
        out[key + "_ii"] = M[:, iu[0][0], iu[0][1]]
        out[key + "_ij"] = M[:, iu[1][0], iu[1][1]]
    return out


class VTKExporter:
    """CBModelExporterVTK with CBSolver::Export's fields for a MechanicsContext."""

    def __init__(self, ctx, mesh, params: dict | None = None, prefix: str | None = None):
        params = params or {}
        prefix = prefix if prefix is not None else str(params.get("Export.Prefix", ""))
        if not prefix:
            raise ValueError("no export directory and prefix given (Export.Prefix)")
        self.ctx, self.mesh = ctx, mesh
        self.time_step = float(params.get("Export.TimeStep", 1e-3))
        self.options = {k: as_bool(params.get("Export.Options." + k, d)) for k, d in DEFAULT_OPTIONS.items()}
        self.dir, self.prefix = os.path.split(prefix)
        self.dir = self.dir or "."
        self.counter = 0
        self.last_time = None
        os.makedirs(os.path.join(self.dir, self.prefix + "_vtu"), exist_ok=True)
        with open(self.pvd_path, "w") as fh:
            fh.write('<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
                     "\t<Collection>\n\t</Collection>\n</VTKFile>\n")
        el = mesh.elements
        types = [np.full(len(el), VTK_QUADRATIC_TETRA if el.shape[1] == 10 else VTK_TETRA, np.uint8)]
        blocks = [el]
        if mesh.surfaces.shape[0]:
            st = VTK_QUADRATIC_TRIANGLE if mesh.surfaces.shape[1] == 6 else VTK_TRIANGLE
            blocks.append(mesh.surfaces)
            types.append(np.full(mesh.surfaces.shape[0], st, np.uint8))
        self._cells = cell_arrays(blocks)
        self._types = np.concatenate(types)
        self.n_solid = len(el)

    @property
    def pvd_path(self):
        return os.path.join(self.dir, self.prefix + ".pvd")

    def due(self, time) -> bool:
        """the reference's export rule (CBSolver.cpp:1808-1810): the start time, then
        whenever Export.TimeStep has passed since the last export"""
        return self.last_time is None or time - self.last_time >= (1 - 1e-12) * self.time_step

    def _pad(self, a):
        """element array extended by zeros for the surface cells"""
        ns = len(self._types) - self.n_solid
        if ns == 0:
            return a
        pad = np.zeros((ns,) + a.shape[1:], a.dtype)
        return np.concatenate([a, pad], axis=0)

    def export(self, time, activation_time=None):
        st = self.ctx.get_state()
        nodes = st["nodes"].reshape(-1, 3)
        opt = self.options
        pdata = {"PointID": np.arange(len(nodes), dtype=np.int32)}
        if opt["Fixation"]:
            pdata["Fixation"] = self.mesh.bc.astype(np.int32)
        if opt["Displacement"]:
            pdata["Displacement"] = st["increment"].reshape(-1, 3)
        if opt["AbsDisplacement"]:
            pdata["AbsDisplacement"] = nodes - self.mesh.coords
        if opt["Velocity"]:
            pdata["Velocity"] = st["velocity"].reshape(-1, 3)
        if opt["Acceleration"]:
            pdata["Acceleration"] = st["acceleration"].reshape(-1, 3)
        Fg, S = self.ctx.element_export()
        fields = element_fields(Fg, S, self.mesh.bases[:, 0])
        cdata = {"CellID": np.arange(len(self._types), dtype=np.int32)}
        if opt["Material"]:
            cdata["Material"] = self._pad(self.mesh.material.astype(np.int32))
        wanted = {"Jacobian": ["J"], "Lambda": ["Lambda"], "Fiber": ["f"], "Sheet": ["s"], "SheetNormal": ["n"],
                  "PK2Stress": ["S_ii", "S_ij"], "GreenLagrangeStrain": ["E_ii", "E_ij"],
                  "Deformation": ["C_ii", "C_ij"]}
        for key, names in wanted.items():
            if opt[key]:
                for nm in names:
                    cdata[nm] = self._pad(np.ascontiguousarray(fields[nm]))
        if opt["LocalActivationTime"] and time == 0.0:
            act = self.mesh.activation if activation_time is None else activation_time
            cdata["LocalActivationTime"] = self._pad(np.asarray(act, np.float64))
        name = f"{self.prefix}.{self.counter}.vtu"
        write_vtu(os.path.join(self.dir, self.prefix + "_vtu", name), nodes, self._cells, self._types, pdata, cdata)
        with open(self.pvd_path) as fh:
            lines = fh.read().splitlines()
        out = []
        for ln in lines:
            if "</Collection>" in ln:
                out.append(f'\t\t<DataSet timestep="{time:.6g}" part="0" group="" '
                           f'file="{self.prefix}_vtu/{name}"/>')
            out.append(ln)
        with open(self.pvd_path, "w") as fh:
            fh.write("\n".join(out) + "\n")
        self.counter += 1
        self.last_time = time
        return os.path.join(self.dir, self.prefix + "_vtu", name)
