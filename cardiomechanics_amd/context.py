"""Host-side handle on one GPU mechanics context (a thin layer over include/cmg.h).

`MechanicsContext` plays the role of the reference's CBSolver + CBFormulation pair
for one mesh: the element loop (CBFormulation::CalcNodalForces /
CalcNodalForcesJacobian, CBFormulation.h:30-41) and the solver step
(CBSolver::SolverStep, CBSolver.h:214) run in libcmg.so on the GPU; this class
only marshals numpy arrays across the C ABI and raises on error codes.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from .materials import Material
from .mesh import Mesh

STATUS_NAMES = {2: "SUCCESS", 0: "NOTHING_DONE", -2: "REPEAT", -3: "FAILED", -4: "INFINITIVE",
                -5: "NOT_A_NUMBER", -6: "CORRUPT_ELEMENT", -8: "SYSTEM_WRONG_STATE",
                -100: "ERR_INVALID", -101: "ERR_HIP", -102: "ERR_NOMEM"}


class CmgError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {STATUS_NAMES.get(code, code)} ({_lib.last_error()})")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def default_options(**kw) -> _lib.CmgOptions:
    o = _lib.CmgOptions()
    _lib.load().cmg_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def to_cmg_material(m: Material) -> _lib.CmgMaterial:
    r = _lib.CmgMaterial()
    r.law = int(m.law)
    r.ignore_corrupt = int(m.ignore_corrupt)
    r.tension_fn = int(m.tension_fn)
    r.density = float(m.density)
    for i in range(12):
        r.params[i] = float(m.params[i]) if i < len(m.params) else 0.0
        r.fn[i] = float(m.fn[i]) if i < len(m.fn) else 0.0
    r.tension_max = float(m.tension_max)
    return r


class MechanicsContext:
    def __init__(self, mesh: Mesh, materials, **options):
        self.L = _lib.load()
        self.mesh = mesh
        self.materials = list(materials)
        self.opt = default_options(**options)
        self.fp32 = self.opt.precision == _lib.PREC_MIXED
        E, npe = mesh.n_elements, mesh.npe
        # keep the arrays alive while the context is created
        self._coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
        self._elements = np.ascontiguousarray(mesh.elements, dtype=np.int32)
        nb = 5 if npe == 10 else 1
        self._bases = np.ascontiguousarray(mesh.bases.reshape(E, nb, 9), dtype=np.float64)
        self._matidx = np.ascontiguousarray(mesh.material, dtype=np.int32)
        self._tact = np.ascontiguousarray(mesh.activation, dtype=np.float64)
        self._bc = np.ascontiguousarray(mesh.bc, dtype=np.uint8)
        self._surf = np.ascontiguousarray(mesh.surfaces, dtype=np.int32)
        groups = np.ascontiguousarray(mesh.surface_id, dtype=np.int32)
        self.group_ids = np.unique(groups) if groups.size else np.zeros(0, np.int32)
        self._sgroup = np.ascontiguousarray(np.searchsorted(self.group_ids, groups), dtype=np.int32)
        cm = _lib.CmgMesh()
        cm.n_nodes, cm.n_elements = mesh.n_nodes, E
        cm.n_surfaces = self._surf.shape[0]
        cm.nodes_per_element = npe
        cm.nodes_per_surface = self._surf.shape[1] if cm.n_surfaces else 0
        cm.n_groups = max(1, len(self.group_ids))
        cm.ref_coords, cm.elements, cm.bases = (self._coords.ctypes.data, self._elements.ctypes.data,
                                                self._bases.ctypes.data)
        cm.material_index, cm.activation_time, cm.bc = (self._matidx.ctypes.data, self._tact.ctypes.data,
                                                        self._bc.ctypes.data)
        cm.surfaces = self._surf.ctypes.data if cm.n_surfaces else None
        cm.surface_group = self._sgroup.ctypes.data if cm.n_surfaces else None
        self.n_groups = cm.n_groups
        mats = (_lib.CmgMaterial * len(self.materials))(*[to_cmg_material(m) for m in self.materials])
        h = ctypes.c_void_p()
        rc = self.L.cmg_create(ctypes.byref(cm), mats, len(self.materials), ctypes.byref(self.opt), ctypes.byref(h))
        if rc != _lib.SUCCESS:
            raise CmgError(rc, "cmg_create")
        self.h = h
        n_nodes, n_el, n_dofs, nnzb = (ctypes.c_int64() for _ in range(4))
        self._check(self.L.cmg_sizes(self.h, ctypes.byref(n_nodes), ctypes.byref(n_el), ctypes.byref(n_dofs),
                                     ctypes.byref(nnzb)), "cmg_sizes")
        self.n_nodes, self.n_dofs, self.nnzb = n_nodes.value, n_dofs.value, nnzb.value
        self.n_elements = n_el.value
        self.rowptr = np.zeros(self.n_nodes + 1, np.int64)
        self.colind = np.zeros(self.nnzb, np.int32)
        self._check(self.L.cmg_pattern(self.h, _ptr(self.rowptr), _ptr(self.colind)), "cmg_pattern")
        for i, m in enumerate(self.materials):
            if getattr(m, "land17", None) is not None:
                p = np.ascontiguousarray(m.land17, dtype=np.float64)
                self._check(self.L.cmg_set_land17_parameters(self.h, i, _ptr(p)), "cmg_set_land17_parameters")

    # -- plumbing -------------------------------------------------------------
    def _check(self, rc, where):
        if rc != _lib.SUCCESS:
            raise CmgError(rc, where)
        return rc

    def close(self):
        # batches and networks made on this context go first
        for dep in list(getattr(self, "_dependents", [])):
            obj = dep()
            if obj is not None:
                obj.close()
        self._dependents = []
        if getattr(self, "h", None):
            self.L.cmg_destroy(self.h)
            self.h = None

    def _add_dependent(self, obj):
        import weakref

        if not hasattr(self, "_dependents"):
            self._dependents = []
        self._dependents.append(weakref.ref(obj))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def value_dtype(self):
        return np.float32 if self.fp32 else np.float64

    def bsr(self, values) -> sp.bsr_matrix:
        v = np.asarray(values).reshape(self.nnzb, 3, 3)
        return sp.bsr_matrix((v, self.colind, self.rowptr), shape=(self.n_dofs, self.n_dofs))

    # -- element loop ------------------------------------------------------------
    def set_pressures(self, pressures):
        p = np.ascontiguousarray(np.resize(np.asarray(pressures, np.float64), self.n_groups))
        self._check(self.L.cmg_set_pressures(self.h, _ptr(p)), "cmg_set_pressures")

    def residual(self, x, time=0.0, include_pressure=True):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        f = np.zeros(self.n_dofs)
        self._check(self.L.cmg_residual(self.h, _ptr(x), time, int(include_pressure), _ptr(f)), "cmg_residual")
        return f

    def tangent(self, x, time=0.0, include_pressure=True):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        v = np.zeros(self.nnzb * 9, self.value_dtype)
        self._check(self.L.cmg_tangent(self.h, _ptr(x), time, int(include_pressure), _ptr(v)), "cmg_tangent")
        return v

    def element_residuals(self, x, time=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        nd = 3 * self.mesh.npe
        fe = np.zeros((self.mesh.n_elements, nd))
        self._check(self.L.cmg_element_residuals(self.h, _ptr(x), time, _ptr(fe)), "cmg_element_residuals")
        return fe

    def element_tangents(self, x, time=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        nd = 3 * self.mesh.npe
        Ke = np.zeros((self.mesh.n_elements, nd, nd), self.value_dtype)
        self._check(self.L.cmg_element_tangents(self.h, _ptr(x), time, _ptr(Ke)), "cmg_element_tangents")
        return Ke

    def surface_forces(self, x, pressures):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        p = np.ascontiguousarray(np.resize(np.asarray(pressures, np.float64), self.n_groups))
        f = np.zeros(self.n_dofs)
        self._check(self.L.cmg_surface_forces(self.h, _ptr(x), _ptr(p), _ptr(f)), "cmg_surface_forces")
        return f

    def cavity_volumes(self, x=None):
        v = np.zeros(self.n_groups)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64).ravel()
        self._check(self.L.cmg_cavity_volumes(self.h, _ptr(xa), _ptr(v)), "cmg_cavity_volumes")
        return v

    def element_export(self, x=None):
        """(F at the centroid, global frame (E,3,3); passive PK2 averaged over the
        quadrature points, fibre frame (E,3,3)) at x or the current nodes"""
        out = np.zeros((18, self.n_elements))
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64).ravel()
        self._check(self.L.cmg_element_export(self.h, _ptr(xa), _ptr(out)), "cmg_element_export")
        return out[:9].T.reshape(-1, 3, 3), out[9:].T.reshape(-1, 3, 3)

    def mass_matrix(self):
        v = np.zeros(self.nnzb * 9)
        self._check(self.L.cmg_mass_matrix(self.h, _ptr(v)), "cmg_mass_matrix")
        return v

    def spmv(self, values, x, lanes=None, layout=0):
        """y = A x; lanes/layout select the kernel (layout 0 block, 1 chunked, 2 flat)"""
        values = np.ascontiguousarray(values, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n_dofs)
        if lanes is None:
            self._check(self.L.cmg_spmv(self.h, _ptr(values), _ptr(x), _ptr(y)), "cmg_spmv")
        else:
            self._check(self.L.cmg_spmv_layout(self.h, _ptr(values), _ptr(x), _ptr(y), lanes, int(layout)),
                        "cmg_spmv_layout")
        return y

    def spmv_tuning(self):
        """(ms per product of the 9 layouts, chosen layout, lanes) after the first solve"""
        ms = np.zeros(9)
        lay, lanes = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.L.cmg_spmv_tuning(self.h, _ptr(ms), ctypes.byref(lay), ctypes.byref(lanes)),
                    "cmg_spmv_tuning")
        return ms, lay.value, lanes.value

    def linear_solve(self, values, b, x0=None, rtol=1e-10):
        values = np.ascontiguousarray(values, dtype=self.value_dtype)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n_dofs) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        st = _lib.CmgSolveStats()
        rc = self.L.cmg_linear_solve(self.h, _ptr(values), _ptr(b), _ptr(x), rtol, ctypes.byref(st))
        if rc not in (_lib.SUCCESS, _lib.FAILED):
            raise CmgError(rc, "cmg_linear_solve")
        return x, st

    # -- solver state / steps -------------------------------------------------------
    def set_state(self, nodes=None, velocity=None, acceleration=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
                for a in (nodes, velocity, acceleration)]
        self._check(self.L.cmg_set_state(self.h, *[_ptr(a) for a in arrs]), "cmg_set_state")

    def get_state(self):
        out = [np.zeros(self.n_dofs) for _ in range(4)]
        self._check(self.L.cmg_get_state(self.h, *[_ptr(a) for a in out]), "cmg_get_state")
        return dict(zip(("nodes", "velocity", "acceleration", "increment"), out))

    def land17_state(self):
        """Land17 state [8][E] (XS XW TRPN TmBlocked ZetaS ZetaW Cd lambda) and its commit time"""
        st = np.zeros((8, self.n_elements))
        t = ctypes.c_double()
        self._check(self.L.cmg_get_land17_state(self.h, _ptr(st), ctypes.byref(t)), "cmg_get_land17_state")
        return st, t.value

    def set_land17_state(self, state, commit_time):
        st = np.ascontiguousarray(state, dtype=np.float64).reshape(8, self.n_elements)
        self._check(self.L.cmg_set_land17_state(self.h, _ptr(st), float(commit_time)), "cmg_set_land17_state")

    def set_initial_guess(self, increment):
        inc = np.ascontiguousarray(increment, dtype=np.float64).ravel()
        self._check(self.L.cmg_set_initial_guess(self.h, _ptr(inc)), "cmg_set_initial_guess")

    def static_step(self, time, dt):
        st = _lib.CmgStepStats()
        rc = self.L.cmg_static_step(self.h, time, dt, ctypes.byref(st))
        if rc not in (_lib.SUCCESS, _lib.FAILED):
            raise CmgError(rc, "cmg_static_step")
        return rc, st

    def newmark_step(self, time, dt):
        st = _lib.CmgStepStats()
        rc = self.L.cmg_newmark_step(self.h, time, dt, ctypes.byref(st))
        if rc not in (_lib.SUCCESS, _lib.FAILED):
            raise CmgError(rc, "cmg_newmark_step")
        return rc, st

    def newmark_energies(self):
        """(kinetic energy after the last committed step, accumulated Rayleigh damping
        dissipation), as the reference's Newmark corrector records them."""
        ke, ed = ctypes.c_double(), ctypes.c_double()
        rc = self.L.cmg_newmark_energies(self.h, ctypes.byref(ke), ctypes.byref(ed))
        if rc != _lib.SUCCESS:
            raise CmgError(rc, "cmg_newmark_energies")
        return ke.value, ed.value

    def newmark_solve(self, time, dt, from_trial=False):
        """Predictor + Newton of a Newmark step, kept pending (see commit_step)."""
        st = _lib.CmgStepStats()
        rc = self.L.cmg_newmark_solve(self.h, time, dt, int(bool(from_trial)), ctypes.byref(st))
        if rc not in (_lib.SUCCESS, _lib.FAILED):
            raise CmgError(rc, "cmg_newmark_solve")
        return rc, st

    def trial_volumes(self):
        """cavity volume per pressure group at the pending step's configuration"""
        v = np.zeros(self.n_groups)
        self._check(self.L.cmg_trial_volumes(self.h, _ptr(v)), "cmg_trial_volumes")
        return v

    def commit_step(self):
        self._check(self.L.cmg_commit_step(self.h), "cmg_commit_step")


class ParameterBatch:
    """S material-parameter sets on the mesh of one `MechanicsContext` (cmg_batch_* in
    include/cmg.h): every set runs the quasi-static step of the context, all sets in
    the same kernel launches.  `material_sets` is a list of S lists of materials, each
    in the context's material order."""

    def __init__(self, ctx: MechanicsContext, material_sets, use_graph=True, spmv_lanes=0, element_variant=0,
                 cg_chunk=16):
        self.ctx = ctx
        self.L = ctx.L
        self.S = len(material_sets)
        flat = [to_cmg_material(m) for ms in material_sets for m in ms]
        if len(flat) != self.S * len(ctx.materials):
            raise ValueError("every set needs one material per context material")
        arr = (_lib.CmgMaterial * len(flat))(*flat)
        h = ctypes.c_void_p()
        rc = self.L.cmg_batch_create(ctx.h, self.S, arr, ctypes.byref(h))
        if rc != _lib.SUCCESS:
            raise CmgError(rc, "cmg_batch_create")
        self.h = h
        ctx._add_dependent(self)
        ctx._check(self.L.cmg_batch_configure(self.h, int(use_graph), spmv_lanes, element_variant, cg_chunk),
                   "cmg_batch_configure")

    def close(self):
        if getattr(self, "h", None):
            self.L.cmg_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = _lib.CmgBatchInfo()
        self.ctx._check(self.L.cmg_batch_info_get(self.h, ctypes.byref(i)), "cmg_batch_info_get")
        return i.as_dict()

    def static_step(self, time, dt):
        st = (_lib.CmgStepStats * self.S)()
        rc = self.L.cmg_batch_static_step(self.h, time, dt, st)
        if rc not in (_lib.SUCCESS, _lib.FAILED):
            raise CmgError(rc, "cmg_batch_static_step")
        return rc, list(st)

    def nodes(self):
        out = np.zeros((self.S, self.ctx.n_dofs))
        self.ctx._check(self.L.cmg_batch_get_nodes(self.h, _ptr(out)), "cmg_batch_get_nodes")
        return out

    def set_nodes(self, nodes):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(nodes, np.float64).reshape(-1, self.ctx.n_dofs),
                                                 (self.S, self.ctx.n_dofs)))
        self.ctx._check(self.L.cmg_batch_set_nodes(self.h, _ptr(a)), "cmg_batch_set_nodes")

    def residual(self, x, time=0.0):
        x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(self.S, self.ctx.n_dofs))
        f = np.zeros_like(x)
        self.ctx._check(self.L.cmg_batch_residual(self.h, _ptr(x), time, _ptr(f)), "cmg_batch_residual")
        return f

    def cavity_volumes(self):
        v = np.zeros((self.S, self.ctx.n_groups))
        self.ctx._check(self.L.cmg_batch_cavity_volumes(self.h, _ptr(v)), "cmg_batch_cavity_volumes")
        return v
