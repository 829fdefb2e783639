"""CPU FP64 oracle of the mechanics hot path -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as the checker.  The product library (cardiomechanics_amd)
never imports it.  `liboracle.so` is built from oracle.cpp by oracle/Makefile.

Parity status of each oracle function (what pins it; see tests/test_oracle_*.py):
  analytic functions        closed forms (sin^2 identity, linear ramp, windows)
  Guccione PK2              S = dW/dE by complex step of the energy; S(I) = 0; symmetry
  Holzapfel PK2             S - 2 dW/dC (complex step) = -(a e^{b(I1-3)} I1/3) C^-1 (the
                            Jm = 1 quirk) computed independently; S(I) = 0
  active stress             S_a = 2 d(T sqrt(I4f))/dC by complex step; F = I gives T
  T10/T4 shape functions    sum dN = 0, sum X (x) dN = I, detJ/6 = volume, patch test
  T10/T4 residual           f = dPi/dx (complex step of the element energy, conservative
                            laws); rigid motions give f = 0; sum_a f_a = 0
  T10/T4 FD tangent         = transpose (T10) / plain (T4) of the complex-step Jacobian
                            to FD noise
  T10/T4 complex-step tangent  exact to rounding by construction; symmetric for
                            conservative laws
  mass matrices             sum = rho V per direction; lumped rho V / 10 (T10), the row
                            sums of the consistent mass (T4)
  T3/T6 forces, volume      cube face force = p A n; closed-surface volume independent
                            of the reference point; tetra volume closed form
  Newton / Newmark drivers  linear spring-mass closed form; residual at convergence
  SinusDriver, DoubleHill   values at chosen times against their closed forms
  TanH tension              limits, window, periodicity and saturation; complex-step
                            element tangent against the reference-style FD tangent
  Land17 tension            closed-form steady state of the ODE system at constant
                            calcium and stretch; closed-form first step from the initial
                            state; calcium lookup rules; dT/dlambda by central differences
  Rayleigh damping          C w = alpha M w + beta_R (reference-layout FD Jacobian) w;
                            the damped oscillator in the modal limit
  CPU reference solver      the FD epsilon (1 + |d|_2) eps and the Newmark / static steps
                            (cpu_solver.cpp) against the drivers' roots
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

LAW_GUCCIONE = 0
LAW_HOLZAPFEL = 1
FN_NONE, FN_LINEAR, FN_SINUS, FN_SINUSDRIVER, FN_DOUBLEHILL = 0, 1, 2, 3, 4
FN_LAND17 = 6
SUCCESS, FAILED, CORRUPT_ELEMENT = 2, -3, -6


class OrcMaterial(ctypes.Structure):
    _fields_ = [("law", ctypes.c_int32), ("ignore_corrupt", ctypes.c_int32),
                ("tension_fn", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("density", ctypes.c_double), ("p", ctypes.c_double * 12),
                ("tension_max", ctypes.c_double), ("fn", ctypes.c_double * 12)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _LIB_PATH


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I32 = ctypes.POINTER(ctypes.c_int32)
_M = ctypes.POINTER(OrcMaterial)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_analytic": (ctypes.c_double, [ctypes.c_int32, _D, ctypes.c_double, ctypes.c_double]),
            "orc_pk2": (ctypes.c_int32, [_M, _D, _D]),
            "orc_energy": (ctypes.c_int32, [_M, _D, _D]),
            "orc_energy_dF_cs": (ctypes.c_int32, [_M, _D, _D]),
            "orc_pk2_dF_cs": (ctypes.c_int32, [_M, _D, _D]),
            "orc_active_stress": (None, [_D, ctypes.c_double, _D]),
            "orc_tanh_tension": (ctypes.c_double, [_M, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "orc_active_stress_model": (None, [_M, _D, ctypes.c_double, ctypes.c_double, _D]),
            "orc_t10_shape": (None, [_D, _D, _D]),
            "orc_shape_batch": (None, [ctypes.c_int32, ctypes.c_int64, _I32, _D, _D, _D]),
            "orc_cpu_solver_create": (ctypes.c_void_p, [ctypes.c_int32, ctypes.c_int64, _I32, ctypes.c_int64, _D, _D,
                                                        _D, _D, _U8, _M, ctypes.c_int32, _I32, _D, ctypes.c_int32,
                                                        ctypes.c_int64, _I32, _I32, ctypes.c_int32]),
            "orc_cpu_solver_destroy": (None, [ctypes.c_void_p]),
            "orc_cpu_set_pressures": (None, [ctypes.c_void_p, _D]),
            "orc_cpu_get_state": (None, [ctypes.c_void_p, _D, _D, _D]),
            "orc_cpu_set_state": (None, [ctypes.c_void_p, _D, _D, _D]),
            "orc_cpu_newmark_step": (ctypes.c_int32, [ctypes.c_void_p] + [ctypes.c_double] * 8 + [ctypes.c_int32, _D]),
            "orc_cpu_static_step": (ctypes.c_int32, [ctypes.c_void_p] + [ctypes.c_double] * 6 + [ctypes.c_int32, _D]),
            "orc_fd_epsilon": (ctypes.c_double, [_D, ctypes.c_int64, ctypes.c_double]),
            "orc_land17_configure": (None, [_D, _D, ctypes.c_double]),
            "orc_land17_cai": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double]),
            "orc_land17_step": (ctypes.c_double, [_D, ctypes.c_double, _D, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_double, _D, _D]),
            "orc_land17_commit_batch": (None, [ctypes.c_int32, ctypes.c_int64, _I32, _D, _D, _D, _M, _I32, _D, _D,
                                               ctypes.c_double, ctypes.c_double, _D]),
            "orc_mass_batch": (None, [ctypes.c_int32, ctypes.c_int64, _I32, _D, _D, _U8, ctypes.c_int32, _D]),
            "orc_t10_residual": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _U8, _M, ctypes.c_double,
                                                  ctypes.c_double, _D]),
            "orc_t10_fd_tangent": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _U8, _M, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double, _D]),
            "orc_t10_cs_tangent": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _U8, _M, ctypes.c_double,
                                                    ctypes.c_double, _D]),
            "orc_t10_mass_consistent": (None, [ctypes.c_double, ctypes.c_double, _U8, _D]),
            "orc_t10_mass_lumped": (None, [ctypes.c_double, ctypes.c_double, _U8, _D]),
            "orc_t10_energy": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _M, _D]),
            "orc_t10_energy_grad_cs": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _M, _D]),
            "orc_t4_shape": (None, [_D, _D, _D]),
            "orc_t4_residual": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _U8, _M, ctypes.c_double,
                                                 ctypes.c_double, _D]),
            "orc_t4_fd_tangent": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _U8, _M, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double, _D]),
            "orc_t4_cs_tangent": (ctypes.c_int32, [_D, _D, _D, ctypes.c_double, _U8, _M, ctypes.c_double,
                                                   ctypes.c_double, _D]),
            "orc_t4_mass_consistent": (None, [ctypes.c_double, ctypes.c_double, _U8, _D]),
            "orc_t4_mass_lumped": (None, [ctypes.c_double, ctypes.c_double, _U8, _D]),
            "orc_t3_pressure_forces": (None, [_D, _U8, ctypes.c_double, _D]),
            "orc_t3_pressure_cs_tangent": (None, [_D, _U8, ctypes.c_double, _D]),
            "orc_t3_volume": (ctypes.c_double, [_D, _D]),
            "orc_t6_pressure_forces": (None, [_D, _U8, ctypes.c_double, _D]),
            "orc_t6_pressure_fd_tangent": (None, [_D, _U8, ctypes.c_double, ctypes.c_double, _D]),
            "orc_t6_pressure_cs_tangent": (None, [_D, _U8, ctypes.c_double, _D]),
            "orc_t6_volume": (ctypes.c_double, [_D, _D]),
            "orc_num_threads": (ctypes.c_int32, []),
        }
        batch = [ctypes.c_int32, ctypes.c_int64, _I32, _D, _D, _D, _D, _U8, _M, _I32, _D, ctypes.c_double]
        sig["orc_residual_batch"] = (ctypes.c_int32, batch + [_D, _I32])
        sig["orc_fd_tangent_batch"] = (ctypes.c_int32, batch + [ctypes.c_double, _D, _I32])
        sig["orc_cs_tangent_batch"] = (ctypes.c_int32, batch + [_D, _I32])
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(_U8)


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I32)


def material(m) -> OrcMaterial:
    """Convert a cardiomechanics_amd.materials.Material (or a dict) to the oracle record."""
    get = (lambda k, d=0.0: m.get(k, d)) if isinstance(m, dict) else (lambda k, d=0.0: getattr(m, k, d))
    r = OrcMaterial()
    r.law = int(get("law"))
    r.ignore_corrupt = int(get("ignore_corrupt", 0))
    r.tension_fn = int(get("tension_fn", 0))
    r.density = float(get("density", 0.0))
    p = list(get("params", []))
    for i in range(12):
        r.p[i] = float(p[i]) if i < len(p) else 0.0
    r.tension_max = float(get("tension_max", 0.0))
    fn = list(get("fn", []))
    for i in range(12):
        r.fn[i] = float(fn[i]) if i < len(fn) else 0.0
    return r


def materials_array(mats):
    arr = (OrcMaterial * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = material(m)
    return arr


# ---- thin numpy front-ends --------------------------------------------------

def analytic(fn_type, fn, time, prev_time=None):
    fn, pf = _d(np.resize(np.asarray(fn, float), 12))
    return lib().orc_analytic(fn_type, pf, time, time if prev_time is None else prev_time)


def pk2(mat, F):
    m = material(mat)
    F, pF = _d(F)
    S = np.zeros(9)
    rc = lib().orc_pk2(ctypes.byref(m), pF, S.ctypes.data_as(_D))
    return rc, S.reshape(3, 3)


def energy(mat, F):
    m = material(mat)
    F, pF = _d(F)
    W = np.zeros(1)
    rc = lib().orc_energy(ctypes.byref(m), pF, W.ctypes.data_as(_D))
    return rc, float(W[0])


def energy_dF_cs(mat, F):
    m = material(mat)
    F, pF = _d(F)
    out = np.zeros(9)
    rc = lib().orc_energy_dF_cs(ctypes.byref(m), pF, out.ctypes.data_as(_D))
    return rc, out.reshape(3, 3)


def pk2_dF_cs(mat, F):
    m = material(mat)
    F, pF = _d(F)
    out = np.zeros(81)
    rc = lib().orc_pk2_dF_cs(ctypes.byref(m), pF, out.ctypes.data_as(_D))
    return rc, out.reshape(9, 9)


def tanh_tension(mat, lam, time, tact=0.0):
    m = material(mat)
    return lib().orc_tanh_tension(ctypes.byref(m), float(lam), float(time), float(tact))


def active_stress_model(mat, F, time, tact=0.0):
    m = material(mat)
    F, pF = _d(F)
    out = np.zeros(9)
    lib().orc_active_stress_model(ctypes.byref(m), pF, float(time), float(tact), out.ctypes.data_as(_D))
    return out.reshape(3, 3)


def active_stress(F, tension):
    F, pF = _d(F)
    out = np.zeros(9)
    lib().orc_active_stress(pF, tension, out.ctypes.data_as(_D))
    return out.reshape(3, 3)


def t10_shape(X30):
    X30, p = _d(X30)
    out = np.zeros(150)
    detJ = np.zeros(1)
    lib().orc_t10_shape(p, out.ctypes.data_as(_D), detJ.ctypes.data_as(_D))
    return out, float(detJ[0])


def t4_shape(X12):
    X12, p = _d(X12)
    out = np.zeros(12)
    detJ = np.zeros(1)
    lib().orc_t4_shape(p, out.ctypes.data_as(_D), detJ.ctypes.data_as(_D))
    return out, float(detJ[0])


def _elem_call(fn, npe, x, dNdX, bases, w, bc, mat, time, tact, extra=(), nout=None):
    m = material(mat)
    x, px = _d(x)
    dNdX, pd = _d(dNdX)
    bases, pb = _d(bases)
    bc, pbc = _u8(bc)
    out = np.zeros(nout)
    rc = fn(px, pd, pb, w, pbc, ctypes.byref(m), time, tact, *extra, out.ctypes.data_as(_D))
    return rc, out


def t10_residual(x, dNdX, bases, detJ, bc, mat, time=0.0, tact=0.0):
    return _elem_call(lib().orc_t10_residual, 10, x, dNdX, bases, detJ, bc, mat, time, tact, nout=30)


def t10_fd_tangent(x, dNdX, bases, detJ, bc, mat, time=0.0, tact=0.0, eps=1e-12):
    rc, K = _elem_call(lib().orc_t10_fd_tangent, 10, x, dNdX, bases, detJ, bc, mat, time, tact, (eps,), 900)
    return rc, K.reshape(30, 30)


def t10_cs_tangent(x, dNdX, bases, detJ, bc, mat, time=0.0, tact=0.0):
    rc, K = _elem_call(lib().orc_t10_cs_tangent, 10, x, dNdX, bases, detJ, bc, mat, time, tact, nout=900)
    return rc, K.reshape(30, 30)


def t4_residual(x, dNdX, basis, volume, bc, mat, time=0.0, tact=0.0):
    return _elem_call(lib().orc_t4_residual, 4, x, dNdX, basis, volume, bc, mat, time, tact, nout=12)


def t4_fd_tangent(x, dNdX, basis, volume, bc, mat, time=0.0, tact=0.0, eps=1e-12):
    rc, K = _elem_call(lib().orc_t4_fd_tangent, 4, x, dNdX, basis, volume, bc, mat, time, tact, (eps,), 144)
    return rc, K.reshape(12, 12)


def t4_cs_tangent(x, dNdX, basis, volume, bc, mat, time=0.0, tact=0.0):
    rc, K = _elem_call(lib().orc_t4_cs_tangent, 4, x, dNdX, basis, volume, bc, mat, time, tact, nout=144)
    return rc, K.reshape(12, 12)


def t10_energy(x, dNdX, bases, detJ, mat):
    m = material(mat)
    x, px = _d(x)
    dNdX, pd = _d(dNdX)
    bases, pb = _d(bases)
    W = np.zeros(1)
    rc = lib().orc_t10_energy(px, pd, pb, detJ, ctypes.byref(m), W.ctypes.data_as(_D))
    return rc, float(W[0])


def t10_energy_grad_cs(x, dNdX, bases, detJ, mat):
    m = material(mat)
    x, px = _d(x)
    dNdX, pd = _d(dNdX)
    bases, pb = _d(bases)
    g = np.zeros(30)
    rc = lib().orc_t10_energy_grad_cs(px, pd, pb, detJ, ctypes.byref(m), g.ctypes.data_as(_D))
    return rc, g


def t10_mass_consistent(detJ, density, bc):
    bc, pbc = _u8(bc)
    M = np.zeros(900)
    lib().orc_t10_mass_consistent(detJ, density, pbc, M.ctypes.data_as(_D))
    return M.reshape(30, 30)


def t10_mass_lumped(volume, density, bc):
    bc, pbc = _u8(bc)
    M = np.zeros(900)
    lib().orc_t10_mass_lumped(volume, density, pbc, M.ctypes.data_as(_D))
    return M.reshape(30, 30)


def t4_mass_consistent(volume, density, bc):
    bc, pbc = _u8(bc)
    M = np.zeros(144)
    lib().orc_t4_mass_consistent(volume, density, pbc, M.ctypes.data_as(_D))
    return M.reshape(12, 12)


def t4_mass_lumped(volume, density, bc):
    bc, pbc = _u8(bc)
    M = np.zeros(144)
    lib().orc_t4_mass_lumped(volume, density, pbc, M.ctypes.data_as(_D))
    return M.reshape(12, 12)


def surface_forces(x, bc, p):
    x, px = _d(x)
    bc, pbc = _u8(bc)
    n = x.size
    f = np.zeros(n)
    fn = lib().orc_t3_pressure_forces if n == 9 else lib().orc_t6_pressure_forces
    fn(px, pbc, p, f.ctypes.data_as(_D))
    return f


def surface_cs_tangent(x, bc, p):
    x, px = _d(x)
    bc, pbc = _u8(bc)
    n = x.size
    K = np.zeros(n * n)
    fn = lib().orc_t3_pressure_cs_tangent if n == 9 else lib().orc_t6_pressure_cs_tangent
    fn(px, pbc, p, K.ctypes.data_as(_D))
    return K.reshape(n, n)


def t6_fd_tangent(x, bc, p, eps=1e-12):
    x, px = _d(x)
    bc, pbc = _u8(bc)
    K = np.zeros(324)
    lib().orc_t6_pressure_fd_tangent(px, pbc, p, eps, K.ctypes.data_as(_D))
    return K.reshape(18, 18)


def surface_volume(x, ref=(0.0, 0.0, 0.0)):
    x, px = _d(x)
    r, pr = _d(np.asarray(ref, float))
    fn = lib().orc_t3_volume if x.size == 9 else lib().orc_t6_volume
    return fn(px, pr)


# ---- mesh-level helpers ------------------------------------------------------

class ElementData:
    """Per-element reference data computed by the oracle's own shape routines."""

    def __init__(self, mesh):
        E, npe = mesh.n_elements, mesh.npe
        self.conn = np.ascontiguousarray(mesh.elements, dtype=np.int32)
        coords, pc = _d(mesh.coords)
        self.dNdX = np.empty((E, 150 if npe == 10 else 12))
        d = np.empty(E)
        # t10_shape / t4_shape of every element (OpenMP)
        lib().orc_shape_batch(npe, E, self.conn.ctypes.data_as(_I32), pc, self.dNdX.ctypes.data_as(_D),
                              d.ctypes.data_as(_D))
        if npe == 10:
            self.w = d
            self.bases = np.ascontiguousarray(mesh.bases.reshape(E, 45))
        else:
            self.w = d / 6.0
            self.bases = np.ascontiguousarray(mesh.bases.reshape(E, 9))
        self.npe = npe
        self.bcmask = np.array(mesh.bc, dtype=np.uint8)          # a copy: callers may edit it
        self.matidx = np.ascontiguousarray(mesh.material, dtype=np.int32)
        self.tact = np.ascontiguousarray(mesh.activation, dtype=np.float64)


def mass_batch(ed: ElementData, density, consistent=True):
    """element mass matrices [E][3npe][3npe] (t10_mass_consistent / t10_mass_lumped /
    t4_mass_consistent / t4_mass_lumped of every element); density per element"""
    E = ed.conn.shape[0]
    nd = 3 * ed.npe
    Me = np.zeros((E, nd, nd))
    rho, prho = _d(density)
    lib().orc_mass_batch(ed.npe, E, ed.conn.ctypes.data_as(_I32), ed.w.ctypes.data_as(_D), prho,
                         ed.bcmask.ctypes.data_as(_U8), int(consistent), Me.ctypes.data_as(_D))
    return Me


def residual_batch(ed: ElementData, coords, mats, time=0.0):
    E = ed.conn.shape[0]
    nd = 3 * ed.npe
    fe = np.zeros((E, nd))
    st = np.zeros(E, np.int32)
    coords, pc = _d(coords)
    marr = materials_array(mats)
    rc = lib().orc_residual_batch(ed.npe, E, ed.conn.ctypes.data_as(_I32), pc, ed.dNdX.ctypes.data_as(_D),
                                  ed.bases.ctypes.data_as(_D), ed.w.ctypes.data_as(_D),
                                  ed.bcmask.ctypes.data_as(_U8), marr, ed.matidx.ctypes.data_as(_I32),
                                  ed.tact.ctypes.data_as(_D), time, fe.ctypes.data_as(_D),
                                  st.ctypes.data_as(_I32))
    return rc, fe, st


def tangent_batch(ed: ElementData, coords, mats, time=0.0, kind="cs", eps=1e-12, sl=None):
    """kind='cs': exact Jacobian, K[r, c] = d f_r / d x_c.  kind='fd': the reference's
    central-difference matrix in its own layout (T10: row = perturbed DOF)."""
    conn, dNdX, bases, w, tact, matidx = ed.conn, ed.dNdX, ed.bases, ed.w, ed.tact, ed.matidx
    if sl is not None:
        conn, dNdX, bases, w, tact, matidx = (np.ascontiguousarray(a[sl]) for a in
                                               (conn, dNdX, bases, w, tact, matidx))
    E = conn.shape[0]
    nd = 3 * ed.npe
    Ke = np.zeros((E, nd, nd))
    st = np.zeros(E, np.int32)
    coords, pc = _d(coords)
    marr = materials_array(mats)
    args = [ed.npe, E, conn.ctypes.data_as(_I32), pc, dNdX.ctypes.data_as(_D), bases.ctypes.data_as(_D),
            w.ctypes.data_as(_D), ed.bcmask.ctypes.data_as(_U8), marr, matidx.ctypes.data_as(_I32),
            tact.ctypes.data_as(_D), time]
    if kind == "fd":
        rc = lib().orc_fd_tangent_batch(*args, eps, Ke.ctypes.data_as(_D), st.ctypes.data_as(_I32))
    else:
        rc = lib().orc_cs_tangent_batch(*args, Ke.ctypes.data_as(_D), st.ctypes.data_as(_I32))
    return rc, Ke, st


# ---- Land17 (CBTensionModelLand17.cpp) -----------------------------------------
def land17_step(params, tmax, state, lam, dt_ms, cai):
    """one evaluation with an explicit Cai: (tension, next state [8], dT/dlambda)"""
    p, pp = _d(params)
    st, pst = _d(state)
    nxt = np.zeros(8)
    d = np.zeros(1)
    t = lib().orc_land17_step(pp, tmax, pst, lam, dt_ms, cai, nxt.ctypes.data_as(_D), d.ctypes.data_as(_D))
    return t, nxt, float(d[0])


def land17_cai(params, time, tact):
    p, pp = _d(params)
    return lib().orc_land17_cai(pp, time, tact)


class Land17State:
    """Per-element Land17 state [E][8] of a mesh (XS XW TRPN TmBlocked ZetaS ZetaW Cd lambda)
    and the time of its last commit; configure() hands it to the element batches for an
    evaluation at `time`, commit() advances it after an accepted step."""

    def __init__(self, mesh, materials):
        self.mats = list(materials)
        E = mesh.n_elements
        self.params = np.zeros((len(self.mats), 40))
        for i, m in enumerate(self.mats):
            if getattr(m, "land17", None) is not None:
                self.params[i] = m.land17
        self.state = np.zeros((E, 8))
        mi = np.asarray(mesh.material, dtype=np.int64)
        self.state[:, :7] = self.params[mi, 27:34]
        self.state[:, 7] = 1.0           # lambda of the undeformed configuration
        self.t_prev = 0.0
        self.active = any(getattr(m, "land17", None) is not None for m in self.mats)

    def configure(self, time):
        if self.active:
            lib().orc_land17_configure(self.params.ctypes.data_as(_D), self.state.ctypes.data_as(_D),
                                       (time - self.t_prev) * 1000.0)

    def commit(self, ed, coords, time):
        if not self.active:
            return
        coords, pc = _d(coords)
        marr = materials_array(self.mats)
        lib().orc_land17_commit_batch(ed.npe, ed.conn.shape[0], ed.conn.ctypes.data_as(_I32), pc,
                                      ed.dNdX.ctypes.data_as(_D), ed.bases.ctypes.data_as(_D), marr,
                                      ed.matidx.ctypes.data_as(_I32), ed.tact.ctypes.data_as(_D),
                                      self.params.ctypes.data_as(_D), time, (time - self.t_prev) * 1000.0,
                                      self.state.ctypes.data_as(_D))
        self.t_prev = time
