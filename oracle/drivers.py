"""CPU FP64 solver drivers -- TEST INFRASTRUCTURE ONLY.

Restates the reference's static and Newmark solver steps on the CPU with the
oracle's element routines and a direct sparse solve (scipy's SuperLU, standing in
for the reference's MUMPS LU, CBSolverNewmarkBeta.cpp:176-192):

  CpuMechanics.residual   CBSolverEquilibrium::CalcNodalForces (CBSolverEquilibrium.cpp:44-108)
                          + CBCirculationCavity::ApplyPressure (CBCirculationCavity.cpp:108-172)
  CpuMechanics.tangent    exact (complex-step) element Jacobians assembled with the
                          identity on constrained DOFs (CBSolver.cpp:413-435)
  static_step             CBSolverEquilibrium::SolverStep / UpdateInitialGuess (:234-363)
  newmark_step            CBSolverNewmarkBeta::SolverStep (CBSolverNewmarkBeta.cpp:829-1019):
                          predictor, Newton on M(d - d~)/(beta dt^2) + f, corrector;
                          with Rayleigh damping C = alpha M + beta_R K(x_n) (:657-702,
                          rebuilt at each new step from the step's start configuration,
                          :857-866) the residual adds C (v~ + gamma/(beta dt)(d - d~))
                          (:384-390) and the tangent gamma/(beta dt) C (:622-626)
Newton iterates to a residual far below the GPU path's tolerance so that both
converge to the same root.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


class CpuMechanics:
    def __init__(self, mesh, materials, pressure_tangent=True, consistent_mass=True, rayleigh=(0.0, 0.0)):
        self.mesh = mesh
        self.materials = list(materials)
        self.ed = O.ElementData(mesh)
        self.N = mesh.n_nodes
        self.n = 3 * self.N
        npe = mesh.npe
        self.dofs = (3 * mesh.elements[:, :, None] + np.arange(3)).reshape(mesh.n_elements, 3 * npe)
        self.bcdof = ((mesh.bc[:, None] >> np.arange(3)) & 1).astype(bool).ravel()
        self.pressure_tangent = pressure_tangent
        self.sdofs = None
        if mesh.surfaces.shape[0]:
            nps = mesh.surfaces.shape[1]
            self.sdofs = (3 * mesh.surfaces[:, :, None] + np.arange(3)).reshape(-1, 3 * nps)
            self.group_ids = np.unique(mesh.surface_id)
            self.sgroup = np.searchsorted(self.group_ids, mesh.surface_id)
        self.pressures = np.zeros(max(1, len(getattr(self, "group_ids", [0]))))
        self.M = self._mass(consistent_mass)
        self.rayleigh = (float(rayleigh[0]), float(rayleigh[1]))
        self.kinetic_energy = 0.0
        self.damping_energy = 0.0
        self.nodes = mesh.coords.ravel().copy()
        self.vel = np.zeros(self.n)
        self.acc = np.zeros(self.n)
        self.d = np.zeros(self.n)
        self.time_last = 0.0
        self.step_last = 0.0
        self.guess = np.zeros(self.n)
        # Land17 cross-bridge state per element, committed after every accepted step
        self.land17 = O.Land17State(mesh, self.materials)

    # -- element / surface loops ------------------------------------------------
    def _mass(self, consistent, chunk=50_000):
        """Assembled mass matrix from the oracle's element mass matrices (t10_mass_consistent
        / t10_mass_lumped / t4_mass_consistent / t4_mass_lumped).  Each is m_ab (x) I3 with the rows of
        constrained DOFs zeroed, so the nodal scalars m_ab of the unconstrained element
        matrices are assembled (in chunks of elements) and the constrained rows cleared
        once: M = D_free (m (x) I3)."""
        rho = np.array([self.materials[k].density for k in self.mesh.material], dtype=np.float64)
        free_ed = O.ElementData.__new__(O.ElementData)
        free_ed.__dict__.update(self.ed.__dict__)
        free_ed.bcmask = np.zeros_like(self.ed.bcmask)
        E, npe = self.mesh.n_elements, self.mesh.npe
        m_node = sp.csr_matrix((self.N, self.N))
        for lo in range(0, E, chunk):
            hi = min(lo + chunk, E)
            sub = O.ElementData.__new__(O.ElementData)
            sub.__dict__.update(free_ed.__dict__)
            sub.conn = np.ascontiguousarray(free_ed.conn[lo:hi])
            sub.w = np.ascontiguousarray(free_ed.w[lo:hi])
            Me = O.mass_batch(sub, rho[lo:hi], consistent).reshape(hi - lo, npe, 3, npe, 3)
            m = Me[:, :, 0, :, 0]
            if lo == 0:   # the element matrices are m_ab (x) I3
                assert np.array_equal(Me, m[:, :, None, :, None] * np.eye(3)[None, None, :, None, :])
            rows = np.repeat(sub.conn, npe, axis=1).ravel()
            cols = np.tile(sub.conn, (1, npe)).ravel()
            m_node = m_node + sp.csr_matrix((m.ravel(), (rows, cols)), shape=(self.N, self.N))
        M = sp.kron(m_node, sp.eye(3), format="csr")
        return (sp.diags((~self.bcdof).astype(float)) @ M).tocsr()

    def surface_forces(self, x):
        f = np.zeros(self.n)
        if self.sdofs is None:
            return f
        xs = x[self.sdofs]
        bcs = self.bcdof[self.sdofs].astype(np.uint8)
        for s in range(xs.shape[0]):
            np.add.at(f, self.sdofs[s], O.surface_forces(xs[s], bcs[s], self.pressures[self.sgroup[s]]))
        return f

    def residual(self, x, time, include_pressure=True):
        self.land17.configure(time)
        rc, fe, st = O.residual_batch(self.ed, x.reshape(-1, 3), self.materials, time)
        if rc != O.SUCCESS:
            raise FloatingPointError(f"corrupt element {int(np.argmax(st != O.SUCCESS))}")
        f = np.bincount(self.dofs.ravel(), weights=fe.ravel(), minlength=self.n)   # in element order
        if include_pressure:
            f += self.surface_forces(x)
        f[self.bcdof] = 0.0
        return f

    def element_tangents(self, x, time):
        self.land17.configure(time)
        rc, Ke, st = O.tangent_batch(self.ed, x.reshape(-1, 3), self.materials, time, kind="cs")
        if rc != O.SUCCESS:
            raise FloatingPointError("corrupt element")
        return Ke

    def tangent(self, x, time, include_pressure=True, cm=0.0, extra=None):
        Ke = self.element_tangents(x, time)
        nd = Ke.shape[1]
        rows = np.repeat(self.dofs, nd, axis=1).ravel()
        cols = np.tile(self.dofs, (1, nd)).ravel()
        vals = Ke.ravel()
        if include_pressure and self.pressure_tangent and self.sdofs is not None:
            xs = x[self.sdofs]
            bcs = self.bcdof[self.sdofs].astype(np.uint8)
            ns = self.sdofs.shape[1]
            Ks = np.array([O.surface_cs_tangent(xs[s], bcs[s], self.pressures[self.sgroup[s]])
                           for s in range(xs.shape[0])])
            rows = np.concatenate([rows, np.repeat(self.sdofs, ns, axis=1).ravel()])
            cols = np.concatenate([cols, np.tile(self.sdofs, (1, ns)).ravel()])
            vals = np.concatenate([vals, Ks.ravel()])
        K = sp.csr_matrix((vals, (rows, cols)), shape=(self.n, self.n))
        if cm != 0.0:
            K = K + cm * self.M
        if extra is not None:
            K = K + extra
        return self.apply_bc(K)

    def apply_bc(self, K):
        K = K.tocsr()
        free = (~self.bcdof).astype(float)
        D = sp.diags(free)
        K = D @ K @ D + sp.diags(self.bcdof.astype(float))
        return K.tocsr()

    # -- Newton -----------------------------------------------------------------
    def damping_matrix(self, time):
        """C = alpha M + beta_R K(x_n): K the element tangent without surface terms at
        the current nodes and the given time, in the layout the reference assembles its
        finite-difference Jacobian in (CBSolverNewmarkBeta.cpp:657-668 builds C from
        it): transposed per T10 element (CBElementSolidT10.cpp:559-565, row = perturbed
        DOF), as is for T4 (CBElementSolidT4.cpp:340-349); constrained rows and columns
        zero."""
        alpha, beta_r = self.rayleigh
        C = alpha * self.M
        if beta_r != 0.0:
            Ke = self.element_tangents(self.nodes, time)
            if self.mesh.npe == 10:
                Ke = np.ascontiguousarray(Ke.transpose(0, 2, 1))
            nd = Ke.shape[1]
            K = sp.csr_matrix((Ke.ravel(), (np.repeat(self.dofs, nd, axis=1).ravel(),
                                            np.tile(self.dofs, (1, nd)).ravel())), shape=(self.n, self.n))
            C = C + beta_r * K
        D = sp.diags((~self.bcdof).astype(float))
        return (D @ C @ D).tocsr()

    def _dynamic_terms(self, d, cm, dpred, damp):
        """inertia cm M (d - d~) and damping C (v~ + cd (d - d~)) on free DOFs"""
        t = cm * (self.M @ (d - dpred))
        if damp is not None:
            C, cd, vpred = damp
            t = t + C @ (vpred + cd * (d - dpred))
        t[self.bcdof] = 0.0
        return t

    def newton(self, time, d, cm=0.0, dpred=None, tol=1e-12, atol=1e-12, maxit=40, damp=None):
        """Full Newton with a direct solve until ||r|| <= max(tol ||r0||, atol) or the
        Newton step falls below 1e-14 ||d|| (the residual's rounding floor).  damp =
        (C, gamma/(beta dt), v~) adds Rayleigh damping."""
        r0 = None
        step = np.inf
        extra = damp[1] * damp[0] if damp is not None else None
        for it in range(maxit + 1):
            x = self.nodes + d
            r = self.residual(x, time)
            if cm != 0.0:
                r = r + self._dynamic_terms(d, cm, dpred, damp)
            rn = np.linalg.norm(r)
            if r0 is None:
                r0 = rn
            if rn <= max(tol * r0, atol) or rn == 0.0 or step <= 1e-14 * np.linalg.norm(d):
                return d, it, rn
            K = self.tangent(x, time, cm=cm, extra=extra)
            delta = spla.spsolve(K.tocsc(), -r)
            lam = 1.0
            while True:
                dn = d + lam * delta
                try:
                    rr = self.residual(self.nodes + dn, time)
                    if cm != 0.0:
                        rr = rr + self._dynamic_terms(dn, cm, dpred, damp)
                    if np.linalg.norm(rr) <= rn or lam < 1.0 / 64:
                        break
                except FloatingPointError:
                    if lam < 2.0 ** -20:     # the step inverts elements however short it is
                        raise RuntimeError("oracle Newton: every step length inverts an element")
                lam *= 0.5
            step = np.linalg.norm(dn - d)
            d = dn
        raise RuntimeError(f"oracle Newton did not converge: |r|={rn:.3e} r0={r0:.3e}")

    def static_step(self, time, dt, tol=1e-12):
        if time > self.time_last:
            if self.step_last > 0:
                self.guess = self.d / self.step_last
                self.d = self.d * (dt / self.step_last)
            else:
                self.d = np.zeros(self.n)
        elif time == self.time_last:
            self.d = np.zeros(self.n)
        else:
            self.d = self.guess * dt
        self.time_last, self.step_last = time, dt
        self.d, it, rn = self.newton(time, self.d, tol=tol)
        self.nodes = self.nodes + self.d
        self.land17.commit(self.ed, self.nodes.reshape(-1, 3), time)
        return it, rn

    def newmark_solve(self, time, dt, beta=0.25, gamma=0.5, tol=1e-12, from_trial=False, atol=1e-12):
        """predictor + Newton (CBSolverNewmarkBeta.cpp:845-923); the increment stays
        pending until commit_step"""
        dpred = dt * self.vel + dt * dt * 0.5 * (1 - 2 * beta) * self.acc
        vpred = 1.0 * self.vel + (1 - gamma) * dt * self.acc
        cm = 1.0 / (beta * dt * dt)
        pend = getattr(self, "_pending", None)
        d0 = pend["d"].copy() if (from_trial and pend is not None and pend["dt"] == dt) else dpred.copy()
        self._pending = None
        damp = None
        if self.rayleigh != (0.0, 0.0):
            damp = (self.damping_matrix(time), gamma / (beta * dt), vpred)
        d, it, rn = self.newton(time, d0, cm=cm, dpred=dpred, tol=tol, atol=atol, damp=damp)
        self._pending = dict(d=d, dpred=dpred, vpred=vpred, dt=dt, beta=beta, gamma=gamma, time=time,
                             C=damp[0] if damp is not None else None)
        return it, rn

    def trial_volumes(self):
        return self.cavity_volumes(self.nodes + self._pending["d"])

    def commit_step(self):
        """corrector (CBSolverNewmarkBeta.cpp:978-1015)"""
        p = self._pending
        dt, beta, gamma = p["dt"], p["beta"], p["gamma"]
        self.acc = 1.0 / (beta * dt * dt) * p["d"] + -1.0 / (beta * dt * dt) * p["dpred"]
        self.vel = p["vpred"] + gamma * dt * self.acc
        # kinetic energy and damping dissipation (CBSolverNewmarkBeta.cpp:990-1003);
        # the mass matrix's constrained rows are zero
        mv = self.M @ self.vel
        mv[self.bcdof] = 0.0
        self.kinetic_energy = 0.5 * float(mv @ self.vel)
        if p["C"] is not None:
            self.damping_energy += float((p["C"] @ self.vel) @ p["d"])
        self.d = p["d"]
        self.nodes = self.nodes + p["d"]
        self.land17.commit(self.ed, self.nodes.reshape(-1, 3), p["time"])
        self._pending = None

    def newmark_step(self, time, dt, beta=0.25, gamma=0.5, tol=1e-12, atol=1e-12):
        it, rn = self.newmark_solve(time, dt, beta, gamma, tol, atol=atol)
        self.commit_step()
        return it, rn

    def cavity_volumes(self, x=None):
        x = self.nodes if x is None else x
        vols = np.zeros(len(self.group_ids))
        xs = x[self.sdofs]
        for s in range(xs.shape[0]):
            vols[self.sgroup[s]] += O.surface_volume(xs[s])
        return vols


class CpuReferenceSolver:
    """The reference's solver steps with its central-difference element tangent on the
    host cores (oracle/cpu_solver.cpp): orc_cpu_newmark_step restates
    CBSolverNewmarkBeta::SolverStep (CBSolverNewmarkBeta.cpp:829-1019), orc_cpu_static_step
    CBSolverEquilibrium::SolverStep (CBSolverEquilibrium.cpp:234-363); the tangent is the
    FD Jacobian with h = CBSolver::CalcFiniteDifferencesEpsilon (CBSolver.cpp:1610-1625);
    the linear solve is BiCGStab with block Jacobi in place of the reference's MUMPS LU.
    The CPU baseline of the benchmark's beat and the FD-tangent solver path."""

    STATS = ("newton", "krylov", "reason", "residual", "initial_residual", "tangent_s", "linear_s", "wall_s")

    def __init__(self, mesh, materials, eps=1e-12):
        import ctypes
        self.mesh = mesh
        self.ed = O.ElementData(mesh)
        self.eps = eps
        self.n = 3 * mesh.n_nodes
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a
        marr = O.materials_array(materials)
        self._keep.append(marr)
        coords = arr(mesh.coords, np.float64)
        if mesh.surfaces.shape[0]:
            sconn = arr(mesh.surfaces, np.int32)
            gids = np.unique(mesh.surface_id)
            sgroup = arr(np.searchsorted(gids, mesh.surface_id), np.int32)
            nps, S, ng = mesh.surfaces.shape[1], mesh.surfaces.shape[0], len(gids)
        else:
            sconn = arr(np.zeros(1), np.int32)
            sgroup = arr(np.zeros(1), np.int32)
            nps, S, ng = 3, 0, 1
        self.n_groups = ng
        D, I32, U8 = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint8)
        ed = self.ed
        self.h = O.lib().orc_cpu_solver_create(
            ed.npe, mesh.n_elements, ed.conn.ctypes.data_as(I32), mesh.n_nodes, coords.ctypes.data_as(D),
            ed.dNdX.ctypes.data_as(D), ed.bases.ctypes.data_as(D), ed.w.ctypes.data_as(D),
            ed.bcmask.ctypes.data_as(U8), marr, len(materials), ed.matidx.ctypes.data_as(I32),
            ed.tact.ctypes.data_as(D), nps, S, sconn.ctypes.data_as(I32), sgroup.ctypes.data_as(I32), ng)

    def close(self):
        if getattr(self, "h", None):
            O.lib().orc_cpu_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_pressures(self, p):
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, float), (self.n_groups,)))
        O.lib().orc_cpu_set_pressures(self.h, p.ctypes.data_as(O._D))

    def state(self):
        out = [np.zeros(self.n) for _ in range(3)]
        O.lib().orc_cpu_get_state(self.h, *[a.ctypes.data_as(O._D) for a in out])
        return dict(zip(("nodes", "velocity", "acceleration"), out))

    def set_state(self, nodes=None, velocity=None, acceleration=None):
        args = []
        for a in (nodes, velocity, acceleration):
            if a is None:
                args.append(None)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                self._keep.append(a)
                args.append(a.ctypes.data_as(O._D))
        O.lib().orc_cpu_set_state(self.h, *args)

    def newmark_step(self, time, dt, beta=0.25, gamma=0.5, atol=1e-8, rtol=1e-8, stol=1e-8, maxit=50):
        st = np.zeros(8)
        rc = O.lib().orc_cpu_newmark_step(self.h, time, dt, beta, gamma, self.eps, atol, rtol, stol, maxit,
                                          st.ctypes.data_as(O._D))
        return rc, dict(zip(self.STATS, st.tolist()))

    def static_step(self, time, dt, atol=1e-8, rtol=1e-8, stol=1e-8, maxit=50):
        st = np.zeros(8)
        rc = O.lib().orc_cpu_static_step(self.h, time, dt, self.eps, atol, rtol, stol, maxit,
                                         st.ctypes.data_as(O._D))
        return rc, dict(zip(self.STATS, st.tolist()))
