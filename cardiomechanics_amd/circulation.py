"""0D circulation coupled to the cavity volumes (host side).

  CircModel          CBCircModel<nC, nS, nR> (mechanics/src/CBCavity/CBCircModel.hpp):
                     state / estimated / previous state variables, RK4 integration in
                     ceil(dt / MaxIntegrationTimeStep) sub-steps at a constant cavity
                     pressure, accept / revert, temporal-difference steady-state check,
                     results file
  CircWindkessel3    CBCircWindkessel3 (mechanics/src/CBCavity/CBCircWindkessel3.cpp):
                     ventricle + arterial compliance behind an ideal aortic valve,
                     peripheral resistance, venous filling through an ideal mitral valve
  Circulation        CBCirculation (mechanics/src/CBSolverPlugins/CBCirculation.cpp):
                     per Newmark step, finds the cavity pressures at which the finite-
                     element cavity volumes equal the circulation's volumes.  Each trial
                     pressure is one mechanics solve; the plugin answers REPEAT until
                     |V_FE - V_circ| <= Tolerance.  Pressure estimates: a small increase
                     or an Adams-Bashforth extrapolation of past pressures, then secant,
                     perturbation + Newton on assembled compliance matrices and Broyden
                     (quasi-Newton) updates of the inverse system compliance; optional
                     coupling damping and the two preloading phases.

Pressures are kept in the plugin's PressureUnit (mmHg, hPa or Pa) and volumes in its
VolumeUnit (ml or m3), as in the reference; the mechanics receives Pascal.  The cavity
of a circulation is the pressure group of its surface index; its volume is the signed
volume of that group's surface about the origin (CBCirculationCavity.cpp:77-94), so a
circulation cavity needs surface normals that point out of the blood volume.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from .settings import as_bool

# CBStatus values the plugin protocol uses (CBStatus.h:20-37)
DACCORD = 6
PREPARING_SIMULATION = 3


class CircModel:
    """Base of the lumped circulation models: n_cav cavities, state names, result names."""
    state_names: tuple = ()
    result_names: tuple = ()
    n_cav = 1

    def __init__(self, params: dict, prefix: str):
        self.prefix = prefix
        self.name = prefix.rsplit(".", 1)[-1]
        self.filename = str(params.get(prefix + ".ExportFile", ""))
        self.max_dt = float(params.get(prefix + ".MaxIntegrationTimeStep", 1e-4))
        nS = len(self.state_names)
        self.state = np.zeros(nS)
        self.prev_state = np.zeros(nS)
        self.est_state = np.zeros(nS)
        self.results = {k: 0.0 for k in self.result_names}
        self.cavity_pressures = np.zeros(self.n_cav)
        self.prev_cavity_pressures = np.zeros(self.n_cav)
        self.est_cavity_pressures = np.zeros(self.n_cav)
        self.cavity_surfaces = [0] * self.n_cav
        self.preloading = np.zeros(2 * self.n_cav)
        self._steady_init(params, prefix)

    # -- model equations (subclasses) ---------------------------------------------
    def algebraics(self, p, u):
        raise NotImplementedError

    def rates(self, p, u):
        raise NotImplementedError

    def set_initial_cavity_volumes(self, v):
        raise NotImplementedError

    def est_cavity_volumes(self, p, dt):
        raise NotImplementedError

    # -- CBCircModel.hpp ------------------------------------------------------------
    def integrate(self, p, dt):
        """RK4 over dt in ceil(dt / MaxIntegrationTimeStep) equal sub-steps with the
        cavity pressures held at p (CBCircModel.hpp:187-236); the results are the
        algebraic variables at the end state."""
        n = math.ceil(dt / self.max_dt)
        h = dt / n
        self.est_cavity_pressures[:] = p
        u = self.state.copy()
        self.est_state = u
        for _ in range(n):
            k1 = self.rates(p, u)
            k2 = self.rates(p, u + h * k1 / 2.0)
            k3 = self.rates(p, u + h * k2 / 2.0)
            k4 = self.rates(p, u + h * k3)
            u += h * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0
        self.algebraics(p, u)

    def accept_est_state(self):
        self.prev_state = self.state.copy()
        self.state = self.est_state.copy()
        self.prev_cavity_pressures = self.cavity_pressures.copy()
        self.cavity_pressures = self.est_cavity_pressures.copy()

    def revert_prev_state(self):
        self.state = self.prev_state.copy()
        self.cavity_pressures = self.prev_cavity_pressures.copy()

    def _steady_init(self, params, prefix):
        """temporal-difference steady-state check (CBCircModel.hpp:26-64)"""
        key = prefix + ".SteadyStateCheck"
        self.steady_active = as_bool(params.get(key + ".Active", False))
        self.prev_check_time = 0.0
        self.prev_check_time_active = 0.0
        self.prev_check_value = 0.0
        if self.steady_active:
            self.steady_start = float(params[key + ".StartTime"])
            self.steady_period = float(params[key + ".Period"])
            self.steady_thresh = float(params[key + ".Threshold"])
            mode = str(params.get(key + ".Mode", "TemporalDifference"))
            if mode != "TemporalDifference":
                raise ValueError(f"{self.name}: only TemporalDifference steady-state checks are available")
            self.steady_param = str(params[key + ".Parameter"])
            if self.steady_param not in self.state_names and self.steady_param not in self.result_names:
                raise ValueError(f"{self.steady_param} is not a state or result of {self.name}")

    def _steady_value(self):
        if self.steady_param in self.state_names:
            return float(self.state[self.state_names.index(self.steady_param)])
        return float(self.results[self.steady_param])

    def steady_state_check(self, time) -> bool:
        """(CBCircModel.hpp:66-103): compare the parameter one period apart"""
        steady = False
        if self.steady_active:
            half = 0.5 * (time - self.prev_check_time)
            if abs(time - self.steady_start) < half:
                self.prev_check_value = self._steady_value()
                self.prev_check_time_active = self.steady_start
            elif time > self.steady_start and abs(time - self.prev_check_time_active - self.steady_period) < half:
                val = abs(self._steady_value() - self.prev_check_value)
                self.prev_check_value = self._steady_value()
                steady = val < self.steady_thresh
                self.prev_check_time_active = time
        self.prev_check_time = time
        return steady

    def write_header(self):
        if self.filename:
            with open(self.filename, "w") as f:
                f.write("%18s" % "time" + "".join("%18s" % k for k in self.state_names + self.result_names))

    def write_results(self, time):
        if self.filename:
            with open(self.filename, "a") as f:
                f.write("\n%18.5f" % time + "".join("%18.9f" % v for v in self.est_state)
                        + "".join("%18.9f" % self.results[k] for k in self.result_names))


class CircWindkessel3(CircModel):
    """Three-element Windkessel behind the ventricle (CBCircWindkessel3.cpp:16-121)."""
    state_names = ("VentrVolume", "ArtVolume")
    result_names = ("VentrPressure", "ArtPressure", "ArtFlow", "PerFlow", "VenFlow")
    n_cav = 1

    def __init__(self, params: dict, prefix: str):
        super().__init__(params, prefix)
        g = lambda k: float(params[prefix + k])  # noqa: E731
        self.art_resist = g(".CircParameters.ArtResist")
        self.art_compli = g(".CircParameters.ArtCompli")
        self.per_resist = g(".CircParameters.PerResist")
        self.ven_resist = g(".CircParameters.VenResist")
        self.ven_pressure = g(".CircParameters.VenPressure")
        self.initial_art_volume = g(".InitialConditions.ArtVolume")
        self.reset_art_volume = as_bool(params.get(prefix + ".InitialConditions.ResetArtVolume", False))
        self.state[1] = self.initial_art_volume
        self.cavity_surfaces = [int(params[prefix + ".Cavities.VentrSurface"])]
        self.preloading = np.array([float(params.get(prefix + ".Cavities.VentrPreloading", 0.0)),
                                    float(params.get(prefix + ".InitialConditions.VentrPressure", 0.0))])
        self.write_header()

    def set_initial_cavity_volumes(self, v):
        self.state[0] = v[0]

    def est_cavity_volumes(self, p, dt):
        self.integrate(p, dt)
        return self.est_state[:1].copy()

    def algebraics(self, p, u):
        pv = p[0]
        pc = u[1] / self.art_compli                      # pressure of the arterial compliance
        r = self.results
        if pv >= pc:                                     # aortic valve open
            r["ArtFlow"] = (pv - pc) / self.art_resist
            r["ArtPressure"] = pv
        else:
            r["ArtFlow"] = 0.0
            r["ArtPressure"] = pc
            if self.reset_art_volume:
                # the arterial volume follows the cavity (an atrial afterload whose valve never opens by drift)
                self.est_state[1] = max(self.initial_art_volume, self.art_compli * pv)
                r["ArtPressure"] = self.est_state[1] / self.art_compli
        r["PerFlow"] = pc / self.per_resist
        r["VenFlow"] = (self.ven_pressure - pv) / self.ven_resist if self.ven_pressure >= pv else 0.0
        r["VentrPressure"] = pv

    def rates(self, p, u):
        self.algebraics(p, u)
        r = self.results
        return np.array([r["VenFlow"] - r["ArtFlow"], r["ArtFlow"] - r["PerFlow"]])


CIRC_TYPES = {"CircWindkessel3": CircWindkessel3}

PRESSURE_UNITS = {"mmHg": 1.33322e2, "hPa": 1e2, "Pa": 1.0}
VOLUME_UNITS = {"ml": 1e6, "m3": 1.0}


class Circulation:
    """CBCirculation: the plugin that couples the circulation models to the cavities.

    groups: {surface index: pressure group of the context} for the cavities used.
    The protocol (called by `solver.run`): apply(time, dt) before each trial solve,
    pressures() for the context, analyze(volumes_m3) after a converged trial solve
    (sets `status`), step_back() when the solver rejects a step, write(time) and
    exit_check(time) after accepted steps."""
    N = 5   # past pressures kept for the extrapolation

    def __init__(self, params: dict, groups: dict, root="Plugins.Circulation"):
        g = lambda k, d=None: params.get(f"{root}.{k}", d)  # noqa: E731
        self.start_time = float(g("StartTime", -math.inf))
        self.status = DACCORD
        self.tolerance = float(g("Tolerance", 1e-4))
        self.max_iterations = int(g("MaxIterations", 100))
        self.n_non_newton = int(g("MaxNonNewtonIterations", 10))
        self.n_secant = int(g("SecantIterations", 3))
        self.perturbation = as_bool(g("Perturbation", True))
        self.preload_t1 = float(g("PreloadingTime1", 0.0))
        self.preload_t2 = float(g("PreloadingTime2", 0.0))
        if self.preload_t1 != 0.0:
            raise NotImplementedError("PreloadingTime1 > 0 relaxes the reference configuration "
                                      "(RelaxElementsAndBases), which this build does not provide")
        self.damping = min(float(g("CouplingDamping.InitialFactor", 0.0)), 0.9)
        self.damping_active = self.damping >= 1e-5
        self.damping_decline = float(g("CouplingDamping.DeclineFactor", 0.0))
        unit_p = str(g("PressureUnit"))
        unit_v = str(g("VolumeUnit"))
        if unit_p not in PRESSURE_UNITS:
            raise ValueError(f"pressure unit {unit_p} unknown")
        if unit_v not in VOLUME_UNITS:
            raise ValueError(f"volume unit {unit_v} unknown")
        self.uP, self.uV = PRESSURE_UNITS[unit_p], VOLUME_UNITS[unit_v]
        self.maxP = 1.33322e3 / self.uP      # 10 mmHg: largest change of an estimate
        self.deltaP = 1.33322 / self.uP      # 0.01 mmHg: increase while no history exists
        self.dP = 1.33322e-6 / self.uP       # 1e-8 mmHg: perturbation
        self.its_file = g("IterationsFile", None)
        self.models = []
        prefix = f"{root}.Circs."
        names = sorted({k[len(prefix):].split(".", 1)[0] for k in params if k.startswith(prefix)})
        for name in names:
            key = prefix + name
            if not as_bool(params.get(key + ".Active", False)):
                continue
            typ = str(params.get(key + ".Type", ""))
            if typ not in CIRC_TYPES:
                raise ValueError(f"circulation model type {typ} is not available")
            self.models.append(CIRC_TYPES[typ](params, key))
        self.surfaces = []
        for m in self.models:
            for sidx in m.cavity_surfaces:
                if sidx in self.surfaces:
                    raise ValueError(f"cavity {sidx} is used more than once")
                if sidx not in groups:
                    raise ValueError(f"no cavity with surface index {sidx}")
                self.surfaces.append(sidx)
        self.groups = [groups[s] for s in self.surfaces]
        nC = self.nC = len(self.surfaces)
        self.est = np.zeros(nC)
        self.prev_est = np.zeros(nC)
        self.pressure = [np.zeros(nC) for _ in range(self.N)]
        self.time_step = np.zeros(self.N)
        self.fine = np.zeros(nC)
        self.prev_fine = np.zeros(nC)
        self.circ = np.zeros(nC)
        self.prev_circ = np.zeros(nC)
        self.preload1 = np.concatenate([m.preloading[:m.n_cav] for m in self.models]) if nC else np.zeros(0)
        self.preload2 = np.concatenate([m.preloading[m.n_cav:] for m in self.models]) if nC else np.zeros(0)
        self.residual = np.zeros(nC)
        self.prev_residual = np.zeros(nC)
        self.fine_compli = np.zeros((nC, nC))
        self.circ_compli = np.zeros((nC, nC))
        self.sys_elast = np.eye(nC)
        self.n = 1
        self.i = 0
        self.j = 0
        self.i_sum = 0
        self.has_started = False
        self.step_back_flag = False
        self.preloading_finished = False
        self.last_preload1 = False
        self.last_preload2 = False
        self.damping_iteration = False
        self.dt = 0.0
        self.prev_dt = 0.0
        self.wants_zero_velocity = False
        self.log = []     # (time, iterations, estimate kind, pressures) of each Apply

    # -- set-up ---------------------------------------------------------------------
    def init_volumes(self, volumes_m3, dt):
        """InitCircModels (CBCirculation.cpp:140-230) after the cavities exist: initial
        FE volumes become the circulation's cavity volumes."""
        self.dt = dt
        self.fine = self.uV * np.asarray([volumes_m3[g] for g in self.groups], float)
        self.circ = self.fine.copy()
        self._set_initial_volumes(self.fine)
        if self.its_file:
            open(self.its_file, "w").close()

    def _set_initial_volumes(self, v):
        pos = 0
        for m in self.models:
            m.set_initial_cavity_volumes(v[pos:pos + m.n_cav])
            pos += m.n_cav

    def _circ_volumes(self, p):
        out = np.zeros(self.nC)
        pos = 0
        for m in self.models:
            out[pos:pos + m.n_cav] = m.est_cavity_volumes(p[pos:pos + m.n_cav], self.dt)
            pos += m.n_cav
        return out

    def pressures(self):
        """{pressure group: Pa} to apply (ApplyToNodalForces, CBCirculation.cpp:282-290)"""
        if not self.has_started:
            return {}
        return {g: self.uP * self.est[k] for k, g in enumerate(self.groups)}

    # -- pressure estimates -----------------------------------------------------------
    def _extrapolate(self):
        """Adams-Bashforth-type extrapolation over the last five accepted pressures
        with their own step sizes (CBCirculation.cpp:518-525)"""
        n, N, P, T = self.n, self.N, self.pressure, self.time_step
        d = [(P[(n - k) % N] - P[(n - k - 1) % N]) / T[(n - k) % N] for k in (1, 2, 3, 4)]
        return P[(n - 1) % N] + self.dt / 24.0 * (55.0 * d[0] - 59.0 * d[1] + 37.0 * d[2] - 9.0 * d[3])

    def _extrapolate_pressure(self):
        self.prev_est = self.est.copy()
        self.est = self._extrapolate()
        if self.damping_active:
            self.est = (self.est - self.damping * self.pressure[(self.n - 1) % self.N]) / (1.0 - self.damping)
        return "extrapolate"

    def _increase_pressure(self):
        self.prev_est = self.est.copy()
        self.est = self.est + self.deltaP
        return "increase"

    def _perturb(self, k):
        self.prev_est[k] = self.est[k]
        self.est[k] += self.dP
        return f"perturb {k}"

    def _assemble_compliance(self, k):
        dp = self.est[k] - self.prev_est[k]
        self.fine_compli[:, k] = (self.fine - self.prev_fine) / dp
        self.circ_compli[:, k] = (self.circ - self.prev_circ) / dp

    def _secant(self):
        dp = self.est - self.prev_est
        dpdr = 1.0 / ((self.fine - self.prev_fine) / dp - (self.circ - self.prev_circ) / dp)
        self.prev_est = self.est.copy()
        self.est = self.est - dpdr * (self.fine - self.circ)
        return "secant"

    def _quasi_newton(self):
        """Broyden update of the inverse system compliance (CBCirculation.cpp:582-600)"""
        self.prev_residual = self.residual.copy()
        self.residual = self.fine - self.circ
        if self.i > 1 or self.n < 10:   # the first iteration of a step reuses the last inverse
            dF = self.residual - self.prev_residual
            dX = self.est - self.prev_est
            v = self.sys_elast @ dF
            self.sys_elast = self.sys_elast + np.outer((dX - v) / (dX @ v), dX) @ self.sys_elast
        self.prev_est = self.est.copy()
        self.est = self.est - self.sys_elast @ self.residual
        return "quasi-newton"

    def _newton(self):
        self.residual = self.fine - self.circ
        self.sys_elast = np.linalg.inv(self.fine_compli - self.circ_compli)
        self.prev_est = self.est.copy()
        self.est = self.est - self.sys_elast @ self.residual
        self.j = 0
        return "newton"

    # -- plugin protocol --------------------------------------------------------------------
    def apply(self, time, dt):
        """CBCirculation::Apply (CBCirculation.cpp:300-420)"""
        if not self.has_started:
            if time > self.start_time:
                self.has_started = True
            else:
                return
        self.prev_dt, self.dt = self.dt, dt
        kind = ""
        if self.preloading_finished:
            nS, nC = self.n_secant, self.nC
            if self.damping_iteration:
                self.est = self.damping * self.pressure[(self.n - 1) % self.N] + (1.0 - self.damping) * self.est
                kind = "damping"
                self.damping *= self.damping_decline
                if self.damping < 1e-5:
                    self.damping_active = False
            elif self.status == _lib.SUCCESS:
                if not self.step_back_flag:
                    for m in self.models:
                        m.accept_est_state()
                    self.pressure[self.n % self.N] = self.est.copy()
                    self.time_step[self.n % self.N] = self.prev_dt
                    self.n += 1
                    self.i = 0
                    self.j = 0
                kind = self._extrapolate_pressure() if self.n > self.N else self._increase_pressure()
            else:
                i, j = self.i, self.j
                if i == 0:
                    kind = self._extrapolate_pressure() if self.n > self.N else self._increase_pressure()
                elif self.perturbation and i <= nS:
                    kind = self._secant()
                elif self.perturbation and i <= nS + nC:
                    kind = self._perturb(i - nS - 1)
                elif self.n_non_newton < j <= self.n_non_newton + nC:
                    kind = self._perturb(j - self.n_non_newton - 1)
                elif (not self.perturbation or i > nS + nC + 1) and j <= self.n_non_newton:
                    kind = self._quasi_newton()
                if self.perturbation or j > self.n_non_newton + 1:
                    if nS + 1 < i <= nS + nC + 1:
                        self._assemble_compliance(i - nS - 2)
                    elif j > self.n_non_newton + 1:
                        self._assemble_compliance(j - self.n_non_newton - 2)
                    if i == nS + nC + 1 or j == self.n_non_newton + nC + 1:
                        kind = self._newton()
            diff = self.est - self.prev_est
            over, under = diff > self.maxP, diff < -self.maxP
            self.est = np.where(over, self.prev_est + self.maxP, np.where(under, self.prev_est - self.maxP, self.est))
            self.step_back_flag = False
            self.i += 1
            self.j += 1
            self.i_sum += 1
        else:
            if self.last_preload1:
                self.wants_zero_velocity = True
                self.last_preload1 = False
            if time < self.preload_t1:
                self.est = -time / self.preload_t1 * self.preload1
                kind = "preloading 1"
            elif abs(time - self.preload_t1) < 0.5 * dt:
                self.est = -1.0 * self.preload1
                self.last_preload1 = True
                kind = "preloading 1, last step"
            elif time <= self.preload_t1 + self.preload_t2:
                self.est = (time - self.preload_t1) / self.preload_t2 * self.preload2
                kind = "preloading 2"
            else:
                self.est = self.preload2.copy()
                self.pressure[0] = self.est.copy()
                self.last_preload2 = True
                kind = "preloading 2, last step"
        self.status = _lib.SUCCESS
        self.log.append((time, self.i, kind, self.est.copy()))

    def analyze(self, volumes_m3):
        """CBCirculation::AnalyzeResults (CBCirculation.cpp:422-500)"""
        if not self.has_started:
            return
        fe = self.uV * np.asarray([volumes_m3[g] for g in self.groups], float)
        if not self.preloading_finished:
            if self.last_preload2:
                self.fine = fe
                self.circ = fe.copy()
                self._set_initial_volumes(fe)
                self.preloading_finished = True
                self.status = DACCORD
            else:
                self.status = _lib.SUCCESS
            return
        self.prev_fine, self.fine = self.fine, fe
        self.prev_circ, self.circ = self.circ, self._circ_volumes(self.est)
        self.status = _lib.SUCCESS
        if self.damping_iteration:
            self.damping_iteration = False
            return
        for k in range(self.nC):
            if abs(self.fine[k] - self.circ[k]) > self.tolerance:
                if self.i > self.max_iterations:
                    self.status = _lib.FAILED
                else:
                    self.status = _lib.REPEAT
                    break
        if self.damping_active and self.status == _lib.SUCCESS:
            self.status = _lib.REPEAT
            self.damping_iteration = True

    def step_back(self):
        """CBCirculation::StepBack: restart the step from the last accepted pressure"""
        if not self.has_started:
            return
        self.est = self.pressure[(self.n - 1) % self.N].copy()
        self.step_back_flag = True
        self.i = 0
        self.j = 0

    def write(self, time):
        if not self.has_started:
            return
        if time > self.preload_t1 + self.preload_t2 + self.dt:
            for m in self.models:
                m.write_results(time)
            if self.its_file:
                with open(self.its_file, "a") as f:
                    f.write("%14.3f%14d\n" % (time, self.i))

    def exit_check(self, time) -> bool:
        if not self.has_started:
            return False
        steady = [m.steady_state_check(time) for m in self.models]
        return bool(steady) and all(steady)

    def coupling_residual(self):
        """|V_FE - V_circ| per cavity at the last analysed trial (volume unit)"""
        return np.abs(self.fine - self.circ)
