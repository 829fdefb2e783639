"""Material, tension and time-function parameters (host side).

Mirrors how the reference reads them (key names kept so settings round-trip):
  Guccione   CBConstitutiveModelGuccione::Init   (CBConstitutiveModelGuccione.cpp:20-71)
  Holzapfel  CBConstitutiveModelHolzapfel::Init  (CBConstitutiveModelHolzapfel.cpp:18-85)
  density    CBMaterial::Init                    (CBMaterial.cpp:19-74)
  tension    CBTensionFromFunction               (CBTensionFromFunction.cpp:36-56)
             CBTensionModelTanH                  (CBTensionModelTanH.cpp:27-125)
  functions  CBAnalyticFunction subclasses       (CBData/CBAnalyticFunction.h:39-176)
Per-material keys fall back to Materials.Mat_Default.<key> when absent.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

from .settings import as_bool

LAW_GUCCIONE = 0
LAW_HOLZAPFEL = 1
FN_NONE, FN_LINEAR, FN_SINUS, FN_SINUSDRIVER, FN_DOUBLEHILL = 0, 1, 2, 3, 4
FN_TANH = 5   # the TanH tension model (a stretch-dependent law, not a time function)
FN_LAND17 = 6  # the Land17 cross-bridge model (ODE state per element)
_FN_NAMES = {"None": FN_NONE, "Linear": FN_LINEAR, "Sinus": FN_SINUS, "SinusDriver": FN_SINUSDRIVER,
             "DoubleHill": FN_DOUBLEHILL}


@dataclass
class Material:
    law: int = LAW_GUCCIONE
    # Guccione: C, bf, bt, bfs, K            Holzapfel: a, b, af, bf, as, bs, afs, bfs, kappa, k
    params: list = field(default_factory=list)
    density: float = 0.0
    ignore_corrupt: int = 0
    tension_fn: int = FN_NONE
    tension_max: float = 0.0
    # startTime, stopTime, keepMaxValue, omega, phi, DoubleHill period, contrTimeOffset,
    # relaxTimeOffset, contrRateConst, relaxRateConst, onsetTime, driverMax
    fn: list = field(default_factory=lambda: [0.0] * 12)
    # Land17 parameter block (LAND17_KEYS order), set by with_land17
    land17: list | None = None


def guccione(C, bf, bt, bfs, K, density=1082.0, a_scale=1.0, b_scale=1.0, **kw) -> Material:
    """Guccione law with the aScale/bScale hooks applied as in Init (:67-70)."""
    return Material(law=LAW_GUCCIONE, params=[C * a_scale, bf * b_scale, bt * b_scale, bfs * b_scale, K],
                    density=density, **kw)


def holzapfel(a, b, af, bf, as_, bs, afs, bfs, kappa, k=100.0, density=1082.0, **kw) -> Material:
    return Material(law=LAW_HOLZAPFEL, params=[a, b, af, bf, as_, bs, afs, bfs, kappa, k], density=density, **kw)


# TanH keys and defaults in the order of the device parameter block fn[0..8]
# (CBTensionModelTanH.cpp:99-118)
TANH_KEYS = (("cycleLength", 1.0), ("MinFiberStretch", 0.7), ("EMD", 0.015), ("PeakTension", 120.0),
             ("Duration", 0.3), ("ContractionTime", 0.1), ("DegreeOfLengthDependence", 5.0),
             ("LengthDependenceUpstrokeTime", 0.5), ("RelaxationTime", 0.1))


def tanh_params(kff=1.0, start_time=0.0, **kw):
    """fn block of the TanH tension model: the nine TANH_KEYS (keyword = key name),
    kff and startTime.  Tmax must be 1000 (the model's output is in kPa)."""
    vals = dict(TANH_KEYS)
    for k, v in kw.items():
        if k not in vals:
            raise KeyError(f"unknown TanH parameter {k}")
        vals[k] = float(v)
    return FN_TANH, [vals[k] for k, _ in TANH_KEYS] + [float(kff), float(start_time), 0.0]


def with_tanh(m: Material, **kw) -> Material:
    """the material with the TanH tension model (Tmax = 1000, TanH parameters as keywords)"""
    m.tension_max = 1000.0
    m.tension_fn, m.fn = tanh_params(**kw)
    return m


# Land17 keys and defaults in the order of the device parameter block (InitParamsFromXml,
# CBTensionModelLand17.cpp:428-500): model constants, then startTime, cycleLength, the
# calcium transient (0 Coppini, 1 Brixius), rateDependancy (0 OFF, 1 ON), kff, then the
# initial state XS XW TRPN TmBlocked ZetaS ZetaW Cd
LAND17_KEYS = (("TRPNn", 2.0), ("TRPNk", 0.1), ("Ca50", 0.805), ("ku", 1.0), ("nTM", 5.0), ("TRPN50", 0.35),
               ("kuw", 0.182), ("kws", 0.012), ("rs", 0.25), ("rw", 0.5), ("gammas", 0.0085), ("gammaw", 0.615),
               ("phi", 2.23), ("Aeff", 25.0), ("beta0", 2.3), ("beta1", -2.4), ("Tref", 120.0), ("a", 2.1),
               ("k", 7.0), ("eta_l", 200.0), ("eta_s", 20.0), ("xi", 1.0), ("startTime", 0.0),
               ("cycleLength", 1.0), ("CalciumTransientType", 0.0), ("rateDependancy", 0.0), ("kff", 1.0),
               ("XS", 0.0), ("XW", 0.0), ("TRPN", 0.0), ("TmBlocked", 1.0), ("ZetaS", 0.0), ("ZetaW", 0.0),
               ("Cd", 0.0))
LAND17_NPARAMS = 40
_CALCIUM = {"Coppini": 0.0, "Brixius": 1.0}


def land17_params(**kw):
    """Land17 parameter block [40] (LAND17_KEYS, keyword = key name; CalciumTransientType
    'Coppini' or 'Brixius', rateDependancy 'ON'/'OFF' or 1/0)"""
    vals = dict(LAND17_KEYS)
    for k, v in kw.items():
        if k not in vals:
            raise KeyError(f"unknown Land17 parameter {k}")
        if k == "CalciumTransientType":
            if v not in _CALCIUM:
                raise ValueError("Land17: CalciumTransientType must be Coppini or Brixius (External and Elphy "
                                 "need a calcium source this build does not have)")
            v = _CALCIUM[v]
        elif k == "rateDependancy" and isinstance(v, str):
            v = 1.0 if v.upper() == "ON" else 0.0
        vals[k] = float(v)
    out = [vals[k] for k, _ in LAND17_KEYS]
    return out + [0.0] * (LAND17_NPARAMS - len(out))


def land17_calcium_tables():
    """the Coppini and Brixius calcium transients the device uses (csrc/land17_calcium.h)"""
    import os
    import re

    text = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "land17_calcium.h")).read()
    body = text[text.index("kLand17Cai[2][kLand17CaSamples] = {"):]
    blocks = body.split("},")
    vals = [[float(v) for v in re.findall(r"\d+\.\d+", b)] for b in blocks[:2]]
    return vals[0], vals[1]


def with_land17(m: Material, tension_max=1000.0, **kw) -> Material:
    """the material with the Land17 tension model (Tmax converts its kPa to Pa: 1000)"""
    m.tension_max = float(tension_max)
    m.tension_fn = FN_LAND17
    m.fn = [0.0] * 12
    m.land17 = land17_params(**kw)
    return m


def function_params(kind: str, start=0.0, stop=0.0, keep_max=True, omega=0.0, phi=0.0,
                    period=0.8, contr_offset=0.215, relax_offset=0.362, contr_rate=1.32, relax_rate=21.9,
                    onset=0.0):
    fn = [start, stop, 1.0 if keep_max else 0.0, omega, phi, period, contr_offset, relax_offset,
          contr_rate, relax_rate, onset, 0.0]
    if kind == "DoubleHill":
        fn[11] = doublehill_max(fn)
    return _FN_NAMES[kind], fn


def doublehill_max(fn) -> float:
    """Normalisation of DoubleHill: max over 1001 samples of one period (CBAnalyticFunction.h:155-165)."""
    best = -math.inf
    for i in range(1001):
        t = i * fn[5] / 1000
        c = (t / fn[6]) ** fn[8]
        c = c / (1.0 + c)
        r = (t / fn[7]) ** fn[9]
        r = 1.0 / (1.0 + r)
        best = max(best, c * r)
    return best


def analytic(kind: int, fn, time: float, prev_time: float | None = None) -> float:
    """Host evaluation of an analytic time function (CBAnalyticFunction.h:53-145)."""
    start, stop = fn[0], fn[1]
    if kind == FN_LINEAR:
        prev = time if prev_time is None else prev_time
        if time < start:
            return 0.0
        if time < stop + 0.5 * (time - prev):
            return (time - start) * 1.0 / (stop - start)
        return 1.0 if fn[2] != 0.0 else 0.0
    if kind == FN_SINUS:
        return 0.5 - 0.5 * math.cos(fn[3] * (time - start) + fn[4]) if start <= time <= stop else 0.0
    if kind == FN_SINUSDRIVER:
        return math.sin(3.14159 * (time - start) / (stop - start)) if start <= time <= stop else 0.0
    if kind == FN_DOUBLEHILL:
        if not (start <= time <= stop):
            return 0.0
        t = time - fn[10]
        t = math.fmod(t, fn[5]) if t > 0.0 else 0.0
        c = (t / fn[6]) ** fn[8]
        c = c / (1.0 + c)
        r = (t / fn[7]) ** fn[9]
        r = 1.0 / (1.0 + r)
        return c * r / (fn[11] if fn[11] > 0 else doublehill_max(fn))
    return 0.0


def _get(params: dict, mat_index, key, default=None):
    own = f"Materials.Mat_{mat_index}.{key}"
    dflt = f"Materials.Mat_Default.{key}"
    if own not in params and dflt in params:
        return params[dflt]
    if own in params:
        return params[own]
    if default is None:
        raise KeyError(own)
    return default


def from_parameters(params: dict, mat_index) -> Material:
    """Build a Material from dotted reference keys (Appendix of the settings files)."""
    law = str(_get(params, mat_index, "Type", "Guccione"))
    density = float(_get(params, mat_index, "Density", 0.0))
    ignore = int(as_bool(_get(params, mat_index, "IgnoreCorruptElements", False)))
    if law == "Guccione":
        m = guccione(*(float(_get(params, mat_index, f"Guccione.{k}")) for k in ("C", "bf", "bt", "bfs", "K")),
                     density=density, a_scale=float(_get(params, mat_index, "Guccione.aScale", 1.0)),
                     b_scale=float(_get(params, mat_index, "Guccione.bScale", 1.0)), ignore_corrupt=ignore)
    elif law == "Holzapfel":
        vals = [float(_get(params, mat_index, f"Holzapfel.{k}")) for k in
                ("a", "b", "af", "bf", "as", "bs", "afs", "bfs", "kappa")]
        m = holzapfel(*vals, k=float(_get(params, mat_index, "Holzapfel.k", 100.0)), density=density,
                      ignore_corrupt=ignore)
    else:
        raise ValueError(f"material law {law} is not on this path (Guccione, Holzapfel)")
    tmodel = str(_get(params, mat_index, "TensionModel", "None"))
    if tmodel == "TanH":
        tmax = float(_get(params, mat_index, "TensionMax", 0.0))
        if tmax != 1000.0:     # CBTensionModelTanH.cpp:32-35
            raise ValueError("CBTensionModelTanH: TensionMax != 1000 is not compatible with this model")
        own, dflt = f"Materials.Mat_{mat_index}.TanH", "Materials.Mat_Default.TanH"

        def key(tail, d):
            if own + "." + tail not in params and dflt + "." + tail in params:
                # the reference reads the fallback key without its tail (InitKey, :18-24),
                # which is no value: its Get throws
                raise KeyError(dflt)
            return float(params.get(own + "." + tail, d))

        for k in ("kss", "knn", "ksn"):
            if key(k, 0.0) != 0.0:
                raise ValueError("TanH stress coefficients kss/knn/ksn are not on this path (only kff)")
        vals = {k: key(k, d) for k, d in TANH_KEYS}
        m.tension_max = tmax
        m.tension_fn, m.fn = tanh_params(kff=key("kff", 1.0), start_time=key("startTime", 0.0), **vals)
    elif tmodel == "FromFunction":
        # Tmax has no Mat_Default fallback (CBTensionFromFunction.cpp:44)
        m.tension_max = float(params[f"Materials.Mat_{mat_index}.TensionMax"])
        base = f"Materials.Mat_{mat_index}.FromFunction"
        kind = str(_get(params, mat_index, "FromFunction.Type", "None"))
        g = lambda k, d: float(params.get(f"{base}.{k}", d))  # noqa: E731
        m.tension_fn, m.fn = function_params(
            kind, start=g("StartTime", 0.0), stop=g("StopTime", 0.0),
            keep_max=as_bool(params.get(f"{base}.KeepMaxValue", True)), omega=g("Sinus.Omega", 0.0),
            phi=g("Sinus.Phi", 0.0), period=g("DoubleHill.Period", 0.8),
            contr_offset=g("DoubleHill.ContrTimeOffset", 0.215), relax_offset=g("DoubleHill.RelaxTimeOffset", 0.362),
            contr_rate=g("DoubleHill.ContrRateConst", 1.32), relax_rate=g("DoubleHill.RelaxRateConst", 21.9),
            onset=g("DoubleHill.OnsetTime", 0.0))
    elif tmodel == "Land17":
        # CBTensionModelLand17::InitParamsFromXml (CBTensionModelLand17.cpp:432-500); Tmax is
        # the material's TensionMax with its Mat_Default fallback (CBMaterial.cpp:39-42)
        own, dflt = f"Materials.Mat_{mat_index}.Land17", "Materials.Mat_Default.Land17"

        def key(tail, d):
            if own + "." + tail not in params and dflt + "." + tail in params:
                # InitKey reads the fallback key without its tail (:289-295): no value, Get throws
                raise KeyError(dflt)
            return float(params.get(own + "." + tail, d))

        for k in ("kss", "knn", "ksn"):
            if key(k, 0.0) != 0.0:
                raise ValueError("Land17 stress coefficients kss/knn/ksn are not on this path (only kff)")
        # the two switches have no Mat_Default fallback (:434, :445)
        vals = {k: key(k, d) for k, d in LAND17_KEYS if k not in ("CalciumTransientType", "rateDependancy")}
        vals["CalciumTransientType"] = str(params.get(own + ".CalciumTransientType", "Coppini"))
        rate = str(params.get(own + ".rateDependancy", "OFF"))
        vals["rateDependancy"] = 0.0 if rate == "OFF" else 1.0   # any other value switches it on (:368)
        with_land17(m, tension_max=float(_get(params, mat_index, "TensionMax", 0.0)), **vals)
    elif tmodel != "None":
        raise ValueError(f"tension model {tmodel} is not on this path (FromFunction, TanH, Land17)")
    return m


# ---- the workloads' materials (SI units) -------------------------------------

def cube_guccione() -> Material:
    """Unit-cube benchmark: C=0.88 kPa, bf=8, bt=6, bfs=12, kappa=1e3 kPa."""
    return guccione(880.0, 8.0, 6.0, 12.0, 1e6, density=1082.0)


def lv_holzapfel(tension_max=60e3, twitch=0.3) -> Material:
    """LV beat: Holzapfel-Ogden a=0.33 kPa, b=9.24, af=18.5 kPa, bf=15.97, as=2.56 kPa, bs=10.45,
    afs=0.42 kPa, bfs=11.6, kappa=1e3 kPa, k=100; active Tmax * sin^2 (Sinus over one twitch)."""
    m = holzapfel(330.0, 9.24, 18500.0, 15.97, 2560.0, 10.45, 420.0, 11.6, 1e6, k=100.0, density=1082.0)
    m.tension_max = tension_max
    m.tension_fn, m.fn = function_params("Sinus", start=0.0, stop=twitch, omega=2 * math.pi / twitch, phi=0.0)
    return m


def guccione_sweep(n_sets=64, seed=0, K=1e5, density=1082.0):
    """Passive-parameter sweep sets: C log-uniform in [0.2, 2] kPa, bf U[4, 16], bt U[2, 8],
    bfs U[4, 16], drawn from numpy.random.default_rng(seed) in that order (one array of
    n_sets per parameter); penalty K = 1e5 Pa as the reference's LV inflation problem
    (examples/benchmark2015/Problem2.xml:38).  With K = 1e6 Pa the softest sets
    (C ~ 0.2 kPa, K/C ~ 5000) stall in damped Newton already at 100 Pa: the Newton
    direction is isochoric only to first order and the penalty of its second-order
    volume change dominates the residual along it."""
    import numpy as np

    rng = np.random.default_rng(seed)
    C = np.exp(rng.uniform(np.log(200.0), np.log(2000.0), n_sets))
    bf = rng.uniform(4.0, 16.0, n_sets)
    bt = rng.uniform(2.0, 8.0, n_sets)
    bfs = rng.uniform(4.0, 16.0, n_sets)
    return [guccione(float(C[i]), float(bf[i]), float(bt[i]), float(bfs[i]), K, density=density)
            for i in range(n_sets)]
