"""Reference-style XML settings files as dotted keys (ParameterMap,
mechanics/src/ParameterMap/ParameterMap.h:29-152): <Solver><Type>NewmarkBeta</Type>
becomes {"Solver.Type": "NewmarkBeta"}; the root tag (<settings>) is dropped and the
values stay text.  The same reading as the host C++ layer's ParameterMap::ParseXML.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET


def parse_xml_string(text: str) -> dict:
    root = ET.fromstring(text)
    out = {}

    def walk(node, prefix):
        for ch in node:
            key = f"{prefix}.{ch.tag}" if prefix else ch.tag
            if len(ch):
                walk(ch, key)
            else:
                out[key] = (ch.text or "").strip()

    walk(root, "")
    return out


def parse_xml(path: str) -> dict:
    with open(path) as fh:
        return parse_xml_string(fh.read())


def as_bool(v) -> bool:
    """ParameterMap's bool conversion: true / 1 / True / TRUE"""
    if isinstance(v, str):
        return v.strip() in ("true", "1", "True", "TRUE")
    return bool(v)


def solver_options(params: dict) -> dict:
    """MechanicsContext options from solver and damping settings, the mapping of the host
    C++ layer (CBSolverGPU): Solver.Precision as Newton abstol/rtol/stol
    (CBSolverNewmarkBeta.cpp:165), Solver.MaxIterations, Solver.NewmarkBeta.Beta /
    Gamma / ConsistentMassMatrix (:206-212), Solver.Epsilon, and Rayleigh damping read
    as the reference does (:274-288): the type under "Material.Global.Damping.Type"
    (only "Rayleigh"), the coefficients under "Materials.Global.Damping.Rayleigh"."""
    g = params.get
    prec = float(g("Solver.Precision", 1e-8))
    dtype = str(g("Material.Global.Damping.Type", "Rayleigh"))
    if dtype != "Rayleigh":
        raise ValueError(f"Global damping type {dtype} is unknown")
    return dict(newton_atol=prec, newton_rtol=prec, newton_stol=prec,
                newton_maxit=int(g("Solver.MaxIterations", 50)),
                newmark_beta=float(g("Solver.NewmarkBeta.Beta", 0.25)),
                newmark_gamma=float(g("Solver.NewmarkBeta.Gamma", 0.5)),
                consistent_mass=int(as_bool(g("Solver.NewmarkBeta.ConsistentMassMatrix", True))),
                fd_epsilon=float(g("Solver.Epsilon", 1e-12)),
                rayleigh_alpha=float(g("Materials.Global.Damping.Rayleigh.Alpha", 0.0)),
                rayleigh_beta=float(g("Materials.Global.Damping.Rayleigh.Beta", 0.0)))
