"""Solver and damping settings read like the reference's settings files."""
import pytest

from cardiomechanics_amd import settings as S
from cardiomechanics_amd import _lib

# the damping block of the reference's EM01 M_1mm.xml and JPhys scenario settings
XML = """<settings>
  <Solver><Type>NewmarkBeta</Type><Precision>1e-10</Precision><MaxIterations>30</MaxIterations>
    <NewmarkBeta><Beta>0.3</Beta><Gamma>0.6</Gamma><ConsistentMassMatrix>false</ConsistentMassMatrix></NewmarkBeta>
  </Solver>
  <Materials><Global><Damping><Rayleigh><Alpha>500.0</Alpha><Beta>0.01</Beta></Rayleigh></Damping></Global></Materials>
</settings>"""


def test_solver_options_from_settings():
    o = S.solver_options(S.parse_xml_string(XML))
    assert o["newton_atol"] == o["newton_rtol"] == o["newton_stol"] == 1e-10
    assert o["newton_maxit"] == 30 and o["newmark_beta"] == 0.3 and o["newmark_gamma"] == 0.6
    assert o["consistent_mass"] == 0
    assert o["rayleigh_alpha"] == 500.0 and o["rayleigh_beta"] == 0.01
    # every key is a field of the C ABI's options struct
    names = {f[0] for f in _lib.CmgOptions._fields_}
    assert set(o) <= names
    d = S.solver_options({})
    assert d["rayleigh_alpha"] == 0.0 and d["rayleigh_beta"] == 0.0 and d["newton_atol"] == 1e-8


def test_damping_type_is_read_under_the_singular_key():
    # the reference reads the type from "Material.Global..." (singular); a type given
    # under "Materials.Global.Damping.Type" is ignored
    p = S.parse_xml_string(XML)
    assert S.solver_options(dict(p, **{"Materials.Global.Damping.Type": "Viscous"}))["rayleigh_alpha"] == 500.0
    with pytest.raises(ValueError):
        S.solver_options(dict(p, **{"Material.Global.Damping.Type": "Viscous"}))
