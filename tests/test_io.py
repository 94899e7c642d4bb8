"""Potential YAML files (gala's potential/potential/io.py), mirroring gala's
tests/potential/potential/test_io.py on the files that test reads, copied
as fixtures into tests/golden/potential_yaml/ (Plummer.yml, Composite.yml,
lm10.yml).  Runs on the CPU: constructing potentials touches no GPU."""
import io
from pathlib import Path

import numpy as np
import pytest

from gala_amd import io as gio
from gala_amd.potential import (CompositePotential, HernquistPotential, MiyamotoNagaiPotential,
                                NFWPotential, PlummerPotential, SCFPotential)
from gala_amd.units import G, galactic

HERE = Path(__file__).parent / "golden" / "potential_yaml"


def test_read_plummer():
    p = gio.load(HERE / "Plummer.yml")
    assert isinstance(p, PlummerPotential)
    assert np.allclose(p.parameters["m"], 1e11)
    assert np.allclose(p.parameters["b"], 0.26)
    assert p.units == galactic


def test_read_composite():
    p = gio.load(HERE / "Composite.yml")
    assert isinstance(p, CompositePotential)
    assert list(p.keys()) == ["halo", "disk"]
    assert isinstance(p["disk"], MiyamotoNagaiPotential)
    assert p["disk"].parameters["a"] == 6.5 and p.units == galactic


def test_gala_ordereddict_tag_is_read_safely():
    """lm10.yml (written by gala) carries !!python/object/apply:collections.
    OrderedDict parameter mappings: they parse as plain mappings; the file's
    LogarithmicPotential halo is not an engine potential and is refused."""
    import yaml
    d = yaml.load((HERE / "lm10.yml").read_text(), Loader=gio._Loader)
    disk = d["components"][0]["parameters"]
    assert disk == {"a": 10.0, "a_unit": "kpc", "b": 0.26, "b_unit": "kpc", "m": 150000.0,
                    "m_unit": "solMass"}
    with pytest.raises(NotImplementedError):
        gio.from_dict(d)
    # any other Python tag is refused, not executed
    with pytest.raises(yaml.constructor.ConstructorError):
        gio.load("!!python/object/apply:os.getcwd []")


def test_round_trip(tmp_path):
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    pot = CompositePotential()
    pot["bulge"] = HernquistPotential(m=5e9, c=1.0)
    pot["disk"] = MiyamotoNagaiPotential(m=6.8e10, a=3.0, b=0.28, R=R)
    pot["halo"] = NFWPotential(m=5.4e11, r_s=15.6)
    pot["blob"] = PlummerPotential(m=1e9, b=0.5, origin=[1.0, 2.0, 3.0])
    f = tmp_path / "pot.yml"
    gio.save(pot, str(f))
    back = gio.load(str(f))
    assert list(back.keys()) == list(pot.keys())
    for k in pot:
        assert type(back[k]) is type(pot[k])
        assert back[k].parameters == pot[k].parameters
        assert np.array_equal(back[k].origin, pot[k].origin)
        assert np.array_equal(back[k].R, pot[k].R)
    # the component tables the C ABI receives are identical
    for a, b in zip(back._components(), pot._components()):
        assert a["type"] == b["type"] and a["pars"] == b["pars"]
        assert a["shift_rotate"] == b["shift_rotate"]
        assert np.array_equal(a["q0"], b["q0"]) and np.array_equal(a["R"], b["R"])
    # file-like objects and YAML text
    buf = io.StringIO()
    gio.save(pot["bulge"], buf)
    assert gio.load(buf.getvalue()).parameters == pot["bulge"].parameters
    buf.seek(0)
    assert gio.load(buf).parameters == pot["bulge"].parameters


def test_scf_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    S = rng.normal(size=(4, 3, 3))
    T = rng.normal(size=(4, 3, 3))
    p = SCFPotential(m=1e12, r_s=20.0, Snlm=S, Tnlm=T)
    f = tmp_path / "scf.yml"
    gio.save(p, f)
    q = gio.load(f)
    assert np.array_equal(q.Snlm, S) and np.array_equal(q.Tnlm, T)
    assert q.parameters == p.parameters


def test_units_convert():
    """<name>_unit values convert into the potential's units; another unit
    system scales G (length^3 / (mass time^2))."""
    text = """
class: HernquistPotential
parameters: {m: 2.0e9, m_unit: solMass, c: 700.0, c_unit: pc}
units: {length: kpc, mass: solMass, time: Myr, angle: rad}
"""
    p = gio.load(text)
    assert p.parameters["c"] == pytest.approx(0.7) and p.parameters["m"] == 2e9
    q = gio.load(text.replace("time: Myr", "time: Gyr"))
    assert q.units.G == pytest.approx(G * 1e6)
    with pytest.raises(ValueError):
        gio.load(text.replace("c_unit: pc", "c_unit: lightyear"))
