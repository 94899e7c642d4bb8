"""Read and write potentials as YAML (gala's potential/potential/io.py:1-366).

The file layout is gala's: ``class``, ``units`` (a mapping of physical type to
unit name), ``parameters`` (values, with an optional ``<name>_unit`` key
each) and, for composites, ``type: composite`` with a ``components`` list of
named component mappings (to_dict / from_dict, io.py:78-262).  Files written
by gala load here and files written here load in gala.

Differences, all because astropy is not part of this engine:
  * units are plain strings converted with a small table (lengths kpc, pc,
    Mpc; times Myr, Gyr, yr; masses solMass/Msun; speeds as length / time or
    km / s; angles rad, deg); anything else raises;
  * a component's origin and rotation, which gala's writer drops, are written
    under ``extra_args`` (which gala's reader passes to the class as keyword
    arguments, io.py:70-73) when they are not the identity;
  * YAML is read with a safe loader; the one Python tag gala's own writer
    emits for parameter mappings (``!!python/object/apply:collections.
    OrderedDict``, see tests/potential/potential/lm10.yml in gala) is accepted
    as a plain mapping, and no other tag constructs objects.
"""
import os

import numpy as np
import yaml

from . import potential as gp
from .units import G as G_GALACTIC
from .units import UnitSystem, dimensionless, galactic

__all__ = ["load", "save", "from_dict", "to_dict"]

# each unit in the base units of the galactic system (kpc, Myr, Msun, rad)
_LENGTH = {"kpc": 1.0, "pc": 1e-3, "Mpc": 1e3}
_TIME = {"Myr": 1.0, "Gyr": 1e3, "yr": 1e-6}
_MASS = {"solMass": 1.0, "Msun": 1.0}
_ANGLE = {"rad": 1.0, "radian": 1.0, "deg": np.pi / 180.0}
_KMS = 1.022712165045695e-3  # km/s in kpc/Myr

_CLASSES = {"HernquistPotential": gp.HernquistPotential,
            "PlummerPotential": gp.PlummerPotential,
            "MiyamotoNagaiPotential": gp.MiyamotoNagaiPotential,
            "NFWPotential": gp.NFWPotential,
            "NullPotential": gp.NullPotential,
            "SCFPotential": gp.SCFPotential}
_COMPOSITES = ("CompositePotential", "CCompositePotential")

# the physical type of each dimensional parameter, written as <name>_unit
_PTYPE = {"HernquistPotential": {"m": "mass", "c": "length"},
          "PlummerPotential": {"m": "mass", "b": "length"},
          "MiyamotoNagaiPotential": {"m": "mass", "a": "length", "b": "length"},
          "NFWPotential": {"m": "mass", "r_s": "length"},
          "SCFPotential": {"m": "mass", "r_s": "length"}}


def _scale(unit):
    """(physical type, factor to the galactic base unit) of a unit string."""
    u = unit.strip()
    if u == "":
        return "dimensionless", 1.0
    if u in _LENGTH:
        return "length", _LENGTH[u]
    if u in _TIME:
        return "time", _TIME[u]
    if u in _MASS:
        return "mass", _MASS[u]
    if u in _ANGLE:
        return "angle", _ANGLE[u]
    if u.replace(" ", "") in ("km/s",):
        return "speed", _KMS
    if "/" in u:
        num, den = (s.strip() for s in u.split("/", 1))
        if num in _LENGTH and den in _TIME:
            return "speed", _LENGTH[num] / _TIME[den]
    raise ValueError(f"unsupported unit {unit!r}")


def _unit_system(d):
    """gala's unit-system mapping -> UnitSystem (None: dimensionless)."""
    if d is None:
        return dimensionless
    d = {str(k): str(v) for k, v in d.items()}
    L, T, M = d.get("length", "kpc"), d.get("time", "Myr"), d.get("mass", "solMass")
    if L not in _LENGTH or T not in _TIME or M not in _MASS:
        raise ValueError(f"unsupported unit system {d}")
    if (L, T, M) == ("kpc", "Myr", "solMass") or (L, T, M) == ("kpc", "Myr", "Msun"):
        return galactic
    # G scales as length^3 / (mass time^2)
    Gs = G_GALACTIC * _MASS[M] * _TIME[T] ** 2 / _LENGTH[L] ** 3
    return UnitSystem(length=L, time=T, mass=M, angle="radian", G=Gs)


def _base_factor(units, ptype):
    """galactic base units per unit of `units` for one physical type"""
    if units is dimensionless or ptype == "dimensionless":
        return 1.0
    names = {"length": units["length"], "time": units["time"], "mass": units["mass"]}
    if ptype in names:
        return _scale(names[ptype])[1]
    if ptype == "speed":
        return _LENGTH[units["length"]] / _TIME[units["time"]]
    return 1.0


def _unpack_params(p, units):
    """Parameter values in the potential's units; <name>_unit keys convert
    (_unpack_params, io.py:15-35)."""
    out = {}
    for key, item in p.items():
        if key.endswith("_unit"):
            continue
        if np.iterable(item) and not isinstance(item, str):
            val = np.array(item, dtype=np.float64)
        else:
            val = float(item)
        unit = p.get(key + "_unit")
        if unit is not None:
            ptype, f = _scale(str(unit))
            val = val * f / _base_factor(units, ptype)
        out[key] = val
    return out


def _parse_component(d):
    cls = d.get("class")
    if cls is None:
        raise KeyError("Potential dictionary must contain a key 'class' for specifying the "
                       "name of the Potential class.")
    if cls not in _CLASSES:
        raise NotImplementedError(f"{cls} is not one of the potentials this engine evaluates "
                                  f"({', '.join(sorted(_CLASSES))})")
    units = _unit_system(d.get("units"))
    params = _unpack_params(dict(d.get("parameters") or {}), units)
    extra = dict(d.get("extra_args") or {})
    for k in ("origin", "R"):
        if k in extra:
            extra[k] = np.array(extra[k], dtype=np.float64)
    return _CLASSES[cls](units=units, **params, **extra)


def from_dict(d):
    """A potential from gala's dictionary form (io.py:78-161)."""
    if d.get("type") == "composite" or d.get("class") in _COMPOSITES:
        comps = [(c.get("name", str(i)), _parse_component(c))
                 for i, c in enumerate(d.get("components", []))]
        units = comps[0][1].units if comps else galactic
        p = gp.CompositePotential(units=units)
        for name, c in comps:
            p[name] = c
        return p
    if d.get("type") == "custom":
        raise NotImplementedError(f"custom composite {d.get('class')} is not supported")
    return _parse_component(d)


def _units_dict(units):
    if units is dimensionless or isinstance(units, type(dimensionless)):
        return None
    L, T, M = units["length"], units["time"], units["mass"]
    return {"angle": "rad", "angular speed": "mas / yr", "length": L,
            "mass": "solMass" if M in ("Msun", "solMass") else M, "speed": "km / s", "time": T}


def _to_dict_help(p):
    d = {"class": type(p).__name__}
    ud = _units_dict(p.units)
    if ud is not None:
        d["units"] = ud
    params = {}
    for k, v in p.parameters.items():
        params[k] = float(v)
        ptype = _PTYPE.get(type(p).__name__, {}).get(k)
        if ud is not None and ptype is not None:
            params[k + "_unit"] = ud[ptype]
    if isinstance(p, gp.SCFPotential):
        params["Snlm"] = p.Snlm.tolist()
        params["Tnlm"] = p.Tnlm.tolist()
    if params:
        d["parameters"] = params
    extra = {}
    if np.any(p.origin != 0.0):
        extra["origin"] = p.origin.tolist()
    if np.any(p.R != np.eye(3)):
        extra["R"] = p.R.tolist()
    if extra:
        d["extra_args"] = extra
    return d


def to_dict(potential):
    """gala's dictionary form of a potential (io.py:165-262)."""
    if isinstance(potential, gp.CompositePotential):
        comps = []
        for k, p in potential.items():
            c = _to_dict_help(p)
            c["name"] = k
            comps.append(c)
        return {"class": "CompositePotential", "components": comps, "type": "composite"}
    return _to_dict_help(potential)


class _Loader(yaml.SafeLoader):
    """Safe YAML plus gala's OrderedDict parameter tag, read as a mapping."""


def _ordered_dict(loader, node):
    # !!python/object/apply:collections.OrderedDict {dictitems: {...}}
    mapping = loader.construct_mapping(node, deep=True)
    items = mapping.get("dictitems", {})
    if not isinstance(items, dict):
        raise yaml.constructor.ConstructorError(None, None, "malformed OrderedDict", node.start_mark)
    return dict(items)


_Loader.add_constructor("tag:yaml.org,2002:python/object/apply:collections.OrderedDict",
                        _ordered_dict)


def load(f):
    """A potential from a YAML file path, a file-like object or YAML text
    (io.py:268-310)."""
    if hasattr(f, "read"):
        text = f.read()
    elif isinstance(f, (str, os.PathLike)) and os.path.exists(str(f)):
        with open(os.path.abspath(str(f)), encoding="utf-8") as fil:
            text = fil.read()
    else:
        text = str(f)
    return from_dict(yaml.load(text, Loader=_Loader))


def save(potential, f):
    """Write a potential as YAML to a path or file-like object (io.py:313-366)."""
    d = to_dict(potential)
    if hasattr(f, "write"):
        yaml.safe_dump(d, f, default_flow_style=None)
    else:
        with open(f, "w", encoding="utf-8") as f2:
            yaml.safe_dump(d, f2, default_flow_style=None)
