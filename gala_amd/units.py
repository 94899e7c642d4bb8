"""Unit system of the engine: gala's ``galactic`` (src/gala/units.py:379-387),
kpc, Myr, Msun, radian.  Without astropy, quantities are plain floats in
these units; velocities are kpc/Myr (multiply km/s by ``KMS_TO_KPCMYR``).
"""

# G in kpc^3 Msun^-1 Myr^-2: the IAU 2015 nominal GM_sun (1.3271244e20 m^3/s^2)
# times (Julian Myr)^2 / kpc^3; astropy's G * M_sun in gala's galactic units.
G = 4.498502151469553e-12

# 1 km/s in kpc/Myr (Julian year, IAU parsec)
KMS_TO_KPCMYR = 1.022712165045695e-3


class UnitSystem:
    """Minimal stand-in for gala.units.UnitSystem: names the base units and
    carries G in them."""

    def __init__(self, length="kpc", time="Myr", mass="Msun", angle="radian", G=G):
        self._units = {"length": length, "time": time, "mass": mass, "angle": angle,
                       "speed": f"{length}/{time}"}
        self.G = G

    def __getitem__(self, key):
        return self._units[key]

    def __repr__(self):
        return "UnitSystem(" + ", ".join(self._units[k] for k in ("length", "time", "mass",
                                                                    "angle")) + ")"

    def __eq__(self, other):
        return isinstance(other, UnitSystem) and self._units == other._units and self.G == other.G

    def __hash__(self):
        return hash(tuple(sorted(self._units.items())) + (self.G,))


galactic = UnitSystem()


class DimensionlessUnitSystem(UnitSystem):
    """Dimensionless units with G = 1."""

    def __init__(self):
        super().__init__("", "", "", "radian", G=1.0)


dimensionless = DimensionlessUnitSystem()
