"""gala_amd: an MI355X-native orbit-integration engine with gala's API.

The hot path of gala (potential forces, Leapfrog and DOP853 orbit
integration, the mock-stream release-and-integrate loop, the SCF basis
expansion, and a learned SCF-force surrogate) runs in hand-written HIP
kernels for gfx950 behind the C ABI of include/gala_amd.h; this package is
the gala-shaped Python host layer above it.
"""
from . import units
from .units import G, galactic

__version__ = "0.1.0"

_EXPORTS = {
    "potential": ["PotentialBase", "HernquistPotential", "PlummerPotential",
                  "MiyamotoNagaiPotential", "NFWPotential", "NullPotential",
                  "CompositePotential", "SCFPotential", "Hamiltonian", "StaticFrame",
                  "MilkyWayLikePotential"],
    "integrate": ["parse_time_specification", "get_integrator", "LeapfrogIntegrator",
                  "DOPRI853Integrator", "Ruth4Integrator", "RK5Integrator"],
    "dynamics": ["PhaseSpacePosition", "Orbit", "MockStream", "FardalStreamDF",
                 "StreaklineStreamDF", "MockStreamGenerator", "DirectNBody"],
}


def __getattr__(name):
    if name == "io":  # potential YAML files
        import importlib
        return importlib.import_module(".io", __name__)
    for mod, names in _EXPORTS.items():
        if name in names:
            import importlib
            return getattr(importlib.import_module(f".{mod}", __name__), name)
    raise AttributeError(name)
