"""Integration methods (hoomd/md/methods/methods.py): ConstantVolume and Langevin."""

from . import thermostats  # noqa: F401
from ..typeparam import TypeParameter


class Method:
    pass


class ConstantVolume(Method):
    """hoomd.md.methods.ConstantVolume(filter, thermostat=None): velocity Verlet
    (TwoStepConstantVolume.cc:7-314), optionally rescaled by a Bussi thermostat."""

    def __init__(self, filter, thermostat=None):
        self.filter = filter
        self.thermostat = thermostat

    def _signature(self):
        return ("ConstantVolume", None if self.thermostat is None else self.thermostat._signature())

    def _method(self):
        if self.thermostat is None:
            return 0, 0.0, [0.0]
        return 1, float(self.thermostat.kT), [float(self.thermostat.tau)]


class Langevin(Method):
    """hoomd.md.methods.Langevin(filter, kT, tally_reservoir_energy=False, default_gamma=1.0)
    (TwoStepLangevin.cc:38-413)."""

    def __init__(self, filter, kT, tally_reservoir_energy=False, default_gamma=1.0,
                 default_gamma_r=(1.0, 1.0, 1.0)):
        if tally_reservoir_energy:
            raise NotImplementedError("tally_reservoir_energy is not supported")
        self.filter = filter
        self.kT = kT
        self.gamma = TypeParameter(float(default_gamma))
        self.gamma_r = TypeParameter(tuple(default_gamma_r))

    def _signature(self):
        return ("Langevin", float(self.kT), tuple(sorted(self.gamma.items())))

    def _method(self, types=None):
        types = types or ["A"]
        return 2, float(self.kT), [float(self.gamma[t]) for t in types]
