"""md.compute.ThermodynamicQuantities (hoomd/md/compute.py -> ComputeThermo.cc:76-292)."""

import numpy as np

from .. import _lib


class ThermodynamicQuantities:
    _log_namespace = ("md", "compute", "ThermodynamicQuantities")
    _loggables = {"kinetic_temperature": ("scalar", True), "pressure": ("scalar", True),
                  "pressure_tensor": ("sequence", True), "kinetic_energy": ("scalar", True),
                  "potential_energy": ("scalar", True), "num_particles": ("scalar", True),
                  "translational_degrees_of_freedom": ("scalar", True),
                  "volume": ("scalar", True)}
    def __init__(self, filter):
        self.filter = filter
        self._simulation = None

    def _needs_virial(self):
        return False  # the virial is evaluated on demand when a pressure is read

    def _thermo(self):
        sim = self._simulation
        if sim is None or sim._system is None:
            raise RuntimeError("ThermodynamicQuantities needs an attached simulation")
        out = np.zeros(15)
        _lib.check(_lib.lib().hm_sys_thermo(sim._system, out.ctypes.data), "thermo")
        return out

    @property
    def num_particles(self):
        return self._simulation.state.N_particles

    @property
    def translational_degrees_of_freedom(self):
        integ = self._simulation.operations.integrator
        from .methods import Langevin
        N = self.num_particles
        if integ is not None and integ.methods and isinstance(integ.methods[0], Langevin):
            return 3.0 * N
        return 3.0 * N - 3.0

    degrees_of_freedom = translational_degrees_of_freedom

    @property
    def volume(self):
        b = self._simulation.state.box
        return b[0] * b[1] * b[2]

    @property
    def kinetic_energy(self):
        return float(self._thermo()[0])

    translational_kinetic_energy = kinetic_energy

    @property
    def potential_energy(self):
        return float(self._thermo()[1])

    @property
    def kinetic_temperature(self):
        return 2.0 * self.kinetic_energy / self.translational_degrees_of_freedom

    @property
    def pressure(self):
        o = self._thermo()
        return float((2.0 * o[0] / 3.0 + o[2]) / self.volume)

    @property
    def pressure_tensor(self):
        """(xx, xy, xz, yy, yz, zz) of (sum m v_a v_b + virial_ab) / V, summed on the device
        (ComputeThermo.cc:236-262)."""
        o = self._thermo()
        return tuple(float((k + w) / self.volume) for k, w in zip(o[9:15], o[3:9]))
