"""Thermostats (hoomd/md/methods/thermostats.py)."""


class Thermostat:
    pass


class Bussi(Thermostat):
    """hoomd.md.methods.thermostats.Bussi(kT, tau=0.0): stochastic velocity rescaling
    (Thermostat.h:337-446). tau = 0 draws the kinetic energy from the canonical distribution
    every step (the reference default, thermostats.py:321)."""

    def __init__(self, kT, tau=0.0):
        self.kT = kT
        self.tau = tau

    def _signature(self):
        return ("Bussi", float(self.kT), float(self.tau))
