"""Triggers (hoomd/trigger.py, hoomd/Trigger.h): boolean functions of the timestep that say when
an operation acts. Each also gives the next step after t on which it fires, so Simulation.run
can hand the native engine whole stretches of steps between two writes."""

_U64 = 1 << 64


class Trigger:
    def __call__(self, timestep):
        return self.compute(int(timestep))

    def next_after(self, t):
        """The smallest step > t on which the trigger fires, or None."""
        raise NotImplementedError


class Periodic(Trigger):
    """Fires when (t - phase) mod period == 0 in unsigned 64-bit arithmetic (Trigger.h:81-84)."""

    def __init__(self, period, phase=0):
        if int(period) < 1:
            raise ValueError("period must be positive")
        self.period = int(period)
        self.phase = int(phase)

    def compute(self, t):
        return ((t - self.phase) % _U64) % self.period == 0

    def next_after(self, t):
        # (s - phase) mod 2^64 grows by one per step on [t+1, phase) and on [phase, ...), and
        # the phase itself always fires
        r = ((t + 1 - self.phase) % _U64) % self.period
        s = t + 1 + ((self.period - r) % self.period)
        return s if t + 1 >= self.phase or s < self.phase else self.phase


class On(Trigger):
    """Fires on one timestep."""

    def __init__(self, timestep):
        self.timestep = int(timestep)

    def compute(self, t):
        return t == self.timestep

    def next_after(self, t):
        return self.timestep if self.timestep > t else None


def as_trigger(trigger):
    """An int stands for Periodic(int), as the reference's trigger type converter does."""
    if isinstance(trigger, Trigger):
        return trigger
    return Periodic(int(trigger))
