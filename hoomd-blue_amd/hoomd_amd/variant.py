"""hoomd.variant subset: scalar functions of the timestep (hoomd/variant/scalar.py,
hoomd/Variant.h). A variant is called with an integer timestep and returns a float; a plain
number converts to Constant."""


class Variant:
    """A scalar of the timestep: ``v(timestep) -> float``."""

    def __call__(self, timestep):
        raise NotImplementedError

    @property
    def min(self):
        raise NotImplementedError

    @property
    def max(self):
        raise NotImplementedError


class Constant(Variant):
    """The same value at every timestep (VariantConstant, hoomd/Variant.h:44-80)."""

    def __init__(self, value):
        self.value = float(value)

    def __call__(self, timestep):
        return self.value

    @property
    def min(self):
        return self.value

    @property
    def max(self):
        return self.value


class Ramp(Variant):
    """A until t_start, then linear to B over t_ramp steps, then B (VariantRamp,
    hoomd/Variant.h:82-170)."""

    def __init__(self, A, B, t_start, t_ramp):
        if t_ramp < 1:
            raise ValueError("t_ramp must be at least 1")
        self.A, self.B = float(A), float(B)
        self.t_start, self.t_ramp = int(t_start), int(t_ramp)

    def __call__(self, timestep):
        t = int(timestep)
        if t < self.t_start:
            return self.A
        if t >= self.t_start + self.t_ramp:
            return self.B
        return self.A + (self.B - self.A) * (t - self.t_start) / self.t_ramp

    @property
    def min(self):
        return min(self.A, self.B)

    @property
    def max(self):
        return max(self.A, self.B)


def as_variant(value):
    if isinstance(value, Variant):
        return value
    if isinstance(value, (int, float)):
        return Constant(value)
    raise TypeError(f"{value!r} is not variant_like")
