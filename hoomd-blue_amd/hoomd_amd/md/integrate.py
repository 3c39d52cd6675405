"""md.Integrator (hoomd/md/integrate.py): forces, one integration method, time step."""

from .. import _lib


class Integrator:
    def __init__(self, dt, integrate_rotational_dof=False, forces=None, constraints=None,
                 methods=None, rigid=None, half_step_hook=None):
        if integrate_rotational_dof or constraints or rigid or half_step_hook:
            raise NotImplementedError("rotational DOF, constraints, rigid bodies and half-step "
                                      "hooks are outside the short-range MD timestep")
        self.dt = float(dt)
        self.forces = list(forces or [])
        self.methods = list(methods or [])

    def _signature(self):
        return (self.dt, tuple(f._signature() for f in self.forces),
                tuple(m._signature() for m in self.methods))

    def _attach(self, sim, h):
        L = _lib.lib()
        if len(self.methods) != 1:
            raise NotImplementedError("hoomd_amd integrates with exactly one method over all "
                                      "particles")
        if len(self.forces) > 4:
            raise NotImplementedError("at most 4 forces")
        nlists = {id(f.nlist): f.nlist for f in self.forces if hasattr(f, "nlist")}
        if len(nlists) > 1:
            raise NotImplementedError("all pair forces must share one neighbour list")
        if nlists:
            nl = next(iter(nlists.values()))
            _lib.check(L.hm_sys_set_nlist(h, nl.buffer, int("bond" in nl.exclusions)), "nlist")
            _lib.check(L.hm_sys_set_nlist_policy(h, int(nl.rebuild_check_delay),
                                                 int(bool(nl.check_dist))), "nlist policy")
        for f in self.forces:
            f._attach(sim, h)
        m = self.methods[0]
        from .methods import Langevin
        if isinstance(m, Langevin):
            kind, kT, param = m._method(sim.state.particle_types)
        else:
            kind, kT, param = m._method()
        import numpy as np
        arr = np.ascontiguousarray(param, np.float64)
        _lib.check(L.hm_sys_set_method(h, kind, self.dt, kT, arr.ctypes.data), "method")
