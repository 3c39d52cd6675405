"""Simulation and Operations (hoomd/simulation.py, hoomd/operations.py).

``Simulation.run`` hands the whole step loop to the native engine (hm_sys_run): the Python layer
only describes the system, exactly as hoomd.Simulation hands it to System::run
(hoomd/System.cc:80-181).
"""

import time
import warnings

import numpy as np

from . import _lib
from .state import State


class Operations:
    """hoomd.Operations: integrator, tuners (a ParticleSorter by default), computes."""

    def __init__(self, simulation):
        from .tune import ParticleSorter
        self._simulation = simulation
        self._integrator = None
        self.tuners = [ParticleSorter()]  # operations.py:75 adds one by default
        self.computes = []
        self.updaters = []
        self.writers = []

    @property
    def integrator(self):
        return self._integrator

    @integrator.setter
    def integrator(self, value):
        self._integrator = value
        # forces know their simulation from here on (methods such as a learned force's
        # descriptors() attach the simulation themselves when first called)
        for f in getattr(value, "forces", None) or []:
            f._simulation = self._simulation
        self._simulation._invalidate()

    def add(self, op):
        from .md.compute import ThermodynamicQuantities
        from .md.tune import NeighborListBuffer
        from .tune import ParticleSorter
        from .write import GSD, Table
        if isinstance(op, (GSD, Table)):
            self.writers.append(op)
            return
        if isinstance(op, (ParticleSorter, NeighborListBuffer)):
            self.tuners.append(op)
        elif isinstance(op, ThermodynamicQuantities):
            # a compute reads the state on demand (its virial included): the running system
            # stays as it is
            self.computes.append(op)
            op._simulation = self._simulation
            return
        else:
            raise TypeError(f"hoomd_amd.Operations cannot hold {type(op).__name__}")
        self._simulation._invalidate()

    def __iadd__(self, op):
        self.add(op)
        return self

    @property
    def is_tuning_complete(self):
        """True once the autotuner has chosen the pair kernel's launch configuration
        (operations.py:346-373); the sweep runs at the start of the first run()."""
        sim = self._simulation
        if sim._system is None:
            return False
        out = np.zeros(4, np.uint32)
        _lib.check(_lib.lib().hm_sys_pair_config(sim._system, out.ctypes.data), "pair config")
        return bool(out[3])

    def tune_kernel_parameters(self):
        """Request a new autotuner sweep at the start of the next run() (operations.py:375-392)."""
        sim = self._simulation
        if sim._system is None:
            raise RuntimeError("Call Simulation.run() before tune_kernel_parameters.")
        _lib.check(_lib.lib().hm_sys_set_autotune(sim._system, 1), "autotune")

    def remove(self, op):
        if op in self.writers:
            self.writers.remove(op)
            op.close()
            return
        if op in self.computes:
            self.computes.remove(op)
            return
        if op in self.tuners:
            self.tuners.remove(op)
        self._simulation._invalidate()


class Simulation:
    """hoomd.Simulation(device, seed)."""

    _log_namespace = ("Simulation",)
    _loggables = {"timestep": ("scalar", True), "seed": ("scalar", True), "tps": ("scalar", True),
                  "walltime": ("scalar", True)}

    def __init__(self, device, seed=None):
        self.device = device
        self._seed = 0 if seed is None else int(seed)
        self._state = None
        self._system = None
        self._attached_key = None
        self._timestep = 0
        self._timestep_set = False
        self._tps = None
        self._walltime = 0.0
        self._configured = False
        self._run_stats = None
        self._run_id = 0
        self.operations = Operations(self)

    # -------------------------------------------------------------- properties
    @property
    def seed(self):
        return self._seed

    @seed.setter
    def seed(self, v):
        self._seed = int(v)
        self._invalidate()

    @property
    def state(self):
        return self._state

    @property
    def timestep(self):
        if self._system is not None:
            return int(_lib.lib().hm_sys_timestep(self._system))
        return self._timestep

    @timestep.setter
    def timestep(self, v):
        if self._state is not None:
            raise RuntimeError("Cannot set timestep after the state is initialised")
        self._timestep = int(v)
        self._timestep_set = True

    @property
    def tps(self):
        return self._tps

    @property
    def walltime(self):
        return self._walltime

    # -------------------------------------------------------------- state
    def create_state_from_snapshot(self, snapshot):
        if self._state is not None:
            raise RuntimeError("Cannot initialize more than once")
        self._state = State(self, snapshot)

    def create_state_from_gsd(self, filename, frame=-1):
        """Simulation.create_state_from_gsd (hoomd/simulation.py:188-240): the state from one
        frame of a GSD file. Unless the timestep was set before, it becomes the frame's
        configuration/step."""
        if self._state is not None:
            raise RuntimeError("Cannot initialize more than once")
        from . import gsd
        snap = gsd.read_snapshot(filename, frame)
        if not self._timestep_set:
            self._timestep = snap._step
        self._state = State(self, snap)

    def _invalidate(self):
        """Operations changed: the native system is rebuilt (from the current state) on next use."""
        self._attached_key = None

    def _destroy_system(self):
        if self._system is not None:
            self._state._sync_from_device()
            # forces whose device state the host copy does not follow (trained parameters)
            for f in getattr(self, "_attached_forces", ()):
                pull = getattr(f, "_pull", None)
                if pull is not None and getattr(f, "_simulation", None) is self:
                    pull()
            self._attached_forces = ()
            self._timestep = int(_lib.lib().hm_sys_timestep(self._system))
            _lib.lib().hm_sys_destroy(self._system)
            self._system = None
            self._configured = False

    def _ensure_system(self):
        if self._state is None:
            raise RuntimeError("Call create_state_from_snapshot first")
        if self._system is not None:
            return
        snap = self._state._snapshot
        p = snap.particles
        N = p.N
        ntypes = len(p.types)
        if ntypes > _lib.MAX_TYPES:
            raise ValueError(f"at most {_lib.MAX_TYPES} particle types are supported")
        pos = np.ascontiguousarray(p.position, np.float64)
        vel = np.ascontiguousarray(p.velocity, np.float64)
        tid = np.ascontiguousarray(p.typeid, np.uint32)
        mass = np.ascontiguousarray(p.mass, np.float64)
        img = np.ascontiguousarray(p.image, np.int32)
        h = _lib.lib().hm_sys_create(N, _lib.box(snap.configuration.box[:3]), ntypes,
                                     pos.ctypes.data, vel.ctypes.data, tid.ctypes.data,
                                     mass.ctypes.data, img.ctypes.data, self._seed & 0xffff)
        if not h:
            raise _lib.HoomdAmdError(-1, _lib.lib().hm_last_error().decode())
        self._system = h
        self._configured = False
        _lib.check(_lib.lib().hm_sys_set_timestep(h, self._timestep), "timestep")

    def _op_key(self):
        integ = self.operations.integrator
        return (id(integ), integ._signature(),
                tuple((type(t).__name__, t._signature()) for t in self.operations.tuners))

    def _set_nlist_buffer(self, nl):
        """A new buffer for the running system's neighbour list (NeighborList::setRBuff): the
        engine rebuilds its cells and list at the next step."""
        if self._system is None or not self._configured:
            return
        _lib.check(_lib.lib().hm_sys_set_nlist(self._system, nl.buffer,
                                               int("bond" in nl.exclusions)), "nlist buffer")
        self._attached_key = self._op_key()

    def _attach(self):
        integ = self.operations.integrator
        if integ is None:
            raise RuntimeError("Simulation.run needs an md.Integrator")
        key = self._op_key()
        for c in self.operations.computes:
            c._simulation = self
        if self._system is not None and self._configured and self._attached_key == key:
            for f in integ.forces:
                if hasattr(f, "_sync"):
                    f._sync(self)
            return
        if self._system is not None and self._configured:
            self._destroy_system()
        self._ensure_system()
        L = _lib.lib()
        h = self._system
        integ._attach(self, h)
        self._attached_forces = tuple(integ.forces)
        sorter = [t for t in self.operations.tuners if type(t).__name__ == "ParticleSorter"]
        if sorter:
            s = sorter[-1]
            _lib.check(L.hm_sys_set_sorter(h, s._period(), s._grid()), "sorter")
        else:
            _lib.check(L.hm_sys_set_sorter(h, 0, 256), "sorter")
        need_virial = any(c._needs_virial() for c in self.operations.computes)
        _lib.check(L.hm_sys_set_compute_virial(h, int(need_virial)), "virial")
        self._attached_key = key
        self._configured = True

    # -------------------------------------------------------------- run
    def run(self, steps, write_at_start=False):
        """Simulation.run (System::run, System.cc:80-181). The engine runs the steps between two
        operation steps in one call: Python tuners act before the step at their trigger's
        steps, writers after the step counter reaches their step (and on the starting step too
        when write_at_start is set)."""
        steps = int(steps)
        if steps < 0:
            raise ValueError("steps must be non-negative")
        self._attach()
        L = _lib.lib()
        writers = self.operations.writers
        tuners = [t for t in self.operations.tuners if hasattr(t, "_act")]
        self._run_stats = None
        self._run_id += 1
        t0 = time.perf_counter()
        if not writers and not tuners:
            _lib.check(L.hm_sys_run(self._system, steps), "run")
        else:
            # the engine restarts its build statistics at each call: keep them over the run
            agg = np.array([0, 100, 0], np.uint64)
            st = np.zeros(9, np.uint64)
            t = self.timestep
            end = t + steps
            if write_at_start:
                self._write(t)
            while t < end:
                for tu in tuners:  # tuners act before the step (System.cc:131-137)
                    if tu.trigger(t):
                        tu._act(self, t)
                nxt = [w.trigger.next_after(t) for w in writers]
                nxt += [tu.trigger.next_after(t) for tu in tuners]
                stop = min([s for s in nxt if s is not None] + [end])
                _lib.check(L.hm_sys_run(self._system, stop - t), "run")
                _lib.check(L.hm_sys_stats(self._system, st.ctypes.data), "stats")
                agg[0] += st[6]
                agg[1] = min(agg[1], st[7])
                agg[2] += st[8]
                t = stop
                self._write(t)
            self._run_stats = agg
        _lib.check(L.hm_synchronize(), "run")
        dt = time.perf_counter() - t0
        self._walltime += dt
        self._tps = steps / dt if dt > 0 and steps else self._tps
        self._warn_dangerous_builds()

    def _write(self, t):
        for w in self.operations.writers:
            if w.trigger(t):
                w._write(self._state, t)

    def _warn_dangerous_builds(self):
        """The reference's notice for builds taken on the first distance check after a build
        (NeighborList.cc:1405-1414)."""
        integ = self.operations.integrator
        for f in (integ.forces if integ is not None else []):
            nl = getattr(f, "nlist", None)
            if nl is not None and nl.num_dangerous_builds:
                warnings.warn("nlist: Dangerous neighborlist build occurred. Continuing this "
                              "simulation may produce incorrect results and/or program crashes. "
                              "Decrease the neighborlist check_period and rerun.", RuntimeWarning)
                return

    def __del__(self):
        h = getattr(self, "_system", None)
        if h is not None and _lib._LIB is not None:
            _lib._LIB.hm_sys_destroy(h)
            self._system = None
