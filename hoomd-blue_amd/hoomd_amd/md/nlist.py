"""Neighbour lists (hoomd/md/nlist.py). ``Cell`` is the binned list (NeighborListGPUBinned)."""

import numpy as np

from .. import _lib


class NeighborList:
    def __init__(self, buffer, exclusions=("bond",), rebuild_check_delay=1, check_dist=True):
        self.buffer = float(buffer)
        self.exclusions = tuple(exclusions)
        if int(rebuild_check_delay) < 0:
            raise ValueError("rebuild_check_delay must be non-negative")
        for e in self.exclusions:
            if e != "bond":
                raise NotImplementedError(f"exclusion '{e}' is not supported (only 'bond')")
        self.rebuild_check_delay = int(rebuild_check_delay)
        self.check_dist = bool(check_dist)
        self._simulation = None

    def _signature(self):
        return (self.buffer, self.exclusions, self.rebuild_check_delay, self.check_dist)

    @property
    def pair_list(self):
        """(tag_i, tag_j) pairs with tag_i < tag_j of the current list (NeighborList.h
        getPairList), as an array [P, 2]."""
        sim = self._simulation
        if sim is None or sim._system is None:
            raise RuntimeError("pair_list is available after the simulation has run")
        n = _lib.lib().hm_sys_get_pairs(sim._system, None, 0)
        _lib.check(n, "pair_list")
        out = np.zeros((max(n, 1), 2), np.uint32)
        _lib.check(_lib.lib().hm_sys_get_pairs(sim._system, out.ctypes.data, n), "pair_list")
        return out[:n]

    def _stats(self):
        sim = self._simulation
        if sim is None or sim._system is None:
            return None
        st = np.zeros(9, np.uint64)
        _lib.check(_lib.lib().hm_sys_stats(sim._system, st.ctypes.data), "stats")
        if sim._run_stats is not None:  # a run split at writer steps
            st[6], st[7], st[8] = sim._run_stats
        return st

    @property
    def num_builds(self):
        """Neighbour-list builds in the last Simulation.run (nlist.py:317-324)."""
        st = self._stats()
        return 0 if st is None else int(st[6])

    @property
    def shortest_rebuild(self):
        """Shortest number of steps between distance-check rebuilds in the last run
        (nlist.py:308-315, NeighborList::getSmallestRebuild)."""
        st = self._stats()
        return 100 if st is None else int(st[7])

    @property
    def num_dangerous_builds(self):
        """Builds taken on the first check after a build with rebuild_check_delay > 1
        (the reference's \"Dangerous neighborlist build\" notice, NeighborList.cc:1405-1414)."""
        st = self._stats()
        return 0 if st is None else int(st[8])


class Cell(NeighborList):
    """hoomd.md.nlist.Cell(buffer, exclusions=('bond',), rebuild_check_delay=1, check_dist=True)."""

    _log_namespace = ("md", "nlist", "Cell")
    _loggables = {"shortest_rebuild": ("scalar", True), "num_builds": ("scalar", False)}

    def __init__(self, buffer, exclusions=("bond",), rebuild_check_delay=1, check_dist=True,
                 deterministic=False, mesh=None):
        super().__init__(buffer, exclusions, rebuild_check_delay, check_dist)
        # the GPU build is deterministic by construction (cells ordered by particle index)
        self.deterministic = deterministic
