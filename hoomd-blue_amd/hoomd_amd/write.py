"""Writers (hoomd/write): GSD files of the system's frames."""

from . import gsd as _gsd
from .filter import All
from .trigger import as_trigger


class GSD:
    """hoomd.write.GSD (hoomd/write/gsd.py, GSDDumpWriter.cc): append a frame to ``filename``
    on every step the trigger fires. ``dynamic`` names the chunks or categories ('property',
    'momentum', 'attribute', 'topology') written on every frame; the rest are written on frame
    0 only. The filter must be All (every particle)."""

    def __init__(self, trigger, filename, filter=All(), mode="ab", truncate=False,
                 dynamic=None, logger=None):
        if not isinstance(filter, All):
            raise ValueError("hoomd_amd.write.GSD writes every particle (filter=All())")
        if logger is not None:
            raise ValueError("hoomd_amd.write.GSD does not write logged quantities")
        if mode not in ("ab", "wb", "xb"):
            raise ValueError("mode must be 'ab', 'wb' or 'xb'")
        self.trigger = as_trigger(trigger)
        self.filename = filename
        self.filter = filter
        self.mode = mode
        self.truncate = bool(truncate)
        self.dynamic = list(dynamic) if dynamic is not None else ["property"]
        self._writer = None

    def _open(self):
        if self._writer is None:
            self._writer = _gsd.GSDWriter(self.filename, self.mode, self.dynamic)

    def _write(self, state, step):
        self._open()
        if self.truncate and self._writer.nframes:
            self._writer.close()
            self._writer = _gsd.GSDWriter(self.filename, "wb", self.dynamic)
        self._writer.write_frame(state.get_snapshot(), step)

    def flush(self):
        if self._writer is not None:
            self._writer._f.flush()

    def close(self):
        if self._writer is not None:
            self._writer.close()
            self._writer = None

    @staticmethod
    def write(state, filename, filter=All(), mode="wb", logger=None):
        """Write the state's current frame to a file (hoomd/write/gsd.py GSD.write)."""
        if not isinstance(filter, All):
            raise ValueError("hoomd_amd.write.GSD writes every particle (filter=All())")
        step = state._simulation.timestep
        w = _gsd.GSDWriter(filename, mode, ["property"])
        w.write_frame(state.get_snapshot(), step)
        w.close()
