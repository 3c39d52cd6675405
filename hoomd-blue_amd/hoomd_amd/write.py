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


class Table:
    """hoomd.write.Table(trigger, logger, output=stdout, header_sep='.', delimiter=' ',
    pretty=True, max_precision=10, max_header_len=None) (hoomd/write/table.py): one row of the
    logger's scalar and string quantities per trigger step, a header row first and whenever the
    set of quantities changes. Columns are at least max(10, max_precision + 6) wide and values
    are centred in them."""

    def __init__(self, trigger, logger, output=None, header_sep=".", delimiter=" ",
                 pretty=True, max_precision=10, max_header_len=None):
        import sys

        from .logging import LoggerCategories
        allowed = LoggerCategories.scalar | LoggerCategories.string
        if logger.categories & ~allowed & LoggerCategories.ALL:
            raise ValueError("Table Logger may only have scalar or string categories set. "
                             "Use hoomd.write.GSD for other categories.")
        self.trigger = as_trigger(trigger)
        self.logger = logger
        self.output = output if output is not None else sys.stdout
        self.header_sep = header_sep
        self.delimiter = delimiter
        self.pretty = bool(pretty)
        self.max_precision = int(max_precision)
        self.max_header_len = max_header_len
        self.min_column_width = max(10, self.max_precision + 6)
        self._widths = {}

    def _header(self, ns):
        if self.max_header_len is None:
            return self.header_sep.join(ns)
        # as many trailing namespace levels as fit (table.py:287-298)
        index, count = -1, len(ns[-1])
        for name in reversed(ns[:-1]):
            count += len(name)
            if count > self.max_header_len:
                break
            index -= 1
        return self.header_sep.join(ns[index:])

    def _fmt(self, value, width):
        import numbers
        if isinstance(value, str):
            return f"{value:^{width}}"
        if isinstance(value, numbers.Integral):
            return f"{int(value):^{width}d}"
        v = float(value)
        mag = len(str(int(abs(v)))) + (1 if v < 0 else 0) + 1 if v == v and abs(v) < 1e300 else 99
        if mag >= 6 or mag > width:
            dec = min(max(width - 6, 1), 5) if self.pretty else max(self.max_precision - 1, 0)
            return f"{v:^{width}.{dec}e}"
        dec = (min(max(width - mag - 2, 1), 5) if self.pretty
               else max(self.max_precision - mag, 0))
        return f"{v:^{width}.{dec}f}"

    def _write(self, state, step):
        from .logging import flatten
        data = {k: v[0] for k, v in flatten(self.logger.log()).items()}
        if data.keys() != self._widths.keys():
            self._widths = {k: max(len(self._header(k)), self.min_column_width) for k in data}
            self.output.write(self.delimiter.join(
                f"{self._header(k):^{w}}" for k, w in self._widths.items()) + "\n")
        self.output.write(self.delimiter.join(
            self._fmt(data[k], w) for k, w in self._widths.items()) + "\n")
        self.output.flush()

    def write(self):
        """Write a row now (hoomd/write/table.py Table.write)."""
        self._write(None, None)

    def flush(self):
        self.output.flush()

    def close(self):
        pass
