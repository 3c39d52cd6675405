"""GSD files: the HOOMD schema's configuration, particle and bond chunks.

The reference reads and writes GSD through the gsd library (hoomd/GSDReader.cc,
hoomd/GSDDumpWriter.cc; hoomd/write/gsd.py). This module reads and writes the same file layout
directly with numpy:

* a 256-byte header: magic, index location and capacity, namelist location and capacity,
  schema version, file-layer version (2.0), application and schema names;
* an index of 32-byte entries (frame, N, location, M, name id, type, flags), sorted by
  (frame, name id): the gsd library finds a chunk by binary search over that order
  (hoomd/extern/gsd.c:780-810);
* a namelist of null-terminated chunk names;
* the chunk data, each an N x M array of the entry's type.

Chunks that a frame does not store take their frame-0 value, then the schema default, as
GSDReader does (GSDReader.cc:97-176).
"""

import struct

import numpy as np

MAGIC = 0x65DF65DF65DF65DF
GSD_VERSION = (2 << 16) | 0
SCHEMA_VERSION = (1 << 16) | 4
HEADER = struct.Struct("<QQQQQII64s64s80s")
ENTRY = struct.Struct("<QQqIHBB")

# gsd_type codes and their numpy types (hoomd/extern/gsd.h:24-51 defines 1-10; code 11, the
# character type of later gsd releases, is read as bytes but never written)
_TYPES = {1: np.uint8, 2: np.uint16, 3: np.uint32, 4: np.uint64, 5: np.int8, 6: np.int16,
          7: np.int32, 8: np.int64, 9: np.float32, 10: np.float64, 11: np.uint8}
_CODES = {np.dtype(v): k for k, v in _TYPES.items() if k != 11}


class GSDFile:
    """Read access to the chunks of a GSD 2.x file."""

    def __init__(self, filename):
        with open(filename, "rb") as f:
            self._data = f.read()
        d = self._data
        if len(d) < HEADER.size:
            raise RuntimeError(f"{filename}: not a GSD file")
        (magic, index_loc, index_n, names_loc, names_n, self.schema_version, self.gsd_version,
         app, schema, _) = HEADER.unpack_from(d, 0)
        if magic != MAGIC:
            raise RuntimeError(f"{filename}: not a GSD file (bad magic)")
        if (self.gsd_version >> 16) != 2:
            raise RuntimeError(f"{filename}: GSD file layer {self.gsd_version >> 16}.x is not "
                               "supported (2.x only)")
        self.application = app.split(b"\0")[0].decode()
        self.schema = schema.split(b"\0")[0].decode()
        raw = d[names_loc:names_loc + names_n * 64]
        self._names = [n.decode() for n in raw.split(b"\0")]
        self._index = {}
        nframes = 0
        for k in range(index_n):
            frame, N, loc, M, nid, typ, _ = ENTRY.unpack_from(d, index_loc + k * ENTRY.size)
            if loc == 0:
                break
            self._index[(frame, self._names[nid])] = (N, M, loc, typ)
            nframes = max(nframes, frame + 1)
        self.nframes = nframes

    def __len__(self):
        return self.nframes

    def chunk_names(self, frame):
        return sorted(n for (f, n) in self._index if f == frame)

    def read_chunk(self, frame, name):
        """The chunk as an array of shape (N,) when M == 1, else (N, M); None if absent."""
        e = self._index.get((frame, name))
        if e is None:
            return None
        N, M, loc, typ = e
        a = np.frombuffer(self._data, dtype=_TYPES[typ], count=N * M, offset=loc).copy()
        return a if M == 1 else a.reshape(N, M)


def _types_from_chunk(a):
    if a is None:
        return None
    rows = a.reshape(a.shape[0], -1) if a.ndim > 1 else a.reshape(-1, 1)
    return [bytes(r).split(b"\0")[0].decode() for r in rows]


def read_snapshot(filename, frame=-1):
    """A Snapshot of one frame (negative frames count from the end), the defaults of the HOOMD
    schema filling chunks that neither this frame nor frame 0 store."""
    from .snapshot import Snapshot
    g = GSDFile(filename)
    if g.nframes == 0:
        raise RuntimeError(f"{filename}: no frames")
    i = frame + g.nframes if frame < 0 else frame
    if not 0 <= i < g.nframes:
        raise IndexError(f"frame {frame} is out of range ({g.nframes} frames)")

    def get(name):
        a = g.read_chunk(i, name)
        return g.read_chunk(0, name) if a is None and i != 0 else a

    s = Snapshot()
    box = get("configuration/box")
    s.configuration.box = [float(x) for x in box] if box is not None else [1, 1, 1, 0, 0, 0]
    dims = get("configuration/dimensions")
    s.configuration.dimensions = int(dims[0]) if dims is not None else 3
    n = get("particles/N")
    N = int(n[0]) if n is not None else 0
    p = s.particles
    p.N = N
    types = _types_from_chunk(get("particles/types"))
    p.types = types if types else ["A"]
    for name, attr, default in (("particles/position", "position", 0.0),
                                ("particles/velocity", "velocity", 0.0),
                                ("particles/image", "image", 0),
                                ("particles/typeid", "typeid", 0),
                                ("particles/mass", "mass", 1.0)):
        a = get(name)
        if a is not None:
            getattr(p, attr)[:] = a.reshape(getattr(p, attr).shape)
    nb = get("bonds/N")
    if nb is not None and int(nb[0]) > 0:
        b = s.bonds
        b.N = int(nb[0])
        b.types = _types_from_chunk(get("bonds/types")) or ["A-A"]
        tid = get("bonds/typeid")
        if tid is not None:
            b.typeid[:] = tid
        grp = get("bonds/group")
        if grp is not None:
            b.group[:] = grp.reshape(b.N, 2)
    s._step = int(get("configuration/step")[0]) if get("configuration/step") is not None else 0
    return s


def _type_chunk(names):
    width = max(1, max((len(n.encode()) for n in names), default=0) + 1)
    a = np.zeros((len(names), width), np.uint8)
    for k, n in enumerate(names):
        b = n.encode()
        a[k, :len(b)] = np.frombuffer(b, np.uint8)
    return a


# dynamic categories (GSDDumpWriter::setDynamic, GSDDumpWriter.cc:149-212)
_CATEGORY = {"configuration/box": "property", "particles/N": "property",
             "particles/position": "property", "particles/velocity": "momentum",
             "particles/image": "momentum", "particles/types": "attribute",
             "particles/typeid": "attribute", "particles/mass": "attribute",
             "bonds/N": "topology", "bonds/types": "topology", "bonds/typeid": "topology",
             "bonds/group": "topology"}


def _is_dynamic(name, dynamic):
    return name in dynamic or _CATEGORY.get(name) in dynamic


class GSDWriter:
    """Frames appended to a GSD 2.0 file, in the manner of GSDDumpWriter (GSDDumpWriter.cc:
    449-700): configuration/step every frame; configuration/dimensions, box, particles/N and
    types on frame 0 and when dynamic; a per-particle chunk only when some value differs from
    the schema default (or differed on frame 0), on frame 0 and when dynamic.

    Chunk data is appended at the end of the file; the index (and the namelist) live in
    allocated blocks that move to the end of the file, twice as large, when they fill."""

    def __init__(self, filename, mode="ab", dynamic=("property",), application="hoomd_amd"):
        if mode not in ("wb", "xb", "ab"):
            raise ValueError("mode must be 'wb', 'xb' or 'ab'")
        self.filename = filename
        self.dynamic = set(dynamic) | {"configuration/box", "particles/N",
                                       "particles/position"}
        self._nondefault = {}
        import os
        if mode == "ab" and os.path.exists(filename) and os.path.getsize(filename) > 0:
            g = GSDFile(filename)
            self._names = [n for n in g._names if n]
            self._entries = [(f, N, loc, M, self._names.index(name), typ)
                             for (f, name), (N, M, loc, typ) in
                             sorted(g._index.items(), key=lambda kv: (kv[0][0], kv[1][2]))]
            self.nframes = g.nframes
            (_, self._index_loc, self._index_alloc, self._names_loc, self._names_alloc,
             *_rest) = HEADER.unpack_from(g._data, 0)
            for name in g.chunk_names(0):
                self._nondefault[name] = True
            self._f = open(filename, "r+b")
        else:
            if mode == "xb" and os.path.exists(filename):
                raise FileExistsError(filename)
            self._names, self._entries, self.nframes = [], [], 0
            self._index_alloc, self._names_alloc = 128, 16
            self._index_loc = HEADER.size
            self._names_loc = self._index_loc + self._index_alloc * ENTRY.size
            self._f = open(filename, "w+b")
            self._f.write(b"\0" * (self._names_loc + self._names_alloc * 64))
            self._application = application
            self._write_header()
            self._write_names()
        self._application = application

    def _write_header(self):
        self._f.seek(0)
        self._f.write(HEADER.pack(MAGIC, self._index_loc, self._index_alloc, self._names_loc,
                                  self._names_alloc, SCHEMA_VERSION, GSD_VERSION,
                                  self._application.encode()[:63], b"hoomd", b""))

    def _write_names(self):
        blob = b"".join(n.encode() + b"\0" for n in self._names)
        if len(blob) > self._names_alloc * 64:
            while len(blob) > self._names_alloc * 64:
                self._names_alloc *= 2
            self._f.seek(0, 2)
            self._names_loc = self._f.tell()
            self._f.write(b"\0" * (self._names_alloc * 64))
            self._write_header()
        self._f.seek(self._names_loc)
        self._f.write(blob + b"\0" * (self._names_alloc * 64 - len(blob)))

    def _write_index(self):
        # gsd_find_chunk binary-searches the index by (frame, name id), so the entries are kept
        # in that order (a name first written in a later frame, or an append to a file whose
        # names were numbered by another writer, would otherwise leave it unsorted)
        self._entries.sort(key=lambda e: (e[0], e[4]))
        if len(self._entries) > self._index_alloc:
            while len(self._entries) > self._index_alloc:
                self._index_alloc *= 2
            self._f.seek(0, 2)
            self._index_loc = self._f.tell()
            self._f.write(b"\0" * (self._index_alloc * ENTRY.size))
            self._write_header()
        self._f.seek(self._index_loc)
        self._f.write(b"".join(ENTRY.pack(f, N, loc, M, nid, typ, 0)
                               for f, N, loc, M, nid, typ in self._entries))

    def _chunk(self, name, a):
        a = np.ascontiguousarray(a)
        if name not in self._names:
            self._names.append(name)
            self._write_names()
        # type names are written as uint8 rows, as GSDDumpWriter does (GSDDumpWriter.cc:436-442)
        typ = 1 if name.endswith("/types") else _CODES[a.dtype]
        N = a.shape[0] if a.ndim else 1
        M = int(np.prod(a.shape[1:])) if a.ndim > 1 else 1
        self._f.seek(0, 2)
        loc = self._f.tell()
        self._f.write(a.tobytes())
        self._entries.append((self.nframes, N, loc, M, self._names.index(name), typ))

    def write_frame(self, snapshot, step):
        first = self.nframes == 0
        p = snapshot.particles

        def want(name):
            return first or _is_dynamic(name, self.dynamic)

        self._chunk("configuration/step", np.array([step], np.uint64))
        if first:
            self._chunk("configuration/dimensions",
                        np.array([snapshot.configuration.dimensions], np.uint8))
        if want("configuration/box"):
            self._chunk("configuration/box", np.asarray(snapshot.configuration.box, np.float32))
        if want("particles/N"):
            self._chunk("particles/N", np.array([p.N], np.uint32))
        if want("particles/types"):
            self._chunk("particles/types", _type_chunk(p.types))
        for name, a, default in (("particles/typeid", np.asarray(p.typeid, np.uint32), 0),
                                 ("particles/mass", np.asarray(p.mass, np.float32), 1),
                                 ("particles/position", np.asarray(p.position, np.float32), 0),
                                 ("particles/velocity", np.asarray(p.velocity, np.float32), 0),
                                 ("particles/image", np.asarray(p.image, np.int32), 0)):
            if not want(name):
                continue
            nondefault = bool((a != default).any())
            if first:
                self._nondefault[name] = nondefault
            if nondefault or self._nondefault.get(name, False):
                self._chunk(name, a)
        b = snapshot.bonds
        if want("bonds/N") and (b.N or not first):
            self._chunk("bonds/N", np.array([b.N], np.uint32))
            self._chunk("bonds/types", _type_chunk(b.types or ["A"]))
            if b.N:
                self._chunk("bonds/typeid", np.asarray(b.typeid, np.uint32))
                self._chunk("bonds/group", np.asarray(b.group, np.uint32))
        self._write_index()
        self.nframes += 1
        self._f.flush()

    def close(self):
        if self._f is not None:
            self._f.close()
            self._f = None

    def __del__(self):
        self.close()


def write_frames(filename, frames, application="hoomd_amd", mode="wb", dynamic=("property",)):
    """Write (snapshot, step) frames to a GSD 2.0 file ('wb' truncates, 'ab' appends)."""
    w = GSDWriter(filename, mode, dynamic, application)
    for s, step in frames:
        w.write_frame(s, step)
    w.close()
