"""CPU checks of the C ABI: the library loads and exports every function include/*.h declares,
and the ctypes binding covers all of them. No compute calls (no GPU here)."""

import glob
import os
import re

from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(hm_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from hoomd_amd import _lib
    L = _lib.lib()
    names = _declared()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    from hoomd_amd import _lib
    assert _declared() == set(_lib.SIGNATURES)


def test_version_and_host_helpers():
    import ctypes as C

    from hoomd_amd import _lib
    L = _lib.lib()
    assert b"gfx950" in L.hm_version()
    dim = (C.c_uint * 3)()
    b = _lib.box(107.494)
    assert L.hm_cell_dims(C.byref(b), 2.9, dim) == 0
    assert list(dim) == [37, 37, 37]
    assert L.hm_cell_n_adj(dim) == 27
    # bussi on the host thread matches the oracle's restatement bit for bit
    import oracle as O
    for t in (0, 7, 123456):
        assert L.hm_bussi_alpha_host(100.0, 299.0, 0.005, 0.0, 1.2, t, 9, 0) == \
            O.bussi_alpha(100.0, 299.0, 0.005, 0.0, 1.2, t, 9, 0)
