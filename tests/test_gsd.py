"""GSD files, Snapshot.replicate and triggers (host logic, no GPU).

The fixture tests/golden/spheres.gsd is the data file of the reference's how-to benchmark
(sphinx-doc/howto/spheres.gsd, written by HOOMD-blue 4.1.0): 2048 particles in a cubic box at
number density 1, with non-zero image flags.
"""

import os

import numpy as np
import pytest

from conftest import GOLDEN
from hoomd_amd import Snapshot, gsd
from hoomd_amd.trigger import On, Periodic, as_trigger

SPHERES = os.path.join(GOLDEN, "spheres.gsd")


def test_reads_reference_file():
    g = gsd.GSDFile(SPHERES)
    assert len(g) == 1
    assert g.application.startswith("HOOMD-blue")
    assert g.schema == "hoomd"
    assert g.chunk_names(0) == ["configuration/box", "configuration/dimensions",
                                "configuration/step", "particles/N", "particles/image",
                                "particles/position", "particles/types"]
    s = gsd.read_snapshot(SPHERES)
    L = s.configuration.box[0]
    assert s.configuration.box[:3] == [L, L, L] and s.configuration.box[3:] == [0, 0, 0]
    assert s.configuration.dimensions == 3
    assert s.particles.N == 2048 and s.particles.types == ["A"]
    assert abs(2048 / L**3 - 1.0) < 1e-5
    p = s.particles.position
    assert (p >= -L / 2).all() and (p < L / 2).all()
    # chunks the file omits take the schema defaults
    assert (s.particles.typeid == 0).all() and (s.particles.mass == 1).all()
    assert (s.particles.velocity == 0).all()
    assert np.abs(s.particles.image).max() > 0
    assert s._step == int(g.read_chunk(0, "configuration/step")[0])


def test_round_trip_is_exact(tmp_path):
    s = gsd.read_snapshot(SPHERES)
    rng = np.random.default_rng(3)
    s.particles.velocity[:] = rng.normal(size=(2048, 3)).astype(np.float32)
    fn = str(tmp_path / "a.gsd")
    gsd.write_frames(fn, [(s, 1234)])
    g = gsd.GSDFile(fn)
    # default-valued chunks are not written (GSDDumpWriter writes non-default data only)
    assert "particles/mass" not in g.chunk_names(0)
    assert "particles/typeid" not in g.chunk_names(0)
    assert "particles/velocity" in g.chunk_names(0)
    t = gsd.read_snapshot(fn)
    assert t._step == 1234
    np.testing.assert_array_equal(t.particles.position, s.particles.position)
    np.testing.assert_array_equal(t.particles.velocity, s.particles.velocity)
    np.testing.assert_array_equal(t.particles.image, s.particles.image)
    assert t.configuration.box == s.configuration.box


def test_append_and_dynamic_chunks(tmp_path):
    s = gsd.read_snapshot(SPHERES)
    fn = str(tmp_path / "b.gsd")
    w = gsd.GSDWriter(fn, "wb", dynamic=["property"])
    steps = []
    for k in range(70):  # 70 frames x 4 chunks grows the 128-entry index twice
        s.particles.position[0, 0] = -6 + 0.1 * k
        w.write_frame(s, 10 * k)
        steps.append(10 * k)
    w.close()
    w = gsd.GSDWriter(fn, "ab", dynamic=["property", "momentum"])
    s.particles.position[0, 0] = 5.5
    w.write_frame(s, 9999)
    w.close()
    g = gsd.GSDFile(fn)
    assert len(g) == 71
    assert g.chunk_names(1) == ["configuration/box", "configuration/step", "particles/N",
                                "particles/position"]
    assert "particles/image" in g.chunk_names(70)
    for k in (0, 1, 37, 69):
        t = gsd.read_snapshot(fn, k)
        assert t._step == steps[k]
        assert t.particles.position[0, 0] == np.float32(-6 + 0.1 * k)
        # images come from frame 0 where a frame omits them
        np.testing.assert_array_equal(t.particles.image, s.particles.image)
    assert gsd.read_snapshot(fn, -1)._step == 9999
    with pytest.raises(IndexError):
        gsd.read_snapshot(fn, 71)


def test_bonds_round_trip(tmp_path):
    s = Snapshot()
    s.configuration.box = [10, 10, 10, 0, 0, 0]
    s.particles.N = 4
    s.particles.types = ["A", "B"]
    s.particles.typeid[:] = [0, 1, 1, 0]
    s.particles.mass[:] = [1, 2, 2, 1]
    s.bonds.N = 3
    s.bonds.types = ["A-A"]
    s.bonds.group[:] = [[0, 1], [1, 2], [2, 3]]
    fn = str(tmp_path / "c.gsd")
    gsd.write_frames(fn, [(s, 0)])
    t = gsd.read_snapshot(fn)
    assert t.particles.types == ["A", "B"] and t.bonds.types == ["A-A"]
    np.testing.assert_array_equal(t.particles.typeid, s.particles.typeid)
    np.testing.assert_array_equal(t.particles.mass, s.particles.mass)
    np.testing.assert_array_equal(t.bonds.group, s.bonds.group)


def test_rejects_other_files(tmp_path):
    fn = tmp_path / "d.gsd"
    fn.write_bytes(b"\0" * 512)
    with pytest.raises(RuntimeError, match="magic"):
        gsd.GSDFile(str(fn))


def test_replicate_lattice_known_answer():
    """hoomd/pytest/test_state.py:263-282: one particle at the origin of a box of side 10,
    replicated 2x2x2, lands at (+-5, +-5, +-5) in this order."""
    s = Snapshot()
    s.configuration.box = [10, 10, 10, 0, 0, 0]
    s.particles.N = 1
    s.replicate(2, 2, 2)
    np.testing.assert_allclose(s.particles.position,
                               [[-5, -5, -5], [-5, -5, 5], [-5, 5, -5], [-5, 5, 5],
                                [5, -5, -5], [5, -5, 5], [5, 5, -5], [5, 5, 5]])
    assert s.configuration.box == [20, 20, 20, 0, 0, 0]


def test_replicate_unwraps_with_images():
    """Copy j = (l, m, n) of particle i (tag j*N + i) sits, unwrapped, at the unwrapped old
    position + (l, m, n) L + (L - L')/2, and inside the new box L'."""
    s = gsd.read_snapshot(SPHERES)
    L = np.array(s.configuration.box[:3])
    N = s.particles.N
    pu = s.particles.position + s.particles.image * L
    s.particles.velocity[:] = np.arange(3 * N).reshape(N, 3)
    r = Snapshot()
    r.configuration.box = list(s.configuration.box)
    r.particles.N = N
    for a in ("position", "image", "velocity", "typeid", "mass"):
        getattr(r.particles, a)[:] = getattr(s.particles, a)
    r.replicate(2, 3, 1)
    Ln = L * [2, 3, 1]
    assert r.particles.N == 6 * N
    q = r.particles.position
    assert (q >= -Ln / 2).all() and (q < Ln / 2).all()
    qu = q + r.particles.image * Ln
    lmn = np.array([[l, m, n] for l in range(2) for m in range(3) for n in range(1)])
    want = (pu[None] + (lmn * L)[:, None] + (L - Ln) / 2).reshape(-1, 3)
    np.testing.assert_allclose(qu, want, atol=1e-9)
    np.testing.assert_array_equal(r.particles.velocity, np.tile(s.particles.velocity, (6, 1)))
    # 1x1x1 is the identity up to rounding, images included
    t = gsd.read_snapshot(SPHERES)
    t.replicate(1, 1, 1)
    np.testing.assert_allclose(t.particles.position, s.particles.position, atol=1e-12)
    np.testing.assert_array_equal(t.particles.image, s.particles.image)


def test_replicate_bonds():
    s = Snapshot()
    s.configuration.box = [4, 4, 4, 0, 0, 0]
    s.particles.N = 2
    s.particles.position[:] = [[0, 0, 0], [1, 0, 0]]
    s.bonds.N = 1
    s.bonds.types = ["A-A"]
    s.bonds.group[:] = [[0, 1]]
    s.replicate(1, 2, 2)
    np.testing.assert_array_equal(s.bonds.group, [[0, 1], [2, 3], [4, 5], [6, 7]])


def test_periodic_trigger():
    for period, phase in ((1, 0), (7, 0), (7, 3), (5, 12)):
        tr = Periodic(period, phase)
        fires = [t for t in range(60) if tr(t)]
        # unsigned arithmetic (Trigger.h:81-84): steps before the phase fire too when
        # (t - phase) mod 2^64 is a multiple of the period
        assert fires == [t for t in range(60) if ((t - phase) % (1 << 64)) % period == 0]
        for t in range(50):
            assert tr.next_after(t) == min(f for f in range(t + 1, 200) if tr(f))
    assert On(10).next_after(3) == 10 and On(10).next_after(10) is None
    assert isinstance(as_trigger(100), Periodic) and as_trigger(100).period == 100


def test_grid_optimizer_closed_form():
    """GridOptimizer (solve.py:511-614) maximising -(x - 0.31)^2 on [0, 1]: 5 bins per round,
    two rounds, the bin centres measured in order; the final bin holds the maximum."""
    from hoomd_amd.tune import GridOptimizer, ManualTuneDefinition
    state = {"x": 0.5}
    seen = []
    tun = ManualTuneDefinition(get_y=lambda: -(state["x"] - 0.31) ** 2, target=0.0,
                               get_x=lambda: state["x"],
                               set_x=lambda v: state.__setitem__("x", v), domain=(0.0, 1.0))
    opt = GridOptimizer(n_bins=5, n_rounds=2)
    done = False
    for _ in range(20):
        done = opt.solve([tun])
        seen.append(state["x"])
        if done:
            break
    assert done
    np.testing.assert_allclose(seen[:5], [0.1, 0.3, 0.5, 0.7, 0.9])
    lo, hi = opt.final_bin(tun)
    assert lo <= 0.31 <= hi and abs((hi - lo) - 0.04) < 1e-12
    assert state["x"] == (lo + hi) / 2
    assert opt.solve([tun])  # stays solved
    # a y of None is skipped and reported as not solved
    opt2 = GridOptimizer()
    none = ManualTuneDefinition(get_y=lambda: None, target=0.0, get_x=lambda: 0.0,
                                set_x=lambda v: None, domain=(0.0, 1.0))
    assert not opt2.solve([none]) and not opt2.is_searching(none)
    # x is clamped into the domain
    tun.x = 5.0
    assert state["x"] == 1.0


def _index_entries(fn):
    with open(fn, "rb") as f:
        d = f.read()
    (_, index_loc, index_n, names_loc, names_n, *_r) = gsd.HEADER.unpack_from(d, 0)
    names = [n.decode() for n in d[names_loc:names_loc + names_n * 64].split(b"\0")]
    out = []
    for k in range(index_n):
        frame, N, loc, M, nid, typ, _ = gsd.ENTRY.unpack_from(d, index_loc + k * gsd.ENTRY.size)
        if loc == 0:
            break
        out.append((frame, nid, names[nid], typ))
    return out


def test_written_index_is_readable_by_the_gsd_library(tmp_path):
    """The index must be sorted by (frame, name id), the order gsd_find_chunk binary-searches
    (hoomd/extern/gsd.c:780-810), and every type code must be one the reference's gsd defines
    (1-10, hoomd/extern/gsd.h:24-51); type names are uint8 rows (GSDDumpWriter.cc:436-442).
    Cases: a chunk first written in frame 1 (velocity becomes non-default only later, with the
    momentum category dynamic) and an append to the reference-written spheres.gsd."""
    import shutil
    s = gsd.read_snapshot(SPHERES)
    fn = str(tmp_path / "b.gsd")
    w = gsd.GSDWriter(fn, "wb", dynamic=("property", "momentum"))
    w.write_frame(s, 0)
    s.particles.velocity[:] = 0.5
    w.write_frame(s, 10)
    w.close()
    ap = str(tmp_path / "spheres_append.gsd")
    shutil.copy(SPHERES, ap)
    gsd.write_frames(ap, [(s, 20), (s, 30)], mode="ab", dynamic=("property", "momentum"))
    for f in (fn, ap):
        ent = _index_entries(f)
        keys = [(fr, nid) for fr, nid, _, _ in ent]
        assert keys == sorted(keys), f
        assert all(1 <= typ <= 10 for _, _, _, typ in ent), f
        assert all(typ == 1 for _, _, name, typ in ent if name.endswith("/types")), f
        assert len(set(keys)) == len(keys)
    g = gsd.GSDFile(fn)
    assert "particles/velocity" not in g.chunk_names(0)
    assert "particles/velocity" in g.chunk_names(1)
    assert gsd.read_snapshot(fn, 1).particles.velocity[0, 0] == np.float32(0.5)
    g = gsd.GSDFile(ap)
    assert len(g) == 3
    t = gsd.read_snapshot(ap, 2)
    np.testing.assert_array_equal(t.particles.velocity, s.particles.velocity)
    assert t.particles.types == ["A"]


def test_gradient_descent_finds_the_maximum():
    """GradientDescent (solve.py:320-480) maximising -(x - 0.31)^2 + 2 from x = 0.1 with the
    momentum weights kappa: a forward-difference gradient and the damped step converge to the
    maximum within tol; the first call perturbs x by 10 %, and a constant y nudges x after
    more than three unchanged measurements."""
    from hoomd_amd.tune import GradientDescent, ManualTuneDefinition
    from hoomd_amd.variant import Ramp
    state = {"x": 0.1}
    tun = ManualTuneDefinition(get_y=lambda: -(state["x"] - 0.31) ** 2 + 2, target=0.0,
                               get_x=lambda: state["x"],
                               set_x=lambda v: state.__setitem__("x", v), domain=(0.0, 1.0))
    opt = GradientDescent(alpha=0.4, kappa=np.array([0.2, 0.1]), tol=1e-7, maximize=True)
    assert not opt.solve([tun])
    assert state["x"] == pytest.approx(0.11)
    for _ in range(500):
        if opt.solve([tun]):
            break
    assert opt.solve([tun])
    assert abs(state["x"] - 0.31) < 1e-3
    # minimising and a step cap
    state["x"] = 0.9
    opt = GradientDescent(alpha=0.5, tol=1e-8, maximize=False, max_delta=0.05)
    tun2 = ManualTuneDefinition(get_y=lambda: (state["x"] - 0.31) ** 2, target=0.0,
                                get_x=lambda: state["x"],
                                set_x=lambda v: state.__setitem__("x", v), domain=(0.0, 1.0))
    xs = []
    for _ in range(500):
        done = opt.solve([tun2])
        xs.append(state["x"])
        if done:
            break
    assert done and abs(state["x"] - 0.31) < 1e-3
    assert max(abs(a - b) for a, b in zip(xs[1:], xs[2:])) <= 0.05 + 1e-12
    # a y that does not change: x stays for three calls, then moves on by half its last change
    state["x"] = 0.5
    flat = ManualTuneDefinition(get_y=lambda: 1.0, target=0.0, get_x=lambda: state["x"],
                                set_x=lambda v: state.__setitem__("x", v), domain=(0.0, 1.0))
    opt = GradientDescent()
    opt.solve([flat])
    assert state["x"] == pytest.approx(0.55)
    for _ in range(3):
        opt.solve([flat])
        assert state["x"] == pytest.approx(0.55)
    opt.solve([flat])
    assert state["x"] == pytest.approx(0.575)
    # alpha as a variant of the number of solve calls
    opt = GradientDescent(alpha=Ramp(1e-5, 1e-6, 0, 30))
    assert opt.alpha == 1e-5
    opt._calls = 15
    assert opt.alpha == pytest.approx(5.5e-6)


def test_variants():
    from hoomd_amd.variant import Constant, Ramp, as_variant
    r = Ramp(0.0, 100.0, 1000, 10000)
    assert r(0) == 0.0 and r(1000) == 0.0 and r(6000) == 50.0 and r(11000) == 100.0
    assert r(10 ** 9) == 100.0 and r.min == 0.0 and r.max == 100.0
    assert as_variant(2)(5) == 2.0 and isinstance(as_variant(Constant(1)), Constant)
    with pytest.raises(TypeError):
        as_variant("x")
