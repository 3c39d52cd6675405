"""Write the golden fixtures used by the parity tests.

Run once in a container that holds the HOOMD-blue sources (default /root/reference):

    python tests/golden/make_fixtures.py [REFERENCE_ROOT]

Every fixture is data only: the inputs and expected outputs that the reference's own tests
hold. Sources:
  * hoomd/md/pytest/forces_and_energies.json  -- the "LJ" entry (Mathematica values at r = 0.75
    and 1.5 for three parameter sets; compared at rtol 5e-6 by test_potential.py:994-1051)
  * hoomd/md/pytest/forces_and_energies.json  -- the "Table" entry (three tables at r = 0.75
    and 1.5, r_cut 2.5)
  * hoomd/md/pytest/true_pair_list.json        -- 1,029 pairs of the 7^3 simple-cubic lattice,
    a = 1, r_cut 1.1, buffer 0 (test_nlist.py:221-231,353-372)
  * hoomd/md/pytest/test_bond.py:39-58          -- FENEWCA force/energy vectors at d = 0.969
    (compared at rel 1e-2 by test_bond.py:134-155); transcribed numbers
  * hoomd/test/test_cell_list.cc:33-148         -- cell-grid dimension known answers
  * hoomd/test/random_numbers_test.cc:289-323,456-501 -- Seed/Counter packing, u01 endpoints
  * Random123 known-answer vectors for philox4x32-10 (published with Random123's kat_vectors;
    the library is an un-vendored submodule of the reference, .gitmodules:10-12)
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    pytest_dir = os.path.join(ref, "hoomd", "md", "pytest")

    with open(os.path.join(pytest_dir, "forces_and_energies.json")) as f:
        fe = json.load(f)
    with open(os.path.join(HERE, "lj_forces_and_energies.json"), "w") as f:
        json.dump({"source": "hoomd/md/pytest/forces_and_energies.json[LJ]",
                   "distances": [0.75, 1.5], "r_cut": 2.5, "mode": "none",
                   "LJ": fe["LJ"]}, f, indent=1)
    with open(os.path.join(HERE, "table_forces_and_energies.json"), "w") as f:
        json.dump({"source": "hoomd/md/pytest/forces_and_energies.json[Table]",
                   "distances": [0.75, 1.5], "r_cut": 2.5, "mode": "none",
                   "Table": fe["Table"]}, f, indent=1)

    with open(os.path.join(pytest_dir, "true_pair_list.json")) as f:
        pairs = json.load(f)
    with open(os.path.join(HERE, "true_pair_list.json"), "w") as f:
        json.dump({"source": "hoomd/md/pytest/true_pair_list.json",
                   "lattice": {"kind": "sc", "n": 7, "a": 1.0}, "r_cut": 1.1, "buffer": 0.0,
                   "pairs": pairs}, f)

    fene = {
        "source": "hoomd/md/pytest/test_bond.py:39-58,100-112",
        "d": 0.969, "L": 5.0,
        "cases": [
            {"params": dict(k=30.0, r0=1.6, epsilon=0.9, sigma=1.1, delta=-0.5),
             "force": 282.296, "energy": 70.5638},
            {"params": dict(k=25.0, r0=1.7, epsilon=1.0, sigma=1.0, delta=-0.5),
             "force": 146.288, "energy": 49.2476},
            {"params": dict(k=20.0, r0=1.8, epsilon=1.1, sigma=0.9, delta=-0.5),
             "force": 88.8238, "energy": 35.3135},
        ],
        "rtol": 1e-2,
    }
    with open(os.path.join(HERE, "fene_bond.json"), "w") as f:
        json.dump(fene, f, indent=1)

    cells = {
        "source": "hoomd/test/test_cell_list.cc:33-148",
        "cases": [
            {"L": [10.0, 10.0, 10.0], "width": 1.0, "dim": [10, 10, 10]},
            {"L": [5.5, 5.5, 5.5], "width": 1.0, "dim": [5, 5, 5]},
            {"L": [5.5, 5.5, 5.5], "width": 0.5, "dim": [11, 11, 11]},
            {"L": [5.5, 3.0, 10.5], "width": 1.0, "dim": [5, 3, 10]},
        ],
    }
    with open(os.path.join(HERE, "cell_dims.json"), "w") as f:
        json.dump(cells, f, indent=1)

    rng = {
        "source_hoomd": "hoomd/test/random_numbers_test.cc:289-323,456-501",
        "seed": {"id": 1, "timestep": "0xabcdef1234567890", "seed": "0x5eed",
                 "key": ["0x015eed12", "0x34567890"]},
        "counters": [
            {"args": [], "ctr": [0, 0, 0, 0]},
            {"args": ["0xabcdef12"], "ctr": [0, 0, 0, "0xabcdef12"]},
            {"args": ["0x1234", "0x5678"], "ctr": [0, 0, "0x5678", "0x1234"]},
            {"args": ["0xabcd", "0xef123", "0x4567"], "ctr": [0, "0x4567", "0xef123", "0xabcd"]},
            {"args": ["0xabcd", "0xef123", "0x4567", "0x1234"],
             "ctr": ["0x12340000", "0x4567", "0xef123", "0xabcd"]},
        ],
        "u01_double": {"smallest": 2.710505431213761e-20, "largest": 1.0},
        "source_philox": "Random123 kat_vectors, philox4x32 R=10",
        "philox4x32_10": [
            {"ctr": ["0x00000000", "0x00000000", "0x00000000", "0x00000000"],
             "key": ["0x00000000", "0x00000000"],
             "out": ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]},
            {"ctr": ["0xffffffff", "0xffffffff", "0xffffffff", "0xffffffff"],
             "key": ["0xffffffff", "0xffffffff"],
             "out": ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]},
            {"ctr": ["0x243f6a88", "0x85a308d3", "0x13198a2e", "0x03707344"],
             "key": ["0xa4093822", "0x299f31d0"],
             "out": ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]},
        ],
    }
    with open(os.path.join(HERE, "rng_known_answers.json"), "w") as f:
        json.dump(rng, f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
