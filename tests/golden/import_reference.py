"""Regenerates the bundled H2 and GRI mechanisms from the gpuFoam test
fixtures (run once, in a checkout that has the reference sources).

Source: gpu_chemistry/unittest/testHelpers/{h2,gri}_{thermos,reactions}.h,
which embed OpenFOAM dictionaries as C++ raw strings.  The dictionaries are
parsed by this package's mechanism compiler and written back in canonical
OpenFOAM form (round-trip exact numbers); the script checks that the
re-emitted text compiles to bit-identical tables.

    python tests/golden/import_reference.py [/path/to/gpuFoam]
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))

from gpufoam_amd.mechanism import MECH_DIR, Mechanism


def _raw_string(path: str) -> str:
    s = open(path).read()
    a = s.index('R"(') + 3
    b = s.rindex(')"')
    return s[a:b]


def main(ref: str = "/root/reference") -> None:
    helpers = os.path.join(ref, "gpu_chemistry", "unittest", "testHelpers")
    for name in ("h2", "gri"):
        thermo = _raw_string(os.path.join(helpers, f"{name}_thermos.h"))
        reactions = _raw_string(os.path.join(helpers, f"{name}_reactions.h"))
        m = Mechanism(thermo, reactions, name)
        t2, r2 = m.to_foam()
        m2 = Mechanism(t2, r2, name)
        for k, v in m.arrays.items():
            if not np.array_equal(v, m2.arrays[k]):
                raise SystemExit(f"{name}: round trip changed '{k}'")
        assert m.equations == m2.equations and m.species == m2.species
        out = os.path.join(MECH_DIR, name)
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "thermo"), "w") as f:
            f.write(t2)
        with open(os.path.join(out, "reactions"), "w") as f:
            f.write(r2)
        print(f"{name}: {m.n_species} species, {m.n_reactions} reactions -> {out}")


if __name__ == "__main__":
    main(*sys.argv[1:])
