"""CHEMKIN-II input (gfc_chemkin_to_foam / gfc_mechanism_from_chemkin, the
chemkinToFoam step of tutorials/h2/Allrun:10-12) and the SENKIN known
answers of the tutorials (CPU).

  * The mechanisms compiled from tutorials/{h2,gri}/chemkin/{chem.inp,
    therm.dat} equal those compiled from the reference's own H2 and GRI
    dictionaries (gpu_chemistry/unittest/testHelpers/{h2,gri}_*.h, which
    chemkinToFoam wrote from these inputs): species, thermo ranges and
    NASA-7 coefficients, reaction types, stoichiometry, exponents,
    efficiencies, Troe parameters and beta exactly; A and the molecular
    weights to 1 ulp (unit conversion, summation order); Ta to 5e-10 (the
    dictionaries print ten digits).
  * The oracle, integrating the tutorials' constant-pressure ignition
    cases with the compiled mechanisms, reproduces CHEMKIN-II SENKIN
    (senk.out): ignition time within 0.5 %, final temperature within 1 K,
    T(t) within 2 K outside the ignition front.
"""
import numpy as np
import pytest

from gpufoam_amd import GfcError, Mechanism, ode_inputs
from gpufoam_amd.mechanism import chemkin_to_foam
from oracle_lib import Oracle

import senkin_case as SC


@pytest.mark.parametrize("name", ["h2", "gri"])
def test_chemkin_tables_equal_reference_dictionaries(name):
    from gpufoam_amd import _capi
    m, _, _ = SC.load(name)
    ref = Mechanism.load(name)
    L = _capi.lib()
    assert m.species == ref.species and m.n_reactions == ref.n_reactions
    a, b = m.arrays, ref.arrays
    exact = ["sTlow", "sThigh", "sTcommon", "type", "irreversible", "n_lhs", "n_rhs",
             "lhs_index", "rhs_index", "lhs_stoich", "rhs_stoich", "lhs_exp", "rhs_exp",
             "lhs_exp_int", "rhs_exp_int", "eff", "troe"]
    for k in exact:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(m.raw_high, ref.raw_high) and np.array_equal(m.raw_low, ref.raw_low)
    assert np.allclose(a["W"], b["W"], rtol=1e-15, atol=0)
    for k in ("k", "kInf"):
        assert np.allclose(a[k][:, 0], b[k][:, 0], rtol=2.3e-16, atol=0), k
        assert np.array_equal(a[k][:, 1], b[k][:, 1]), k
        assert np.allclose(a[k][:, 2], b[k][:, 2], rtol=5e-10, atol=0), k
    for r in range(m.n_reactions):
        assert (L.gfc_mechanism_reaction_equation(m.handle, r) ==
                L.gfc_mechanism_reaction_equation(ref.handle, r))


@pytest.mark.parametrize("name", ["h2", "gri"])
def test_senkin_ignition_known_answer(name):
    m, y, rec = SC.load(name)
    nS = m.n_species
    ode = ode_inputs("Rosenbrock34", relTol=1e-6, absTol=1e-10)
    o = Oracle(m)
    dtc = np.array([1e-8])
    cs = SC.CASES[name]
    n = int(round(cs["end"] / cs["sample"]))
    ts, Ts = [0.0], [y[nS, 0]]
    for k in range(n):
        y, dtc, _, st, _ = o.solve_cells(ode, y, cs["sample"], dtc, threads=1)
        assert st[0][0] == 0
        ts.append((k + 1) * cs["sample"])
        Ts.append(y[nS, 0])
    dtig, dTend, dT = SC.compare(ts, Ts, rec)
    print(f"{name}: ignition {dtig:+.3%}, final T {dTend:+.2f} K, |dT| away from the front "
          f"{dT:.2f} K")
    assert abs(dtig) < 5e-3 and abs(dTend) < 1.0 and dT < 2.0


CHEM = """ELEMENTS
H O N
END
SPECIES
H2 O2 H2O N2
END
REACTIONS
{body}
END
"""


def _therm():
    with open(SC.os.path.join(SC.HERE, "golden", "chemkin", "h2", "therm.dat")) as f:
        return f.read()


def test_chemkin_reactions_map_to_openfoam_types():
    """+M, (+M) with LOW/TROE, (+N2), explicit units and an irreversible
    reaction, read back from the written dictionaries."""
    body = """2H2+O2=>2H2O   1.0E+10 0.5 1000.0
2H2+O2+M<=>2H2O+M   2.0E+15 0.0 0.0
   H2O/6.0/ N2/0.5/
2H2+O2(+M)=2H2O(+M)   3.0E+12 0.0 0.0
   LOW / 4.0E+18 -1.0 0.0 /
   TROE / 0.5 100.0 1000.0 /
2H2+O2(+N2)=2H2O(+N2)   5.0E+12 0.0 500.0
   LOW / 6.0E+17 0.0 0.0 /"""
    thermo, reactions = chemkin_to_foam(CHEM.format(body=body), _therm())
    m = Mechanism(thermo, reactions)
    a = m.arrays
    assert list(a["type"]) == [0, 1, 3, 2] and list(a["irreversible"]) == [1, 0, 0, 0]
    assert a["k"][0, 0] == 1.0e10 / 1e6                # third order in cm^3/mol
    assert a["k"][0, 2] == 1000.0 / 1.987316          # Ta = Ea / R (cal/mol)
    assert a["k"][1, 0] == 2.0e15 / 1e9                # +M adds one order
    assert list(a["eff"][1]) == [1.0, 1.0, 6.0, 0.5]
    assert a["k"][2, 0] == 4.0e18 / 1e9                # k0 = LOW, one order more
    assert a["kInf"][2, 0] == 3.0e12 / 1e6
    assert a["troe"][2, 3] == 1.0 / np.finfo(np.float64).eps  # T** absent
    assert list(a["eff"][3]) == [0.0, 0.0, 0.0, 1.0]    # (+N2)
    units = CHEM.format(body="2H2+O2=2H2O  1.0E+10 0.0 2.0").replace("REACTIONS",
                                                                  "REACTIONS KCAL/MOLE")
    t2, r2 = chemkin_to_foam(units, _therm())
    assert Mechanism(t2, r2).arrays["k"][0, 2] == 2.0 * 1000.0 / 1.987316


@pytest.mark.parametrize("body,status", [
    ("H2+O2=H2O+X  1.0 0.0 0.0", 2),                      # unknown specie
    ("H2+O2=H2O  1.0 0.0 0.0\n   PLOG / 1.0 1.0 0.0 0.0 /", 3),  # unsupported keyword
    ("H2+O2(+M)=H2O(+M)  1.0 0.0 0.0", 2),                # fall-off without LOW
    ("H2+O2=H2O  1.0 0.0", 2),                             # missing Ea
])
def test_chemkin_errors_are_status_codes(body, status):
    with pytest.raises(GfcError) as e:
        chemkin_to_foam(CHEM.format(body=body), _therm())
    assert e.value.status == status, str(e.value)


def test_chemkin_missing_thermo_record():
    chem = CHEM.format(body="H2+O2=H2O  1.0 0.0 0.0").replace("H2 O2 H2O N2", "H2 O2 H2O N2 XY")
    with pytest.raises(GfcError) as e:
        chemkin_to_foam(chem, _therm())
    assert e.value.status == 2 and "XY" in str(e.value)
