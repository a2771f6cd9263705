"""Mechanism compiler and chemistryProperties reader (CPU).

Mirrors the reference's host-side failure cases (makeGpuReactionRate.H,
makeGpuReactions.H, makeGpuOdeSolver.H) and the odeCoeffs defaults
(readGpuOdeInputs.H:12-22)."""
import ctypes as C

import numpy as np
import pytest

from gpufoam_amd import GfcError, Mechanism, _capi, ode_inputs, ode_inputs_from_foam

THERMO = """
species 3 ( H2 O2 H2O );
H2 { specie { molWeight 2.01594; }
     thermodynamics { Tlow 200; Thigh 3500; Tcommon 1000;
       highCpCoeffs ( 3.3372792 -4.94024731e-05 4.99456778e-07 -1.79566394e-10 2.00255376e-14 -950.158922 -3.20502331 );
       lowCpCoeffs ( 2.34433112 0.00798052075 -1.9478151e-05 2.01572094e-08 -7.37611761e-12 -917.935173 0.683010238 ); } }
O2 { specie { molWeight 31.9988; }
     thermodynamics { Tlow 200; Thigh 3500; Tcommon 1000;
       highCpCoeffs ( 3.28253784 0.00148308754 -7.57966669e-07 2.09470555e-10 -2.16717794e-14 -1088.45772 5.45323129 );
       lowCpCoeffs ( 3.78245636 -0.00299673416 9.84730201e-06 -9.68129509e-09 3.24372837e-12 -1063.94356 3.65767573 ); } }
H2O { specie { molWeight 18.01534; }
     thermodynamics { Tlow 200; Thigh 3500; Tcommon 1000;
       highCpCoeffs ( 3.03399249 0.00217691804 -1.64072518e-07 -9.7041987e-11 1.68200992e-14 -30004.2971 4.9667701 );
       lowCpCoeffs ( 4.19864056 -0.0020364341 6.52040211e-06 -5.48797062e-09 1.77197817e-12 -30293.7267 -0.849032208 ); } }
"""


def reactions(body: str, extra: str = "") -> str:
    return "reactions {\n" + body + "\n}\n" + extra


R_OK = reactions("""
r0 { type reversibleArrhenius; reaction "2H2 + O2 = 2H2O"; A 1e10; beta 0.5; Ta 1000; }
r1 { type irreversibleArrheniusReaction; reaction "H2O = H2O"; A 1; beta 0; Ta 0; }
r2 { type reversibleThirdBodyArrhenius; reaction "2H2O = 2H2 + O2"; A 2; beta 0; Ta 10;
     coeffs 3 ( (H2 2.5) (O2 1) (H2O 12) ); }
r3 { type reversibleArrheniusTroeFallOff; reaction "2H2 + O2 = 2H2O";
     k0 { A 1; beta 0; Ta 0; } kInf { A 2; beta 0; Ta 0; }
     F { alpha 0.5; Tsss 100; Ts 1000; Tss 5000; }
     thirdBodyEfficiencies { defaultEfficiency 1.5; } }
r4 { type reversibleArrheniusLindemannFallOff; reaction "O2 + 2H2^1.5 = 2H2O";
     k0 { A 1; beta 0; Ta 0; } kInf { A 2; beta 0; Ta 0; } F { }
     thirdBodyEfficiencies { defaultEfficiency 1; } }
""", "Tlow 250;\nThigh 4000;")


def test_compile_small_mechanism():
    m = Mechanism(THERMO, R_OK)
    a = m.arrays
    assert m.n_species == 3 and m.n_reactions == 5
    assert m.species == ["H2", "O2", "H2O"]
    assert list(a["type"]) == [0, 0, 1, 3, 2]
    assert list(a["irreversible"]) == [0, 1, 0, 0, 0]
    assert (m.Tlow, m.Thigh) == (250.0, 4000.0)
    # stoichiometry and integer/real exponents (make_specie_coeffs)
    assert list(a["lhs_index"][0, :2]) == [0, 1] and list(a["lhs_stoich"][0, :2]) == [2, 1]
    assert a["lhs_exp_int"][0, 0] == 2
    assert a["lhs_exp_int"][4, 1] == _capi.NO_INT_EXPONENT and a["lhs_exp"][4, 1] == 1.5
    # efficiencies: by name and by default
    assert list(a["eff"][2]) == [2.5, 1.0, 12.0]
    assert list(a["eff"][3]) == [1.5, 1.5, 1.5]
    assert list(a["troe"][3]) == [0.5, 100, 1000, 5000]
    # specie coefficients are multiplied by R = RR/W (makeGpuThermo.H:41-44)
    RR = 6.02214e26 * 1.38065e-23
    assert np.array_equal(a["high"][0], (RR / 2.01594) * m.raw_high[0])


def test_composite_thermo_is_products_minus_reactants():
    """Y*coeffs of the reaction thermo equals sum(nu W a) of products minus
    reactants (the Gibbs energy change used by Kc)."""
    m = Mechanism(THERMO, R_OK)
    a = m.arrays
    W = a["W"]
    for r in (0, 2):
        want = np.zeros(7)
        for side, sgn in (("lhs", -1), ("rhs", 1)):
            for q in range(int(a[f"n_{side}"][r])):
                i = a[f"{side}_index"][r, q]
                want += sgn * a[f"{side}_stoich"][r, q] * W[i] * a["low"][i]
        got = a["rY"][r] * a["rlow"][r]
        assert np.allclose(got, want, rtol=1e-9, atol=1e-6 * np.abs(want).max())


@pytest.mark.parametrize("body, status", [
    ('r { type reversibleSRIFallOff; reaction "H2 = H2"; }', 3),            # unsupported type
    ('r { type reversibleArrhenius; reaction "H2 = O2"; A 1; beta 0; Ta 0; }', 4),  # mass imbalance
    ('r { type reversibleArrhenius; reaction "H2 = XX"; A 1; beta 0; Ta 0; }', 2),  # unknown specie
    ('r { type reversibleThirdBodyArrhenius; reaction "H2O = H2O"; A 1; beta 0; Ta 0;'
     ' coeffs 2 ( (H2 1) (O2 1) ); }', 7),                                  # efficiency count
    ('r { type reversibleArrhenius; reaction "H2O = H2O"; beta 0; Ta 0; }', 2),     # missing A
])
def test_mechanism_errors(body, status):
    with pytest.raises(GfcError) as e:
        Mechanism(THERMO, reactions(body))
    assert e.value.status == status


def test_too_long_reaction():
    lhs = " + ".join(["H2"] * 11)
    rhs = " + ".join(["H2"] * 11)
    with pytest.raises(GfcError) as e:
        Mechanism(THERMO, reactions(
            f'r {{ type reversibleArrhenius; reaction "{lhs} = {rhs}"; A 1; beta 0; Ta 0; }}'))
    assert e.value.status == 5


def test_third_body_marker_dropped_and_defaults():
    m = Mechanism(THERMO, reactions(
        'r { type reversibleThirdBodyArrhenius; reaction "2H2O + M = 2H2 + O2 + M"; '
        'A 1; beta 0; Ta 0; defaultEfficiency 1; }'))
    assert m.arrays["n_lhs"][0] == 1 and m.arrays["n_rhs"][0] == 2
    assert m.Tlow == 0.0 and m.Thigh == np.finfo(float).max / 10


def test_bundled_mechanisms():
    h2 = Mechanism.load("h2")
    gri = Mechanism.load("gri")
    assert (h2.n_species, h2.n_reactions) == (10, 27)
    assert (gri.n_species, gri.n_reactions) == (53, 325)
    assert list(np.bincount(gri.arrays["type"])) == [284, 12, 3, 26]
    assert int(gri.arrays["irreversible"].sum()) == 16
    assert gri.species[0] == "CH4" and gri.species[47] == "N2"


def test_ode_inputs_defaults_and_names():
    o = ode_inputs("Rosenbrock23")
    assert o.solver == 23 and o.relTol == 1e-4 and o.maxSteps == 10000
    assert o.absTol == np.finfo(float).eps
    assert (o.safeScale, o.alphaInc, o.alphaDec, o.minScale, o.maxScale) == (0.9, 0.2, 0.25, 0.2, 10)
    with pytest.raises(GfcError) as e:
        ode_inputs("Rosenbrock45")
    assert e.value.status == 8


def test_chemistry_properties():
    txt = """FoamFile { format ascii; class dictionary; object chemistryProperties; }
    type gpuChemistryModel;
    chemistry on;
    initialChemicalTimeStep 1e-7;
    Treact 295;
    odeCoeffs { solver Rosenbrock34; absTol 1e-12; relTol 1e-1; maxSteps 500; alphaIncrease 0.3; }
    #include "reactions"
    """
    o, Treact = ode_inputs_from_foam(txt)
    assert o.solver == 34 and o.absTol == 1e-12 and o.relTol == 0.1
    assert o.maxSteps == 500 and o.alphaInc == 0.3 and o.alphaDec == 0.25
    assert Treact == 295.0
    o2, T2 = ode_inputs_from_foam("odeCoeffs { solver Rosenbrock12; }")
    assert o2.solver == 12 and T2 == 0.0


def test_library_exports_every_declared_symbol():
    """Every entry point declared in include/gpufoam_amd.h is exported."""
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "gpufoam_amd.h")).read()
    declared = set(re.findall(r"^[a-z][\w *]*?\b(gfc_[a-z_]+)\(", hdr, re.M))
    L = C.CDLL(_capi.LIB_PATH)
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert declared >= set(_capi.exported_symbols())
