"""The 16-lane-team kernels at their second size (N = 16) with padding rows:
H2 with four inert copies of N2 (14 species, n = 15 < N), so the team's last
member holds no component and the LU carries a padding row.  The four-cell
solve is checked against the oracle, and the team adjoint against the
one-lane adjoint, to the bars of test_gpu_parity.py and
test_gpu_adjoint_team.py."""
import os
import re

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import gpufoam_amd  # noqa: E402
from gpufoam_amd import ode_inputs  # noqa: E402
from gpufoam_amd.mechanism import Mechanism  # noqa: E402
from gpufoam_amd.workloads import make_states  # noqa: E402

from test_gpu_adjoint_team import _team_against_lane  # noqa: E402
from test_gpu_parity import _check, _device_run  # noqa: E402


def _h2_with_inert(k=4):
    """The bundled H2 mechanism plus k copies of N2 (named N2x0...), with
    third-body efficiency 1 in every efficiency list."""
    d = os.path.join(os.path.dirname(gpufoam_amd.__file__), "mechanisms", "h2")
    with open(os.path.join(d, "thermo")) as f:
        thermo = f.read()
    with open(os.path.join(d, "reactions")) as f:
        reactions = f.read()
    new = [f"N2x{i}" for i in range(k)]
    head = "species 10 ( H2 H O2 O OH HO2 H2O2 H2O AR N2 );"
    assert head in thermo
    thermo = thermo.replace(head, f"species {10 + k} ( H2 H O2 O OH HO2 H2O2 H2O AR N2 "
                                  f"{' '.join(new)} );")
    i = thermo.index("\nN2\n{") + 1
    depth, j = 0, thermo.index("{", i)
    while True:
        if thermo[j] == "{":
            depth += 1
        elif thermo[j] == "}":
            depth -= 1
            if depth == 0:
                break
        j += 1
    block = thermo[i:j + 1]
    thermo += "\n" + "\n".join(block.replace("N2\n{", f"{s}\n{{", 1) for s in new) + "\n"

    def coeffs(mt):
        return (f"coeffs\n{10 + k}\n(\n{mt.group(1)}\n" +
                "\n".join(f"({s} 1)" for s in new) + "\n)")
    reactions, nsub = re.subn(r"coeffs\n10\n\(\n(.*?)\n\)", coeffs, reactions, flags=re.S)
    assert nsub > 0
    return Mechanism(thermo, reactions, "h2x")


def _states(m, n):
    """h2_small states with 1 % of the mass in each added specie."""
    _, Y0, dt0, c = make_states("h2_small", n)
    nS = m.n_species
    Y = np.zeros((nS + 2, n))
    Y[:10] = Y0[:10] * (1.0 - 0.01 * (nS - 10))
    Y[10:nS] = 0.01
    Y[nS:] = Y0[10:]
    return Y, dt0, c


def test_padded_four_cell_solve_matches_oracle():
    from oracle_lib import Oracle
    m = _h2_with_inert()
    assert m.n_species == 14
    Y0, dt0, c = _states(m, 128)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    rho = np.linspace(0.5, 1.5, 128)
    ref = Oracle(m).solve_cells(ode, Y0, c["deltaT"], dt0, rho=rho, want_rr=True)
    print("case h2x (14 species) n=128 cells per wavefront 4", flush=True)
    got = _device_run(m, Y0, dt0, c, ode, rho, batched=True, cells_per_wave=4)
    _check(m, got, ref, tight_dt=False, Y0=Y0, rho=rho, deltaT=c["deltaT"])


def test_padded_team_adjoint_matches_lane_adjoint():
    m = _h2_with_inert()
    Y0, dt0, c = _states(m, 64)
    ode = ode_inputs(c["solver"], relTol=c["relTol"], absTol=c["absTol"])
    _team_against_lane(m, Y0, dt0, c, ode, 64, "h2x (14 species)")
