"""gpufoam_amd: MI355X (gfx950) batched per-cell chemistry integrator with
the gpuFoam chemistry-model surface (solver names, odeCoeffs, Treact)."""
from ._capi import GfcError, lib  # noqa: F401
from .mechanism import Mechanism, ode_inputs, ode_inputs_from_foam  # noqa: F401
