"""Chemistry mechanisms: OpenFOAM dictionaries compiled by the C++ mechanism
compiler of libgpufoam_amd (the counterpart of gpuFoam's makeGpuThermos /
makeGpuReactions, gpu_chemistry/src/gpuThermo/makeGpuThermo.H:16-66 and
gpu_chemistry/src/gpuReaction/makeGpuReactions.H:90-142).

Bundled mechanisms live under ``gpufoam_amd/mechanisms/<name>/`` as the two
OpenFOAM files ``thermo`` (species list + NASA-7 sub-dictionaries) and
``reactions``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _capi

MECH_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mechanisms")

_BASE_NAMES = {
    _capi.RATE_ARRHENIUS: "Arrhenius",
    _capi.RATE_THIRD_BODY: "ThirdBodyArrhenius",
    _capi.RATE_LINDEMANN: "ArrheniusLindemannFallOff",
    _capi.RATE_TROE: "ArrheniusTroeFallOff",
    _capi.RATE_SRI: "ArrheniusSRIFallOff",
    _capi.RATE_LINDEMANN_CA: "ArrheniusLindemannChemicallyActivated",
    _capi.RATE_TROE_CA: "ArrheniusTroeChemicallyActivated",
    _capi.RATE_SRI_CA: "ArrheniusSRIChemicallyActivated",
    _capi.RATE_LANDAU_TELLER: "LandauTeller",
    _capi.RATE_JANEV: "Janev",
    _capi.RATE_POWER_SERIES: "powerSeries",
}
_TYPE_NAMES = {(t, irr): ("irreversible" if irr else "reversible") + b
               for t, b in _BASE_NAMES.items() for irr in (0, 1)}
_PRESSURE_DEPENDENT = (_capi.RATE_LINDEMANN, _capi.RATE_TROE, _capi.RATE_SRI,
                       _capi.RATE_LINDEMANN_CA, _capi.RATE_TROE_CA, _capi.RATE_SRI_CA)


def _arr(ptr, n, dtype=np.float64):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


class Mechanism:
    """A compiled mechanism (owns a gfc_mechanism handle)."""

    def __init__(self, thermo_text: str, reactions_text: str, name: str = ""):
        L = _capi.lib()
        h = C.c_void_p()
        _capi.check(L.gfc_mechanism_from_foam(thermo_text.encode(),
                                              reactions_text.encode(), C.byref(h)))
        self._h = h
        self.name = name
        self.thermo_text = thermo_text
        self.reactions_text = reactions_text
        v = _capi.MechView()
        _capi.check(L.gfc_mechanism_view(h, C.byref(v)))
        self._view = v
        nS, nR = v.n_species, v.n_reactions
        W = _capi.MAX_REACTION_WIDTH
        self.n_species, self.n_reactions = nS, nR
        self.Tlow, self.Thigh = v.Tlow, v.Thigh
        self.species = [L.gfc_mechanism_species_name(h, i).decode() for i in range(nS)]
        self.equations = [L.gfc_mechanism_reaction_equation(h, r).decode() for r in range(nR)]
        a = {}
        for f in ("W", "sTlow", "sThigh", "sTcommon", "rY", "rW", "rTcommon"):
            a[f] = _arr(getattr(v, f), nS if f[0] in "Ws" else nR)
        a["high"] = _arr(v.high, 7 * nS).reshape(nS, 7)
        a["low"] = _arr(v.low, 7 * nS).reshape(nS, 7)
        a["type"] = _arr(v.type, nR, np.int32)
        a["irreversible"] = _arr(v.irreversible, nR, np.int32)
        a["k"] = _arr(v.k, 3 * nR).reshape(nR, 3)
        a["kInf"] = _arr(v.kInf, 3 * nR).reshape(nR, 3)
        a["troe"] = _arr(v.troe, 4 * nR).reshape(nR, 4)
        a["extra"] = _arr(v.extra, _capi.EXTRA_PARAMS * nR).reshape(nR, _capi.EXTRA_PARAMS)
        a["eff"] = _arr(v.eff, nS * nR).reshape(nR, nS)
        a["n_lhs"] = _arr(v.n_lhs, nR, np.int32)
        a["n_rhs"] = _arr(v.n_rhs, nR, np.int32)
        for side in ("lhs", "rhs"):
            a[f"{side}_index"] = _arr(getattr(v, f"{side}_index"), W * nR, np.int32).reshape(nR, W)
            a[f"{side}_stoich"] = _arr(getattr(v, f"{side}_stoich"), W * nR).reshape(nR, W)
            a[f"{side}_exp"] = _arr(getattr(v, f"{side}_exp"), W * nR).reshape(nR, W)
            a[f"{side}_exp_int"] = _arr(getattr(v, f"{side}_exp_int"), W * nR, np.int32).reshape(nR, W)
        a["rhigh"] = _arr(v.rhigh, 7 * nR).reshape(nR, 7)
        a["rlow"] = _arr(v.rlow, 7 * nR).reshape(nR, 7)
        self.arrays = a
        raw_h = np.zeros((nS, 7))
        raw_l = np.zeros((nS, 7))
        hh = (C.c_double * 7)()
        ll = (C.c_double * 7)()
        for i in range(nS):
            _capi.check(L.gfc_mechanism_raw_thermo(h, i, hh, ll))
            raw_h[i] = np.frombuffer(hh, dtype=np.float64)
            raw_l[i] = np.frombuffer(ll, dtype=np.float64)
        self.raw_high, self.raw_low = raw_h, raw_l

    # ---------------------------------------------------------------- misc
    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def view(self) -> _capi.MechView:
        """The C view (pointers valid while this object lives)."""
        return self._view

    def index(self, name: str) -> int:
        return self.species.index(name)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().gfc_mechanism_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------- loading
    @classmethod
    def from_dir(cls, path: str, name: str = "") -> "Mechanism":
        with open(os.path.join(path, "thermo")) as f:
            t = f.read()
        with open(os.path.join(path, "reactions")) as f:
            r = f.read()
        return cls(t, r, name or os.path.basename(path))

    @classmethod
    def from_chemkin(cls, chem_inp: str, therm_dat: str, name: str = "") -> "Mechanism":
        """A mechanism from CHEMKIN-II texts (chem.inp, therm.dat), converted
        as chemkinToFoam does (gfc_chemkin_to_foam) and compiled."""
        thermo, reactions = chemkin_to_foam(chem_inp, therm_dat)
        return cls(thermo, reactions, name)

    @classmethod
    def load(cls, name: str) -> "Mechanism":
        """A bundled mechanism: 'h2' (10 species, 27 reactions), 'gri'
        (GRI-Mech 3.0, 53 species, 325 reactions), 'syn200' (the seeded
        synthetic 200-species / 1200-reaction mechanism, generated) or
        'syn100' (a 100-species / 600-reaction one of the same generator,
        seed 7)."""
        if name == "syn200":
            from .synthetic import syn_mechanism
            return syn_mechanism()
        if name == "syn100":
            from .synthetic import syn_texts
            t, r = syn_texts(seed=7, nS=100, nR=600)
            return cls(t, r, "syn100")
        return cls.from_dir(os.path.join(MECH_DIR, name), name)

    # ------------------------------------------------------------- writing
    def to_foam(self) -> tuple[str, str]:
        """Re-emit (thermo, reactions) as OpenFOAM dictionaries with
        round-trip exact numbers."""
        return foam_thermo_text(self.species, self.arrays["W"], self.arrays["sTlow"],
                                self.arrays["sThigh"], self.arrays["sTcommon"],
                                self.raw_high, self.raw_low), \
            foam_reactions_text(self)


def chemkin_to_foam(chem_inp: str, therm_dat: str) -> tuple[str, str]:
    """(thermo, reactions) OpenFOAM dictionary texts of a CHEMKIN-II
    mechanism (the chemkinToFoam conversion, gfc_chemkin_to_foam)."""
    L = _capi.lib()
    a, b = chem_inp.encode(), therm_dat.encode()
    nt, nr = C.c_size_t(), C.c_size_t()
    _capi.check(L.gfc_chemkin_to_foam(a, b, None, 0, C.byref(nt), None, 0, C.byref(nr)))
    bt, br = C.create_string_buffer(nt.value), C.create_string_buffer(nr.value)
    _capi.check(L.gfc_chemkin_to_foam(a, b, bt, nt.value, C.byref(nt), br, nr.value,
                                      C.byref(nr)))
    return bt.value.decode(), br.value.decode()


def _num(x: float) -> str:
    r = repr(float(x))
    return r[:-2] if r.endswith(".0") else r


def foam_thermo_text(species, W, Tlow, Thigh, Tcommon, high, low) -> str:
    out = [f"species {len(species)} ( {' '.join(species)} );", ""]
    for i, s in enumerate(species):
        out += [
            s, "{",
            "    specie", "    {", f"        molWeight       {_num(W[i])};", "    }",
            "    thermodynamics", "    {",
            f"        Tlow            {_num(Tlow[i])};",
            f"        Thigh           {_num(Thigh[i])};",
            f"        Tcommon         {_num(Tcommon[i])};",
            "        highCpCoeffs    ( " + " ".join(_num(x) for x in high[i]) + " );",
            "        lowCpCoeffs     ( " + " ".join(_num(x) for x in low[i]) + " );",
            "    }", "}", ""]
    return "\n".join(out)


def foam_reactions_text(m: "Mechanism", names=None) -> str:
    a = m.arrays
    out = ["reactions", "{"]

    def arr(prefix, k):
        return [f"{prefix}A               {_num(k[0])};",
                f"{prefix}beta            {_num(k[1])};",
                f"{prefix}Ta              {_num(k[2])};"]

    def effs(prefix, r):
        lines = [f"{prefix}coeffs", f"{m.n_species}", "("]
        lines += [f"({s} {_num(a['eff'][r, i])})" for i, s in enumerate(m.species)]
        lines += [")", ";"]
        return lines

    for r in range(m.n_reactions):
        t = int(a["type"][r])
        tn = _TYPE_NAMES[(t, int(a["irreversible"][r]))]
        nm = names[r] if names else f"un-named-reaction-{r}"
        out += [f"    {nm}", "    {", f"        type            {tn};",
                f"        reaction        \"{m.equations[r]}\";"]
        p = "        "
        x = a["extra"][r]
        if t not in _PRESSURE_DEPENDENT:
            out += arr(p, a["k"][r])
            if t == _capi.RATE_THIRD_BODY:
                out += effs(p, r)
            elif t == _capi.RATE_LANDAU_TELLER:
                out += [f"{p}B               {_num(x[0])};", f"{p}C               {_num(x[1])};"]
            elif t == _capi.RATE_JANEV:
                out += [f"{p}b               9 ( " + " ".join(_num(v) for v in x[:9]) + " );"]
            elif t == _capi.RATE_POWER_SERIES:
                out += [f"{p}coeffs          4 ( " + " ".join(_num(v) for v in x[:4]) + " );"]
        else:
            out += [p + "k0", p + "{"] + arr(p + "    ", a["k"][r]) + [p + "}"]
            out += [p + "kInf", p + "{"] + arr(p + "    ", a["kInf"][r]) + [p + "}"]
            out += [p + "F", p + "{"]
            if t in (_capi.RATE_SRI, _capi.RATE_SRI_CA):
                out += [f"{p}    {k}               {_num(v)};" for k, v in zip("abcde", x[:5])]
            if t in (_capi.RATE_TROE, _capi.RATE_TROE_CA):
                tr = a["troe"][r]
                out += [f"{p}    alpha           {_num(tr[0])};",
                        f"{p}    Tsss            {_num(tr[1])};",
                        f"{p}    Ts              {_num(tr[2])};",
                        f"{p}    Tss             {_num(tr[3])};"]
            out += [p + "}", p + "thirdBodyEfficiencies", p + "{"] + effs(p + "    ", r) + [p + "}"]
        out += ["    }"]
    out += ["}", f"Tlow            {_num(m.Tlow)};", f"Thigh           {_num(m.Thigh)};", ""]
    return "\n".join(out)


def ode_inputs(solver: str = "Rosenbrock34", **kw) -> _capi.OdeInputs:
    """gfc_ode_inputs with the reference defaults (readGpuOdeInputs.H:12-22)
    overridden by keyword (absTol, relTol, maxSteps, safeScale, alphaInc,
    alphaDec, minScale, maxScale)."""
    L = _capi.lib()
    s = C.c_int()
    _capi.check(L.gfc_solver_from_name(solver.encode(), C.byref(s)))
    o = _capi.OdeInputs()
    L.gfc_ode_inputs_default(s.value, C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def ode_inputs_from_foam(chemistry_properties: str) -> tuple[_capi.OdeInputs, float]:
    """odeCoeffs and Treact of a chemistryProperties dictionary text."""
    o = _capi.OdeInputs()
    t = C.c_double()
    _capi.check(_capi.lib().gfc_ode_inputs_from_foam(chemistry_properties.encode(),
                                                     C.byref(o), C.byref(t)))
    return o, t.value
