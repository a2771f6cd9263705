// Mechanism compiler: parses OpenFOAM thermo and reaction dictionaries and
// builds the structure-of-arrays tables the kernels read.
//
// Follows the host converters of gpuFoam:
//   makeGpuThermo            gpu_chemistry/src/gpuThermo/makeGpuThermo.H:16-66
//   make_reaction_rate       gpu_chemistry/src/gpuReaction/makeGpuReactionRate.H:9-135
//   make_specie_coeffs       gpu_chemistry/src/gpuReaction/makeGpuReactions.H:61-88
//   set_reaction_thermo      gpu_chemistry/src/gpuReaction/makeGpuReactions.H:18-59
//   makeGpuReactions         gpu_chemistry/src/gpuReaction/makeGpuReactions.H:90-142
// The reaction strings follow OpenFOAM's specieCoeffs syntax
// ("2H + O2 = H2O^1.5 + ..."; the exponent defaults to the stoichiometric
// coefficient; a third body "M" that is not a specie is dropped).
#include "mechanism.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <map>

#include "constants.hpp"
#include "foam_dict.hpp"

namespace gfc {

namespace {

thread_local std::string tLastError;

// NASA-7 thermo of one specie or of a reaction composite, with the
// perfect-gas "Y" and molecular weight (gpuThermo / gpuPerfectGas data).
struct Thermo {
    double Y = 0, W = 0, Tlow = 0, Thigh = 0, Tcommon = 0;
    double high[7] = {0}, low[7] = {0};
};

// s * thermo: scales the mass fraction only (gpuThermo.H:428-436,
// gpuPerfectGas.H:300-302).
Thermo scaled(double s, const Thermo& t) {
    Thermo r = t;
    r.Y = s * t.Y;
    return r;
}

// a += b (gpuThermo.H:354-378 with gpuPerfectGas.H:103-112).
void addInPlace(Thermo& a, const Thermo& b) {
    double Y1 = a.Y;
    const double sumY = a.Y + b.Y;
    if (std::abs(sumY) > kSmall) a.W = sumY / (a.Y / a.W + b.Y / b.W);
    a.Y = sumY;
    if (std::abs(a.Y) > kSmall) {
        Y1 /= a.Y;
        const double Y2 = b.Y / a.Y;
        a.Tlow = std::max(a.Tlow, b.Tlow);
        a.Thigh = std::min(a.Thigh, b.Thigh);
        for (int c = 0; c < 7; ++c) {
            a.high[c] = Y1 * a.high[c] + Y2 * b.high[c];
            a.low[c] = Y1 * a.low[c] + Y2 * b.low[c];
        }
    }
}

// lhs == rhs: the reaction thermo whose Gibbs energy is that of the
// products minus that of the reactants (gpuThermo.H:442-470 with
// gpuPerfectGas.H:312-325).
Thermo difference(const Thermo& lhs, const Thermo& rhs) {
    Thermo r;
    double diffY = rhs.Y - lhs.Y;
    if (std::abs(diffY) < kSmall) diffY = kSmall;
    const double diffRW = rhs.Y / rhs.W - lhs.Y / lhs.W;
    double W = kGreat;
    if (std::abs(diffRW) > kSmall) W = diffY / diffRW;
    r.Y = diffY;
    r.W = W;
    const double Y1 = rhs.Y / r.Y;
    const double Y2 = lhs.Y / r.Y;
    for (int c = 0; c < 7; ++c) {
        r.high[c] = Y1 * rhs.high[c] - Y2 * lhs.high[c];
        r.low[c] = Y1 * rhs.low[c] - Y2 * lhs.low[c];
    }
    r.Tlow = std::max(rhs.Tlow, lhs.Tlow);
    r.Thigh = std::min(rhs.Thigh, lhs.Thigh);
    r.Tcommon = rhs.Tcommon;
    return r;
}

// gpuSpecieExponent.H:21-36: a value is "integer" unless it exceeds its
// truncation.
bool isIntegerExponent(double N) {
    const int X = static_cast<int>(N);
    const double d = N - X;
    return !(d > 0);
}

struct SpecieCoeff {
    int index;
    double stoich;
    double exponent;
};

bool isNumberStart(char c) {
    return std::isdigit(static_cast<unsigned char>(c)) || c == '.';
}

// "2H + O2 = HO2 + H" -> lhs/rhs lists in written order (duplicates kept).
void parseEquation(const std::string& eq,
                   const std::map<std::string, int>& index,
                   std::vector<SpecieCoeff>& lhs,
                   std::vector<SpecieCoeff>& rhs) {
    std::vector<SpecieCoeff>* side = &lhs;
    bool sawEquals = false;
    double pending = -1; // a coefficient written apart from its specie
    size_t i = 0;
    const size_t n = eq.size();
    while (i < n) {
        const char c = eq[i];
        if (std::isspace(static_cast<unsigned char>(c)) || c == '+') {
            ++i;
            continue;
        }
        if (c == '=') {
            if (sawEquals)
                throw MechError{GFC_ERR_PARSE, "two '=' in reaction '" + eq + "'"};
            sawEquals = true;
            side = &rhs;
            ++i;
            continue;
        }
        size_t j = i;
        while (j < n && !std::isspace(static_cast<unsigned char>(eq[j])) &&
               eq[j] != '+' && eq[j] != '=')
            ++j;
        std::string term = eq.substr(i, j - i);
        i = j;
        double stoich = 1.0;
        size_t p = 0;
        while (p < term.size() && isNumberStart(term[p])) ++p;
        if (p > 0) {
            stoich = toScalar(term.substr(0, p));
            term = term.substr(p);
            if (term.empty()) {
                pending = stoich;
                continue;
            }
        } else if (pending > 0) {
            stoich = pending;
        }
        pending = -1;
        double exponent = stoich;
        const size_t caret = term.find('^');
        if (caret != std::string::npos) {
            exponent = toScalar(term.substr(caret + 1));
            term = term.substr(0, caret);
        }
        auto it = index.find(term);
        if (it == index.end()) {
            if (term == "M") continue; // third body marker
            throw MechError{GFC_ERR_PARSE, "unknown specie '" + term +
                                               "' in reaction '" + eq + "'"};
        }
        side->push_back(SpecieCoeff{it->second, stoich, exponent});
    }
    if (!sawEquals || lhs.empty() || rhs.empty())
        throw MechError{GFC_ERR_PARSE, "malformed reaction '" + eq + "'"};
}

void readArrhenius(const Dict& d, double* out) {
    out[0] = d.scalar("A");
    out[1] = d.scalar("beta");
    out[2] = d.scalar("Ta");
}

void readEfficiencies(const Dict& d, const std::map<std::string, int>& index,
                      int nS, double* eff) {
    if (nS > GFC_MAX_SPECIES)
        throw MechError{GFC_ERR_TOO_MANY_SPECIES,
                        "Only up to 300 species supported in "
                        "ThirdBodyEfficiencies."};
    if (d.found("coeffs")) {
        const auto coeffs = d.wordScalarList("coeffs");
        if (static_cast<int>(coeffs.size()) != nS)
            throw MechError{GFC_ERR_EFFICIENCIES,
                            "Number of efficiencies is not equal to the "
                            "number of species."};
        for (const auto& [name, value] : coeffs) {
            auto it = index.find(name);
            if (it == index.end())
                throw MechError{GFC_ERR_PARSE,
                                "unknown specie '" + name + "' in coeffs"};
            eff[it->second] = value;
        }
    } else {
        const double dflt = d.scalar("defaultEfficiency");
        for (int i = 0; i < nS; ++i) eff[i] = dflt;
    }
}

} // namespace

void setLastError(const std::string& msg) { tLastError = msg; }

gfc_mechanism* compileMechanism(const std::string& thermoText,
                                const std::string& reactionsText) {
    Dict thermoDict, reactDict;
    try {
        thermoDict = parseDict(thermoText);
        reactDict = parseDict(reactionsText);
    } catch (const ParseError& e) {
        throw MechError{GFC_ERR_PARSE, e.what()};
    }

    auto m = std::make_unique<gfc_mechanism>();
    try {
        m->species = thermoDict.wordList("species");
    } catch (const ParseError& e) {
        throw MechError{GFC_ERR_PARSE, e.what()};
    }
    const int nS = static_cast<int>(m->species.size());
    if (nS == 0) throw MechError{GFC_ERR_PARSE, "empty species list"};
    if (nS > GFC_MAX_SPECIES)
        throw MechError{GFC_ERR_TOO_MANY_SPECIES,
                        "more than 300 species are not supported"};
    m->nS = nS;
    std::map<std::string, int> index;
    for (int i = 0; i < nS; ++i) {
        if (!index.emplace(m->species[i], i).second)
            throw MechError{GFC_ERR_PARSE,
                            "duplicate specie '" + m->species[i] + "'"};
    }

    // ---- species thermo (makeGpuThermo.H:16-48) ----
    m->W.resize(nS);
    m->sTlow.resize(nS);
    m->sThigh.resize(nS);
    m->sTcommon.resize(nS);
    m->high.resize(7 * nS);
    m->low.resize(7 * nS);
    m->rawHigh.resize(7 * nS);
    m->rawLow.resize(7 * nS);
    std::vector<Thermo> th(nS);
    try {
        for (int i = 0; i < nS; ++i) {
            const Dict& sd = thermoDict.requireDict(m->species[i]);
            const double W = sd.requireDict("specie").scalar("molWeight");
            const Dict& td = sd.requireDict("thermodynamics");
            const auto ch = td.scalarList("highCpCoeffs");
            const auto cl = td.scalarList("lowCpCoeffs");
            if (ch.size() != 7 || cl.size() != 7)
                throw MechError{GFC_ERR_PARSE,
                                "NASA-7 list of '" + m->species[i] +
                                    "' does not have 7 coefficients"};
            Thermo& t = th[i];
            t.Y = 1.0;
            t.W = W;
            t.Tlow = td.scalar("Tlow");
            t.Thigh = td.scalar("Thigh");
            t.Tcommon = td.scalar("Tcommon");
            const double R = kRR / W; // specie::R()
            for (int c = 0; c < 7; ++c) {
                t.high[c] = R * ch[c];
                t.low[c] = R * cl[c];
                m->rawHigh[7 * i + c] = ch[c];
                m->rawLow[7 * i + c] = cl[c];
                m->high[7 * i + c] = t.high[c];
                m->low[7 * i + c] = t.low[c];
            }
            m->W[i] = W;
            m->sTlow[i] = t.Tlow;
            m->sThigh[i] = t.Thigh;
            m->sTcommon[i] = t.Tcommon;
        }
    } catch (const ParseError& e) {
        throw MechError{GFC_ERR_PARSE, e.what()};
    }

    // ---- reactions (makeGpuReactions.H:90-142) ----
    const Dict* reactions = reactDict.subDict("reactions");
    if (!reactions) throw MechError{GFC_ERR_PARSE, "missing 'reactions'"};
    try {
        m->Tlow = reactDict.scalarOr("Tlow", 0.0);
        m->Thigh = reactDict.scalarOr("Thigh", DBL_MAX / 10);
    } catch (const ParseError& e) {
        throw MechError{GFC_ERR_PARSE, e.what()};
    }
    const int W_ = GFC_MAX_REACTION_WIDTH;
    for (const Entry& re : reactions->entries) {
        if (!re.dict)
            throw MechError{GFC_ERR_PARSE,
                            "reaction '" + re.key + "' is not a dictionary"};
        const Dict& rd = *re.dict;
        ++m->nR;
        m->reactionNames.push_back(re.key);
        try {
            std::string typeName = rd.word("type");
            const auto pos = typeName.find("Reaction");
            if (pos != std::string::npos) typeName = typeName.substr(0, pos);

            double k[3] = {0, 0, 0}, kInf[3] = {0, 0, 0},
                   troe[4] = {0, 0, 0, 0}, extra[GFC_EXTRA_PARAMS] = {};
            std::vector<double> eff(nS, 0.0);
            int type = -1, irreversible = 0;
            // gpuFoam's types (makeGpuReactionRate.H:115-134) and the
            // OpenFOAM types it rejects (makeMoreGpuChemistryModels.C:59-95)
            std::string base;
            if (typeName.rfind("irreversible", 0) == 0) {
                irreversible = 1;
                base = typeName.substr(12);
            } else if (typeName.rfind("reversible", 0) == 0) {
                base = typeName.substr(10);
            }
            auto pressureDependent = [&](int t, const std::string& fn) {
                type = t;
                readArrhenius(rd.requireDict("k0"), k);
                readArrhenius(rd.requireDict("kInf"), kInf);
                readEfficiencies(rd.requireDict("thirdBodyEfficiencies"),
                                 index, nS, eff.data());
                const Dict& F = rd.requireDict("F");
                if (fn == "Troe") {
                    troe[0] = F.scalar("alpha");
                    troe[1] = F.scalar("Tsss");
                    troe[2] = F.scalar("Ts");
                    troe[3] = F.scalar("Tss");
                } else if (fn == "SRI") {
                    extra[0] = F.scalar("a");
                    extra[1] = F.scalar("b");
                    extra[2] = F.scalar("c");
                    extra[3] = F.scalarOr("d", 1.0);
                    extra[4] = F.scalarOr("e", 0.0);
                }
            };
            if (base == "Arrhenius") {
                type = GFC_RATE_ARRHENIUS;
                readArrhenius(rd, k);
            } else if (base == "ThirdBodyArrhenius") {
                type = GFC_RATE_THIRD_BODY;
                readArrhenius(rd, k);
                readEfficiencies(rd, index, nS, eff.data());
            } else if (base == "ArrheniusLindemannFallOff") {
                pressureDependent(GFC_RATE_LINDEMANN, "Lindemann");
            } else if (base == "ArrheniusTroeFallOff") {
                pressureDependent(GFC_RATE_TROE, "Troe");
            } else if (base == "ArrheniusSRIFallOff") {
                pressureDependent(GFC_RATE_SRI, "SRI");
            } else if (base == "ArrheniusLindemannChemicallyActivated") {
                pressureDependent(GFC_RATE_LINDEMANN_CA, "Lindemann");
            } else if (base == "ArrheniusTroeChemicallyActivated") {
                pressureDependent(GFC_RATE_TROE_CA, "Troe");
            } else if (base == "ArrheniusSRIChemicallyActivated") {
                pressureDependent(GFC_RATE_SRI_CA, "SRI");
            } else if (base == "LandauTeller") {
                type = GFC_RATE_LANDAU_TELLER;
                readArrhenius(rd, k);
                extra[0] = rd.scalar("B");
                extra[1] = rd.scalar("C");
            } else if (base == "Janev") {
                type = GFC_RATE_JANEV;
                readArrhenius(rd, k);
                const auto b = rd.scalarList("b");
                if (b.size() != 9)
                    throw MechError{GFC_ERR_PARSE, "Janev rate needs 9 coefficients b"};
                for (int q = 0; q < 9; ++q) extra[q] = b[q];
            } else if (base == "powerSeries") {
                type = GFC_RATE_POWER_SERIES;
                readArrhenius(rd, k);
                const auto c = rd.scalarList("coeffs");
                if (c.size() != 4)
                    throw MechError{GFC_ERR_PARSE, "powerSeries rate needs 4 coeffs"};
                for (int q = 0; q < 4; ++q) extra[q] = c[q];
            } else {
                throw MechError{GFC_ERR_UNSUPPORTED,
                                "Invalid reaction type: " + rd.word("type")};
            }

            const std::string eq = rd.word("reaction");
            m->equations.push_back(eq);
            std::vector<SpecieCoeff> lhs, rhs;
            parseEquation(eq, index, lhs, rhs);
            if (static_cast<int>(lhs.size()) > W_ ||
                static_cast<int>(rhs.size()) > W_)
                throw MechError{GFC_ERR_TOO_LONG, "Too long reaction"};

            m->type.push_back(type);
            m->irreversible.push_back(irreversible);
            m->k.insert(m->k.end(), k, k + 3);
            m->kInf.insert(m->kInf.end(), kInf, kInf + 3);
            m->troe.insert(m->troe.end(), troe, troe + 4);
            m->extra.insert(m->extra.end(), extra, extra + GFC_EXTRA_PARAMS);
            m->eff.insert(m->eff.end(), eff.begin(), eff.end());
            m->nLhs.push_back(static_cast<int>(lhs.size()));
            m->nRhs.push_back(static_cast<int>(rhs.size()));
            auto pushSide = [&](const std::vector<SpecieCoeff>& s,
                                std::vector<int>& idx,
                                std::vector<double>& st,
                                std::vector<double>& ex,
                                std::vector<int>& exi) {
                for (int q = 0; q < W_; ++q) {
                    if (q < static_cast<int>(s.size())) {
                        idx.push_back(s[q].index);
                        st.push_back(s[q].stoich);
                        if (isIntegerExponent(s[q].exponent)) {
                            const int e = static_cast<int>(s[q].exponent);
                            exi.push_back(e);
                            ex.push_back(static_cast<double>(e));
                        } else {
                            exi.push_back(GFC_NO_INT_EXPONENT);
                            ex.push_back(s[q].exponent);
                        }
                    } else {
                        idx.push_back(0);
                        st.push_back(0.0);
                        ex.push_back(0.0);
                        exi.push_back(0);
                    }
                }
            };
            pushSide(lhs, m->lhsIndex, m->lhsStoich, m->lhsExp, m->lhsExpInt);
            pushSide(rhs, m->rhsIndex, m->rhsStoich, m->rhsExp, m->rhsExpInt);

            // Composite thermo (makeGpuReactions.H:18-59).
            auto sideThermo = [&](const std::vector<SpecieCoeff>& s) {
                Thermo t = scaled(s[0].stoich * th[s[0].index].W,
                                  th[s[0].index]);
                for (size_t q = 1; q < s.size(); ++q)
                    addInPlace(t, scaled(s[q].stoich * th[s[q].index].W,
                                         th[s[q].index]));
                return t;
            };
            const Thermo rhsT = sideThermo(rhs);
            const Thermo lhsT = sideThermo(lhs);
            if (std::abs(lhsT.Y - rhsT.Y) > 0.1)
                throw MechError{GFC_ERR_MASS_IMBALANCE,
                                "Mass imbalance for reaction '" + eq + "'"};
            const Thermo rt = difference(lhsT, rhsT);
            m->rY.push_back(rt.Y);
            m->rW.push_back(rt.W);
            m->rTcommon.push_back(rt.Tcommon);
            m->rhigh.insert(m->rhigh.end(), rt.high, rt.high + 7);
            m->rlow.insert(m->rlow.end(), rt.low, rt.low + 7);
        } catch (const ParseError& e) {
            throw MechError{GFC_ERR_PARSE,
                            "reaction '" + re.key + "': " + e.what()};
        }
    }
    if (m->nR == 0) throw MechError{GFC_ERR_PARSE, "no reactions"};
    return m.release();
}

} // namespace gfc

gfc_mech_view gfc_mechanism::view() const {
    gfc_mech_view v;
    std::memset(&v, 0, sizeof(v));
    v.n_species = nS;
    v.n_reactions = nR;
    v.Tlow = Tlow;
    v.Thigh = Thigh;
    v.W = W.data();
    v.sTlow = sTlow.data();
    v.sThigh = sThigh.data();
    v.sTcommon = sTcommon.data();
    v.high = high.data();
    v.low = low.data();
    v.type = type.data();
    v.irreversible = irreversible.data();
    v.k = k.data();
    v.kInf = kInf.data();
    v.troe = troe.data();
    v.eff = eff.data();
    v.n_lhs = nLhs.data();
    v.n_rhs = nRhs.data();
    v.lhs_index = lhsIndex.data();
    v.lhs_stoich = lhsStoich.data();
    v.lhs_exp = lhsExp.data();
    v.lhs_exp_int = lhsExpInt.data();
    v.rhs_index = rhsIndex.data();
    v.rhs_stoich = rhsStoich.data();
    v.rhs_exp = rhsExp.data();
    v.rhs_exp_int = rhsExpInt.data();
    v.rY = rY.data();
    v.rW = rW.data();
    v.rTcommon = rTcommon.data();
    v.rhigh = rhigh.data();
    v.rlow = rlow.data();
    v.extra = extra.data();
    return v;
}

// ---------------------------------------------------------------- C ABI --

extern "C" {

const char* gfc_last_error(void) { return gfc::tLastError.c_str(); }

int gfc_mechanism_from_foam(const char* thermo_text, const char* reactions_text,
                            gfc_mechanism** out) {
    if (!thermo_text || !reactions_text || !out) {
        gfc::setLastError("gfc_mechanism_from_foam: null argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    try {
        *out = gfc::compileMechanism(thermo_text, reactions_text);
        gfc::setLastError("");
        return GFC_OK;
    } catch (const gfc::MechError& e) {
        gfc::setLastError(e.message);
        return e.status;
    } catch (const std::exception& e) {
        gfc::setLastError(e.what());
        return GFC_ERR_PARSE;
    }
}

void gfc_mechanism_destroy(gfc_mechanism* mech) { delete mech; }

int gfc_mechanism_view(const gfc_mechanism* mech, gfc_mech_view* out) {
    if (!mech || !out) {
        gfc::setLastError("gfc_mechanism_view: null argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    *out = mech->view();
    return GFC_OK;
}

const char* gfc_mechanism_species_name(const gfc_mechanism* mech, int i) {
    if (!mech || i < 0 || i >= mech->nS) return nullptr;
    return mech->species[i].c_str();
}

const char* gfc_mechanism_reaction_equation(const gfc_mechanism* mech, int r) {
    if (!mech || r < 0 || r >= mech->nR) return nullptr;
    return mech->equations[r].c_str();
}

int gfc_mechanism_species_index(const gfc_mechanism* mech, const char* name) {
    if (!mech || !name) return -1;
    for (int i = 0; i < mech->nS; ++i)
        if (mech->species[i] == name) return i;
    return -1;
}

int gfc_mechanism_raw_thermo(const gfc_mechanism* mech, int i, double* high7,
                             double* low7) {
    if (!mech || i < 0 || i >= mech->nS || !high7 || !low7) {
        gfc::setLastError("gfc_mechanism_raw_thermo: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    for (int c = 0; c < 7; ++c) {
        high7[c] = mech->rawHigh[7 * i + c];
        low7[c] = mech->rawLow[7 * i + c];
    }
    return GFC_OK;
}

int gfc_solver_from_name(const char* name, int* solver) {
    if (!name || !solver) {
        gfc::setLastError("gfc_solver_from_name: null argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    const std::string n(name);
    if (n == "Rosenbrock12") *solver = GFC_ROSENBROCK12;
    else if (n == "Rosenbrock23") *solver = GFC_ROSENBROCK23;
    else if (n == "Rosenbrock34") *solver = GFC_ROSENBROCK34;
    else {
        gfc::setLastError("Invalid ODE Solver type: " + n);
        return GFC_ERR_SOLVER_NAME;
    }
    return GFC_OK;
}

void gfc_ode_inputs_default(int solver, gfc_ode_inputs* out) {
    if (!out) return;
    out->solver = solver;
    out->absTol = gfc::kSmall;
    out->relTol = 1e-4;
    out->maxSteps = 10000;
    out->safeScale = 0.9;
    out->alphaInc = 0.2;
    out->alphaDec = 0.25;
    out->minScale = 0.2;
    out->maxScale = 10;
}

int gfc_ode_inputs_from_foam(const char* text, gfc_ode_inputs* out,
                             double* Treact) {
    if (!text || !out) {
        gfc::setLastError("gfc_ode_inputs_from_foam: null argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    try {
        const gfc::Dict d = gfc::parseDict(text);
        const gfc::Dict& oc = d.requireDict("odeCoeffs");
        int solver = 0;
        const std::string name = oc.word("solver");
        int st = gfc_solver_from_name(name.c_str(), &solver);
        if (st != GFC_OK) return st;
        gfc_ode_inputs_default(solver, out);
        out->absTol = oc.scalarOr("absTol", out->absTol);
        out->relTol = oc.scalarOr("relTol", out->relTol);
        out->maxSteps = static_cast<int>(oc.scalarOr("maxSteps", out->maxSteps));
        out->safeScale = oc.scalarOr("safeScale", out->safeScale);
        out->alphaInc = oc.scalarOr("alphaIncrease", out->alphaInc);
        out->alphaDec = oc.scalarOr("alphaDecrease", out->alphaDec);
        out->minScale = oc.scalarOr("minScale", out->minScale);
        out->maxScale = oc.scalarOr("maxScale", out->maxScale);
        if (Treact) *Treact = d.scalarOr("Treact", 0.0);
        gfc::setLastError("");
        return GFC_OK;
    } catch (const std::exception& e) {
        gfc::setLastError(e.what());
        return GFC_ERR_PARSE;
    }
}

// OpenFOAM's Switch words (on/off, yes/no, true/false, y/n, t/f, none).
static int switch_value(const std::string& w) {
    for (const char* t : {"on", "yes", "true", "y", "t"})
        if (w == t) return 1;
    for (const char* f : {"off", "no", "false", "n", "f", "none"})
        if (w == f) return 0;
    return -1;
}

int gfc_chemistry_switches_from_foam(const char* text, int* chemistry,
                                     double* initialChemicalTimeStep,
                                     double* maxChemicalTimeStep) {
    if (!text) {
        gfc::setLastError("gfc_chemistry_switches_from_foam: null argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    try {
        const gfc::Dict d = gfc::parseDict(text);
        const int on = switch_value(d.word("chemistry"));
        if (on < 0) throw gfc::ParseError("chemistry: expected a switch (on/off)");
        if (chemistry) *chemistry = on;
        if (initialChemicalTimeStep)
            *initialChemicalTimeStep = d.scalar("initialChemicalTimeStep");
        if (maxChemicalTimeStep)
            *maxChemicalTimeStep = d.scalarOr("maxChemicalTimeStep", gfc::kGreat);
        gfc::setLastError("");
        return GFC_OK;
    } catch (const std::exception& e) {
        gfc::setLastError(e.what());
        return GFC_ERR_PARSE;
    }
}

} // extern "C"
