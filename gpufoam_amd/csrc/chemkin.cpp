// CHEMKIN-II mechanism reader: chem.inp (ELEMENTS, SPECIES, REACTIONS) and
// a NASA-7 thermo database (therm.dat) -> the OpenFOAM thermo and
// reactions dictionaries that OpenFOAM's chemkinToFoam writes for gpuFoam's
// tutorials (tutorials/h2/Allrun:10-12: chemkinToFoam chem.inp therm.dat
// ... constant/reactions constant/speciesThermo), and from them a compiled
// mechanism.
//
// Conventions of that conversion, which the compiled tables depend on:
//   * species order of the SPECIES block; molecular weights from the
//     element counts of the thermo record and the atomic weights below
//     (OpenFOAM's table, e.g. H 1.00797, C 12.01115);
//   * NASA-7 coefficients as read, Tcommon from the record (columns
//     66-73) or the THERMO ALL default line;
//   * rate constants converted to kmol, m^3, s: A times (1e-3)^(order-1)
//     with the order the sum of the forward stoichiometric coefficients,
//     plus one for "+M" and for the low-pressure limit of a fall-off;
//     Ta = Ea / R with R = 1.987316 cal/(mol K) for CAL/MOLE (the value
//     behind every Ta of the reference's H2 and GRI dictionaries, which
//     chemkinToFoam wrote from these inputs), 1.987316e-3 for KCAL/MOLE,
//     RR/1000 J/(mol K) for JOULES/MOLE and RR/1e6 for KJOULES/MOLE;
//   * "+M" -> reversible/irreversibleThirdBodyArrhenius with efficiencies
//     1 unless listed; "(+M)" with LOW -> Lindemann fall-off, with TROE ->
//     Troe fall-off, with SRI -> SRI fall-off; "(+SP)" -> efficiency 1 for
//     SP and 0 for the others; HIGH (chemically activated) and REV (explicit
//     reverse rate) map to their own OpenFOAM types, and so do LT
//     (Landau-Teller), JAN (Janev) and FIT1 (power series);
//   * terms are written as in the input ("H + H + H2"), reactions named
//     un-named-reaction-<index>.
#include "mechanism.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "constants.hpp"

namespace gfc {

namespace {

struct ChemkinError {
    int status;
    std::string message;
};

[[noreturn]] void fail(const std::string& m, int status = GFC_ERR_PARSE) {
    throw ChemkinError{status, m};
}

std::string upper(std::string s) {
    for (char& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    return s;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

std::vector<std::string> words(const std::string& s) {
    std::istringstream in(s);
    std::vector<std::string> out;
    std::string w;
    while (in >> w) out.push_back(w);
    return out;
}

std::string noComment(const std::string& line) {
    const size_t p = line.find('!');
    return p == std::string::npos ? line : line.substr(0, p);
}

std::vector<std::string> lines(const std::string& text) {
    std::vector<std::string> out;
    std::istringstream in(text);
    std::string l;
    while (std::getline(in, l)) {
        if (!l.empty() && l.back() == '\r') l.pop_back();
        out.push_back(l);
    }
    return out;
}

double number(const std::string& s, const std::string& what) {
    std::string t = trim(s);
    for (char& c : t)
        if (c == 'D' || c == 'd') c = 'E'; // Fortran exponents
    if (t.empty()) fail("missing number in " + what);
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') fail("bad number '" + s + "' in " + what);
    return v;
}

// Atomic weights [kg/kmol] of OpenFOAM's element table for the elements of
// combustion mechanisms.
const std::map<std::string, double>& atomicWeights() {
    static const std::map<std::string, double> w = {
        {"H", 1.00797},  {"D", 2.01410},   {"HE", 4.0026},  {"C", 12.01115},
        {"N", 14.0067},  {"O", 15.9994},   {"F", 18.9984},  {"NE", 20.183},
        {"NA", 22.9898}, {"MG", 24.312},   {"AL", 26.9815}, {"SI", 28.086},
        {"P", 30.9738},  {"S", 32.064},    {"CL", 35.453},  {"AR", 39.948},
        {"K", 39.102},   {"CA", 40.08},    {"FE", 55.847},  {"KR", 83.80},
        {"XE", 131.30},  {"E", 5.45e-4}};
    return w;
}

struct SpecieThermo {
    double W = 0;
    double Tlow = 0, Thigh = 0, Tcommon = 0;
    double high[7] = {}, low[7] = {};
};

struct Term {
    double coeff;
    std::string specie;
    std::string text; // as written, coefficient included
};

struct Reaction {
    std::string equation;
    bool reversible = true;
    std::vector<Term> lhs, rhs;
    bool thirdBody = false;     // "+M"
    std::string fallOffBody;    // "" none, "M", or a specie
    double A = 0, beta = 0, Ea = 0;
    std::vector<double> low, high, troe, sri, rev, lt, jan, fit1;
    std::vector<std::pair<std::string, double>> eff;
    bool duplicate = false;
};

// Reads the thermo records of `names` from a CHEMKIN thermo database (the
// first record of a specie wins).
std::map<std::string, SpecieThermo> readThermo(const std::string& text,
                                               const std::set<std::string>& names) {
    std::map<std::string, SpecieThermo> out;
    const auto L = lines(text);
    size_t i = 0;
    double dTlow = 300, dTcommon = 1000, dThigh = 5000;
    // skip to THERMO
    while (i < L.size() && upper(trim(noComment(L[i]))).rfind("THERMO", 0) != 0) ++i;
    if (i == L.size()) fail("therm.dat: no THERMO line");
    ++i;
    // optional default temperature line
    while (i < L.size() && trim(noComment(L[i])).empty()) ++i;
    if (i < L.size()) {
        const auto w = words(noComment(L[i]));
        if (w.size() >= 3) {
            char* e1 = nullptr;
            std::strtod(w[0].c_str(), &e1);
            if (*e1 == '\0') {
                dTlow = number(w[0], "THERMO temperatures");
                dTcommon = number(w[1], "THERMO temperatures");
                dThigh = number(w[2], "THERMO temperatures");
                ++i;
            }
        }
    }
    const auto& aw = atomicWeights();
    while (i < L.size()) {
        const std::string& l1 = L[i];
        const std::string t1 = upper(trim(l1));
        if (t1.rfind("END", 0) == 0) break;
        if (t1.empty() || t1[0] == '!') {
            ++i;
            continue;
        }
        if (i + 3 >= L.size()) fail("therm.dat: truncated record at line " + std::to_string(i + 1));
        const auto nameWords = words(l1.substr(0, std::min<size_t>(18, l1.size())));
        if (nameWords.empty()) fail("therm.dat: record without a name at line " + std::to_string(i + 1));
        const std::string name = nameWords[0];
        auto field = [&](const std::string& l, size_t a, size_t n) {
            return a < l.size() ? l.substr(a, std::min(n, l.size() - a)) : std::string();
        };
        SpecieThermo th;
        // elements: four (symbol, count) pairs in columns 25-44, a fifth in 74-78
        double W = 0;
        auto element = [&](const std::string& sym, const std::string& cnt) {
            const std::string s = upper(trim(sym));
            if (s.empty() || s == "0") return;
            const std::string c = trim(cnt);
            const double n = c.empty() ? 0.0 : number(c, "element count of " + name);
            if (n == 0.0) return;
            auto it = aw.find(s);
            if (it == aw.end()) fail("therm.dat: unknown element '" + s + "' in " + name);
            W += n * it->second;
        };
        for (int q = 0; q < 4; ++q) element(field(l1, 24 + 5 * q, 2), field(l1, 26 + 5 * q, 3));
        element(field(l1, 73, 2), field(l1, 75, 3));
        th.W = W;
        const std::string sTl = trim(field(l1, 45, 10)), sTh = trim(field(l1, 55, 10)),
                          sTc = trim(field(l1, 65, 8));
        th.Tlow = sTl.empty() ? dTlow : number(sTl, "Tlow of " + name);
        th.Thigh = sTh.empty() ? dThigh : number(sTh, "Thigh of " + name);
        th.Tcommon = sTc.empty() ? dTcommon : number(sTc, "Tcommon of " + name);
        double a[15];
        for (int q = 0; q < 15; ++q) {
            const std::string& l = L[i + 1 + q / 5];
            a[q] = q < 14 ? number(field(l, 15 * (q % 5), 15), "coefficients of " + name) : 0.0;
        }
        for (int q = 0; q < 7; ++q) {
            th.high[q] = a[q];
            th.low[q] = a[7 + q];
        }
        if (names.count(name) && !out.count(name)) out[name] = th;
        i += 4;
    }
    return out;
}

// "2OH" -> (2, "OH"); species names may start with a digit only if the
// whole token is a specie.
Term parseTerm(const std::string& tok, const std::set<std::string>& species) {
    if (species.count(tok)) return Term{1.0, tok, tok};
    size_t p = 0;
    while (p < tok.size() &&
           (std::isdigit(static_cast<unsigned char>(tok[p])) || tok[p] == '.'))
        ++p;
    if (p == 0 || p == tok.size()) fail("unknown specie '" + tok + "'");
    const std::string sp = tok.substr(p);
    if (!species.count(sp)) fail("unknown specie '" + sp + "'");
    return Term{number(tok.substr(0, p), "coefficient"), sp, tok};
}

// Removes "(+X)" from one side of an equation; returns X or "".
std::string takeFallOff(std::string& side) {
    const size_t p = side.find("(+");
    if (p == std::string::npos) return "";
    const size_t q = side.find(')', p);
    if (q == std::string::npos) fail("unbalanced '(+' in '" + side + "'");
    const std::string body = trim(side.substr(p + 2, q - p - 2));
    side = side.substr(0, p) + side.substr(q + 1);
    return body;
}

std::vector<Term> parseSide(std::string side, const std::set<std::string>& species,
                            bool& thirdBody) {
    std::vector<Term> out;
    // split on '+' that is not the first character of a term
    std::string cur;
    std::vector<std::string> toks;
    for (size_t i = 0; i < side.size(); ++i) {
        const char c = side[i];
        if (std::isspace(static_cast<unsigned char>(c))) continue;
        if (c == '+' && !cur.empty()) {
            toks.push_back(cur);
            cur.clear();
            continue;
        }
        cur += c;
    }
    if (!cur.empty()) toks.push_back(cur);
    for (const auto& t : toks) {
        if (upper(t) == "M") {
            thirdBody = true;
            continue;
        }
        out.push_back(parseTerm(t, species));
    }
    if (out.empty()) fail("empty reaction side in '" + side + "'");
    return out;
}

// Slash-delimited auxiliary data: "LOW / a b c /", "H2O/6.4/", "DUP".
void parseAuxiliary(const std::string& line, Reaction& r,
                    const std::set<std::string>& species) {
    std::string s = line;
    size_t i = 0;
    while (true) {
        while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
        if (i >= s.size()) break;
        size_t j = i;
        while (j < s.size() && s[j] != '/' && !std::isspace(static_cast<unsigned char>(s[j])))
            ++j;
        std::string key = s.substr(i, j - i);
        // the next non-blank character decides whether values follow
        size_t k = j;
        while (k < s.size() && std::isspace(static_cast<unsigned char>(s[k]))) ++k;
        std::vector<double> vals;
        bool hasVals = k < s.size() && s[k] == '/';
        if (hasVals) {
            const size_t e = s.find('/', k + 1);
            if (e == std::string::npos) fail("unterminated '/' in '" + line + "'");
            for (const auto& w : words(s.substr(k + 1, e - k - 1)))
                vals.push_back(number(w, key));
            i = e + 1;
        } else {
            i = j;
        }
        const std::string K = upper(key);
        if (K == "DUP" || K == "DUPLICATE") {
            r.duplicate = true;
        } else if (K == "LOW") {
            r.low = vals;
        } else if (K == "HIGH") {
            r.high = vals;
        } else if (K == "TROE") {
            r.troe = vals;
        } else if (K == "SRI") {
            r.sri = vals;
        } else if (K == "REV") {
            r.rev = vals;
        } else if (K == "LT") {
            r.lt = vals;
        } else if (K == "JAN") {
            r.jan = vals;
        } else if (K == "FIT1") {
            r.fit1 = vals;
        } else if (hasVals && species.count(key) && vals.size() == 1) {
            r.eff.emplace_back(key, vals[0]);
        } else if (K == "PLOG" || K == "FORD" || K == "RORD" || K == "CHEB" ||
                   K == "RLT" || K == "TDEP" || K == "EXCI" ||
                   K == "MOME" || K == "XSMI" || K == "UNITS" || K == "HV") {
            fail("CHEMKIN keyword " + K + " is not supported (reaction '" + r.equation + "')",
                 GFC_ERR_UNSUPPORTED);
        } else {
            fail("unrecognised auxiliary data '" + key + "' (reaction '" + r.equation + "')");
        }
    }
}

std::string num(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

void writeArrhenius(std::ostringstream& o, const std::string& ind, double A, double beta,
                    double Ta) {
    o << ind << "A " << num(A) << ";\n" << ind << "beta " << num(beta) << ";\n"
      << ind << "Ta " << num(Ta) << ";\n";
}

struct Translation {
    std::string thermo, reactions;
};

Translation translate(const std::string& chem, const std::string& thermoDb) {
    const auto L = lines(chem);
    std::vector<std::string> species;
    std::set<std::string> spSet;
    std::vector<Reaction> rx;
    constexpr double kRcal = 1.987316; // cal/(mol K)
    double eScale = 1.0, eR = kRcal;   // Ta = Ea * eScale / eR
    std::string thermoInline;
    enum { None, Elem, Spec, Therm, React } block = None;
    for (size_t li = 0; li < L.size(); ++li) {
        const std::string raw = L[li];
        const std::string line = noComment(raw);
        const auto w = words(line);
        if (block == Therm) {
            thermoInline += raw + "\n";
            if (!w.empty() && upper(w[0]) == "END") block = None;
            continue;
        }
        if (w.empty()) continue;
        const std::string W0 = upper(w[0]);
        if (block == None) {
            if (W0.rfind("ELEM", 0) == 0) {
                block = Elem;
            } else if (W0.rfind("SPEC", 0) == 0) {
                block = Spec;
                for (size_t q = 1; q < w.size(); ++q) {
                    if (upper(w[q]) == "END") { block = None; break; }
                    if (spSet.insert(w[q]).second) species.push_back(w[q]);
                }
            } else if (W0.rfind("THER", 0) == 0) {
                block = Therm;
                thermoInline = raw + "\n";
            } else if (W0.rfind("REAC", 0) == 0) {
                block = React;
                for (size_t q = 1; q < w.size(); ++q) {
                    const std::string u = upper(w[q]);
                    if (u == "CAL/MOLE") { eScale = 1.0; eR = kRcal; }
                    else if (u == "KCAL/MOLE") { eScale = 1000.0; eR = kRcal; }
                    else if (u == "JOULES/MOLE") { eScale = 1000.0; eR = kRR; }
                    else if (u == "KJOULES/MOLE") { eScale = 1.0e6; eR = kRR; }
                    else if (u == "KELVINS") { eScale = 1.0; eR = 1.0; }
                    else if (u == "MOLES") continue;
                    else fail("REACTIONS units '" + w[q] + "' are not supported",
                              GFC_ERR_UNSUPPORTED);
                }
            } else {
                fail("unexpected line outside a block: '" + trim(raw) + "'");
            }
            continue;
        }
        if (W0 == "END") {
            block = None;
            continue;
        }
        if (block == Elem) continue;
        if (block == Spec) {
            for (const auto& s : w) {
                if (upper(s) == "END") { block = None; break; }
                if (spSet.insert(s).second) species.push_back(s);
            }
            continue;
        }
        // REACTIONS: a reaction line holds '=' and ends with A beta Ea
        if (line.find('=') != std::string::npos) {
            if (w.size() < 4) fail("reaction line without A, beta, Ea: '" + trim(raw) + "'");
            Reaction r;
            r.A = number(w[w.size() - 3], "A");
            r.beta = number(w[w.size() - 2], "beta");
            r.Ea = number(w[w.size() - 1], "Ea");
            std::string eq;
            for (size_t q = 0; q + 3 < w.size(); ++q) eq += w[q];
            r.equation = eq;
            std::string lhs, rhs;
            size_t p;
            if ((p = eq.find("<=>")) != std::string::npos) {
                lhs = eq.substr(0, p);
                rhs = eq.substr(p + 3);
            } else if ((p = eq.find("=>")) != std::string::npos) {
                r.reversible = false;
                lhs = eq.substr(0, p);
                rhs = eq.substr(p + 2);
            } else {
                p = eq.find('=');
                lhs = eq.substr(0, p);
                rhs = eq.substr(p + 1);
            }
            const std::string fl = takeFallOff(lhs), fr = takeFallOff(rhs);
            if (fl != fr) fail("fall-off bodies differ in '" + eq + "'");
            r.fallOffBody = fl;
            bool tbl = false, tbr = false;
            r.lhs = parseSide(lhs, spSet, tbl);
            r.rhs = parseSide(rhs, spSet, tbr);
            if (tbl != tbr) fail("third body on one side only in '" + eq + "'");
            r.thirdBody = tbl;
            if (r.thirdBody && !r.fallOffBody.empty())
                fail("both +M and (+M) in '" + eq + "'");
            rx.push_back(std::move(r));
            continue;
        }
        if (rx.empty()) fail("auxiliary data before the first reaction: '" + trim(raw) + "'");
        parseAuxiliary(line, rx.back(), spSet);
    }
    if (species.empty()) fail("no SPECIES block");

    const auto thermo = readThermo(thermoInline.empty() ? thermoDb : thermoInline + thermoDb,
                                   spSet);
    std::ostringstream T;
    T << "species " << species.size() << " (";
    for (const auto& s : species) T << " " << s;
    T << " );\n\n";
    for (const auto& s : species) {
        auto it = thermo.find(s);
        if (it == thermo.end()) fail("no thermo record for specie '" + s + "'");
        const SpecieThermo& t = it->second;
        T << s << "\n{\n    specie { molWeight " << num(t.W) << "; }\n"
          << "    thermodynamics\n    {\n        Tlow " << num(t.Tlow) << ";\n        Thigh "
          << num(t.Thigh) << ";\n        Tcommon " << num(t.Tcommon)
          << ";\n        highCpCoeffs (";
        for (double v : t.high) T << " " << num(v);
        T << " );\n        lowCpCoeffs (";
        for (double v : t.low) T << " " << num(v);
        T << " );\n    }\n}\n\n";
    }

    std::ostringstream R;
    R << "reactions\n{\n";
    for (size_t k = 0; k < rx.size(); ++k) {
        const Reaction& r = rx[k];
        double order = 0;
        for (const auto& t : r.lhs) order += t.coeff;
        // divide by the exact power of 1000 (multiplying by the inexact
        // 1e-3 would move A by an ulp)
        auto Aconv = [&](double A, double ord) { return A / std::pow(1e3, ord - 1.0); };
        auto Ta = [&](double Ea) { return Ea * eScale / eR; };
        auto side = [](const std::vector<Term>& s) {
            std::string o;
            for (size_t q = 0; q < s.size(); ++q) o += (q ? " + " : "") + s[q].text;
            return o;
        };
        const std::string eqText = side(r.lhs) + " = " + side(r.rhs);
        const std::string pre = r.reversible ? "reversible" : "irreversible";
        R << "    un-named-reaction-" << k << "\n    {\n";
        auto effs = [&](const std::string& ind, const std::string& body) {
            R << ind << "coeffs " << species.size() << " (";
            for (const auto& s : species) {
                double e = 1.0;
                if (!body.empty() && body != "M") e = (s == body) ? 1.0 : 0.0;
                for (const auto& pe : r.eff)
                    if (pe.first == s) e = pe.second;
                R << " (" << s << " " << num(e) << ")";
            }
            R << " );\n";
        };
        if (!r.rev.empty()) {
            if (r.rev.size() != 3) fail("REV needs three values in '" + r.equation + "'");
            if (!r.fallOffBody.empty() || r.thirdBody)
                fail("REV with a pressure-dependent rate is not supported ('" + r.equation + "')",
                     GFC_ERR_UNSUPPORTED);
            double rorder = 0;
            for (const auto& t : r.rhs) rorder += t.coeff;
            R << "        type nonEquilibriumReversibleArrhenius;\n"
              << "        reaction \"" << eqText << "\";\n        forward\n        {\n";
            writeArrhenius(R, "            ", Aconv(r.A, order), r.beta, Ta(r.Ea));
            R << "        }\n        reverse\n        {\n";
            writeArrhenius(R, "            ", Aconv(r.rev[0], rorder), r.rev[1], Ta(r.rev[2]));
            R << "        }\n    }\n";
            continue;
        }
        const int nT = !r.lt.empty() + !r.jan.empty() + !r.fit1.empty();
        if (nT > 0) {
            // temperature-dependent forms: LT (Landau-Teller), JAN (Janev),
            // FIT1 (power series)
            if (nT > 1 || r.thirdBody || !r.fallOffBody.empty())
                fail("LT/JAN/FIT1 combine with neither each other nor a third body ('" +
                     r.equation + "')", GFC_ERR_UNSUPPORTED);
            const char* type = !r.lt.empty() ? "LandauTeller" : !r.jan.empty() ? "Janev"
                                                                               : "powerSeries";
            const auto& v = !r.lt.empty() ? r.lt : !r.jan.empty() ? r.jan : r.fit1;
            const size_t want = !r.lt.empty() ? 2 : !r.jan.empty() ? 9 : 4;
            if (v.size() != want) fail(std::string(type) + " needs " + std::to_string(want) +
                                       " values in '" + r.equation + "'");
            R << "        type " << pre << type << ";\n"
              << "        reaction \"" << eqText << "\";\n";
            writeArrhenius(R, "        ", Aconv(r.A, order), r.beta, Ta(r.Ea));
            if (!r.lt.empty()) {
                R << "        B " << num(v[0]) << ";\n        C " << num(v[1]) << ";\n";
            } else {
                R << "        " << (r.jan.empty() ? "coeffs " : "b ") << want << " (";
                for (double x : v) R << " " << num(x);
                R << " );\n";
            }
            R << "    }\n";
            continue;
        }
        if (r.fallOffBody.empty()) {
            if (!r.low.empty() || !r.high.empty() || !r.troe.empty() || !r.sri.empty())
                fail("LOW/HIGH/TROE/SRI without (+M) in '" + r.equation + "'");
            if (r.thirdBody) {
                R << "        type " << pre << "ThirdBodyArrhenius;\n"
                  << "        reaction \"" << eqText << "\";\n";
                writeArrhenius(R, "        ", Aconv(r.A, order + 1), r.beta, Ta(r.Ea));
                effs("        ", "M");
            } else {
                R << "        type " << pre << "Arrhenius;\n"
                  << "        reaction \"" << eqText << "\";\n";
                writeArrhenius(R, "        ", Aconv(r.A, order), r.beta, Ta(r.Ea));
            }
            R << "    }\n";
            continue;
        }
        // pressure-dependent: fall-off (LOW) or chemically activated (HIGH)
        const bool fallOff = !r.low.empty();
        const bool activated = !r.high.empty();
        if (fallOff == activated)
            fail("(+M) reaction needs exactly one of LOW and HIGH: '" + r.equation + "'");
        const auto& other = fallOff ? r.low : r.high;
        if (other.size() != 3) fail("LOW/HIGH need three values in '" + r.equation + "'");
        std::string fn = "Lindemann";
        if (!r.troe.empty()) fn = "Troe";
        if (!r.sri.empty()) fn = "SRI";
        // the main line is the high-pressure limit of a fall-off and the
        // low-pressure limit of a chemically activated rate
        double k0[3], kInf[3];
        if (fallOff) {
            k0[0] = Aconv(other[0], order + 1); k0[1] = other[1]; k0[2] = Ta(other[2]);
            kInf[0] = Aconv(r.A, order); kInf[1] = r.beta; kInf[2] = Ta(r.Ea);
        } else {
            k0[0] = Aconv(r.A, order + 1); k0[1] = r.beta; k0[2] = Ta(r.Ea);
            kInf[0] = Aconv(other[0], order); kInf[1] = other[1]; kInf[2] = Ta(other[2]);
        }
        R << "        type " << pre << "Arrhenius" << fn
          << (fallOff ? "FallOff" : "ChemicallyActivated") << ";\n"
          << "        reaction \"" << eqText << "\";\n        k0\n        {\n";
        writeArrhenius(R, "            ", k0[0], k0[1], k0[2]);
        R << "        }\n        kInf\n        {\n";
        writeArrhenius(R, "            ", kInf[0], kInf[1], kInf[2]);
        R << "        }\n        F\n        {\n";
        if (fn == "Troe") {
            if (r.troe.size() != 3 && r.troe.size() != 4)
                fail("TROE needs three or four values in '" + r.equation + "'");
            R << "            alpha " << num(r.troe[0]) << ";\n            Tsss "
              << num(r.troe[1]) << ";\n            Ts " << num(r.troe[2])
              << ";\n            Tss " << num(r.troe.size() == 4 ? r.troe[3] : kGreat)
              << ";\n";
        } else if (fn == "SRI") {
            if (r.sri.size() != 3 && r.sri.size() != 5)
                fail("SRI needs three or five values in '" + r.equation + "'");
            R << "            a " << num(r.sri[0]) << ";\n            b " << num(r.sri[1])
              << ";\n            c " << num(r.sri[2]) << ";\n            d "
              << num(r.sri.size() == 5 ? r.sri[3] : 1.0) << ";\n            e "
              << num(r.sri.size() == 5 ? r.sri[4] : 0.0) << ";\n";
        }
        R << "        }\n        thirdBodyEfficiencies\n        {\n";
        effs("            ", r.fallOffBody);
        R << "        }\n    }\n";
    }
    R << "}\n";
    return Translation{T.str(), R.str()};
}

int copyOut(const std::string& s, char* buf, size_t cap, size_t* len) {
    if (len) *len = s.size() + 1;
    if (!buf) return GFC_OK;
    if (cap < s.size() + 1) {
        setLastError("gfc_chemkin_to_foam: buffer too small");
        return GFC_ERR_CAPACITY;
    }
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return GFC_OK;
}

} // namespace
} // namespace gfc

extern "C" {

int gfc_chemkin_to_foam(const char* chem_inp, const char* therm_dat, char* thermo_out,
                        size_t thermo_cap, size_t* thermo_len, char* reactions_out,
                        size_t reactions_cap, size_t* reactions_len) {
    if (!chem_inp || !therm_dat) {
        gfc::setLastError("gfc_chemkin_to_foam: null argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    try {
        const auto t = gfc::translate(chem_inp, therm_dat);
        int st = gfc::copyOut(t.thermo, thermo_out, thermo_cap, thermo_len);
        if (st != GFC_OK) return st;
        st = gfc::copyOut(t.reactions, reactions_out, reactions_cap, reactions_len);
        if (st != GFC_OK) return st;
        gfc::setLastError("");
        return GFC_OK;
    } catch (const gfc::ChemkinError& e) {
        gfc::setLastError(e.message);
        return e.status;
    } catch (const std::exception& e) {
        gfc::setLastError(e.what());
        return GFC_ERR_PARSE;
    }
}

int gfc_mechanism_from_chemkin(const char* chem_inp, const char* therm_dat,
                               gfc_mechanism** out) {
    if (!chem_inp || !therm_dat || !out) {
        gfc::setLastError("gfc_mechanism_from_chemkin: null argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    try {
        const auto t = gfc::translate(chem_inp, therm_dat);
        return gfc_mechanism_from_foam(t.thermo.c_str(), t.reactions.c_str(), out);
    } catch (const gfc::ChemkinError& e) {
        gfc::setLastError(e.message);
        return e.status;
    } catch (const std::exception& e) {
        gfc::setLastError(e.what());
        return GFC_ERR_PARSE;
    }
}

} // extern "C"
