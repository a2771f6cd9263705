// Operation counts of the oracle (test infrastructure): oracle.c compiled as
// C++ with every double replaced by a counting scalar, so that each +, -, *,
// / and each transcendental call (exp, log, log10, pow, sqrt) of the
// restated reference arithmetic is counted.  Comparisons, fabs/fmax/fmin
// and sign changes are not counted.  Single-threaded.
#include <stddef.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace opc {
struct Counts {
    uint64_t add, mul, div, trans;
};
static Counts g{0, 0, 0, 0};

struct cd {
    double v;
    cd() = default;
    constexpr cd(double x) : v(x) {}
    explicit constexpr operator double() const { return v; }
    explicit operator long() const { return (long)v; }
    explicit operator int() const { return (int)v; }
    cd& operator+=(cd o) { ++g.add; v += o.v; return *this; }
    cd& operator-=(cd o) { ++g.add; v -= o.v; return *this; }
    cd& operator*=(cd o) { ++g.mul; v *= o.v; return *this; }
    cd& operator/=(cd o) { ++g.div; v /= o.v; return *this; }
    cd operator-() const { return cd(-v); }
    cd operator+() const { return *this; }
};
inline cd operator+(cd a, cd b) { ++g.add; return cd(a.v + b.v); }
inline cd operator-(cd a, cd b) { ++g.add; return cd(a.v - b.v); }
inline cd operator*(cd a, cd b) { ++g.mul; return cd(a.v * b.v); }
inline cd operator/(cd a, cd b) { ++g.div; return cd(a.v / b.v); }
#define OPC_CMP(op) \
    inline bool operator op(cd a, cd b) { return a.v op b.v; } \
    inline bool operator op(cd a, double b) { return a.v op b; } \
    inline bool operator op(double a, cd b) { return a op b.v; }
OPC_CMP(<) OPC_CMP(>) OPC_CMP(<=) OPC_CMP(>=) OPC_CMP(==) OPC_CMP(!=)
#undef OPC_CMP
inline cd operator+(cd a, double b) { return a + cd(b); }
inline cd operator+(double a, cd b) { return cd(a) + b; }
inline cd operator-(cd a, double b) { return a - cd(b); }
inline cd operator-(double a, cd b) { return cd(a) - b; }
inline cd operator*(cd a, double b) { return a * cd(b); }
inline cd operator*(double a, cd b) { return cd(a) * b; }
inline cd operator/(cd a, double b) { return a / cd(b); }
inline cd operator/(double a, cd b) { return cd(a) / b; }
inline cd exp(cd a) { ++g.trans; return cd(std::exp(a.v)); }
inline cd log(cd a) { ++g.trans; return cd(std::log(a.v)); }
inline cd log10(cd a) { ++g.trans; return cd(std::log10(a.v)); }
inline cd sqrt(cd a) { ++g.trans; return cd(std::sqrt(a.v)); }
inline cd pow(cd a, cd b) { ++g.trans; return cd(std::pow(a.v, b.v)); }
inline cd pow(cd a, double b) { return pow(a, cd(b)); }
inline cd pow(double a, cd b) { return pow(cd(a), b); }
inline cd fabs(cd a) { return cd(std::fabs(a.v)); }
inline cd fmax(cd a, cd b) { return cd(std::fmax(a.v, b.v)); }
inline cd fmin(cd a, cd b) { return cd(std::fmin(a.v, b.v)); }
inline cd fmax(cd a, double b) { return fmax(a, cd(b)); }
inline cd fmax(double a, cd b) { return fmax(cd(a), b); }
inline cd fmin(cd a, double b) { return fmin(a, cd(b)); }
inline cd fmin(double a, cd b) { return fmin(cd(a), b); }
} // namespace opc

using opc::cd;
using opc::exp;
using opc::fabs;
using opc::fmax;
using opc::fmin;
using opc::log;
using opc::log10;
using opc::pow;
using opc::sqrt;

#define double cd
#define orc_species_thermo opc_species_thermo
#define orc_reaction_Kc opc_reaction_Kc
#define orc_reaction_rates opc_reaction_rates
#define orc_reaction_parameters opc_reaction_parameters
#define orc_workspace_doubles opc_workspace_doubles
#define orc_derivatives opc_derivatives
#define orc_jacobian opc_jacobian
#define orc_lu_decompose opc_lu_decompose
#define orc_lu_backsubstitute opc_lu_backsubstitute
#define orc_rosenbrock_step opc_rosenbrock_step
#define orc_ode_solve opc_ode_solve
#define orc_solve_cells opc_solve_cells
#undef _OPENMP
#include "oracle.c"
#undef double

extern "C" {
// Counts of one call on state y (nS+2 values); out: add, mul, div, trans.
static void take(uint64_t* out) {
    out[0] = opc::g.add;
    out[1] = opc::g.mul;
    out[2] = opc::g.div;
    out[3] = opc::g.trans;
    opc::g = {0, 0, 0, 0};
}
void opc_count_derivatives(const gfc_mech_view* m, const double* y, uint64_t* out) {
    const int nE = m->n_species + 2;
    cd* yy = (cd*)y;  // same layout (one double member)
    cd* d = (cd*)calloc(nE, sizeof(cd));
    cd* w = (cd*)calloc(opc_workspace_doubles(m->n_species) + 16, sizeof(cd));
    opc::g = {0, 0, 0, 0};
    opc_derivatives(m, yy, d, w);
    take(out);
    free(d);
    free(w);
}
void opc_count_jacobian(const gfc_mech_view* m, const double* y, uint64_t* out) {
    const int nE = m->n_species + 2;
    cd* yy = (cd*)y;
    cd* d = (cd*)calloc(nE, sizeof(cd));
    cd* J = (cd*)calloc((size_t)nE * nE, sizeof(cd));
    cd* w = (cd*)calloc(opc_workspace_doubles(m->n_species) + 16, sizeof(cd));
    opc::g = {0, 0, 0, 0};
    opc_jacobian(m, yy, d, J, w);
    take(out);
    free(d);
    free(J);
    free(w);
}
// LU of an n x n matrix a (row-major) and one back-substitution.
void opc_count_lu(int n, const double* a, uint64_t* lu, uint64_t* solve) {
    cd* A = (cd*)malloc(sizeof(cd) * n * n);
    memcpy(A, a, sizeof(cd) * n * n);
    int* piv = (int*)malloc(sizeof(int) * n);
    cd* vv = (cd*)malloc(sizeof(cd) * n);
    cd* b = (cd*)malloc(sizeof(cd) * n);
    for (int i = 0; i < n; ++i) b[i] = cd(1.0 + i);
    opc::g = {0, 0, 0, 0};
    opc_lu_decompose(n, A, piv, vv);
    take(lu);
    opc_lu_backsubstitute(n, A, piv, b);
    take(solve);
    free(A);
    free(piv);
    free(vv);
    free(b);
}
}
