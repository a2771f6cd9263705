/*
 * CPU oracle for the gpufoam_amd chemistry integrator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of gpuFoam's
 * per-cell chemistry path (the same arithmetic, in the same order, as the
 * reference's header-only device code compiled for the host).  It is used by
 * tests/ as the checker, by bench.py as the CPU baseline, and by
 * __graft_entry__.smoke(); the product library never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative
 * to the gpuFoam repository root).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - against the reference's own headers built for the host (oracle/_ref),
 *     bit for bit, on thermo, rates, RHS, Jacobian, LU and ODE solves;
 *   - against golden vectors committed under tests/golden/;
 *   - against closed forms and invariants (dHa/dT = Cp, mass and energy
 *     conservation of the RHS, finite-difference Jacobian, LU residual,
 *     Rosenbrock order on linear problems).
 */
#ifndef GFC_ORACLE_H
#define GFC_ORACLE_H

#include "../include/gpufoam_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scratch size in doubles for one cell of a mechanism with nS species. */
long orc_workspace_doubles(int nS);

/* --- thermo (gpuThermo.H) --- */
void orc_species_thermo(const gfc_mech_view* m, int i, double p, double T,
                        double out[6]); /* Cp, Ha, Hf, S, Gstd, dCpdT */
void orc_reaction_Kc(const gfc_mech_view* m, int r, double p, double T,
                     double out[2]); /* Kc, dKcdTbyKc (KcAnddKcTbyKc) */

/* --- reactions (gpuReaction.H) --- */
/* out: kf, kr, omega (omega clips T; kf/kr at the clipped T) */
void orc_reaction_rates(const gfc_mech_view* m, int r, double p, double T,
                        const double* c, double out[3]);
/* computeReactionParameters: out[0]=omega, out[1]=dwdT,
 * dCf[10], dCr[10], ddc[nS] (ddc only written when the rate depends on c) */
void orc_reaction_parameters(const gfc_mech_view* m, int r, double p,
                             double T, const double* c, double out[2],
                             double* dCf, double* dCr, double* ddc);

/* --- ODE system (gpuODESystem.H), y and dydx of length nS+2 --- */
void orc_derivatives(const gfc_mech_view* m, const double* y, double* dydx,
                     double* work);
/* J is (nS+2)x(nS+2) row-major; dfdx receives the derivative */
void orc_jacobian(const gfc_mech_view* m, const double* y, double* dfdx,
                  double* J, double* work);

/* --- LU (ludecompose.H) --- */
void orc_lu_decompose(int n, double* a, int* piv, double* vv);
void orc_lu_backsubstitute(int n, const double* a, const int* piv, double* b);

/* --- ODE solvers --- */
/* One Rosenbrock step of size dx from y0 (dydx0 = f(y0)); writes y and
 * returns the normalised error (gpuRosenbrock{12,23,34}.H::solve). */
double orc_rosenbrock_step(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                           const double* y0, const double* dydx0, double dx,
                           double* y, double* work);
/* odeSolve (gpuODESolverBase.H:67-111); stats[0]+=attempts, stats[1]+=accepted;
 * returns GFC_CELL_* bits */
int orc_ode_solve(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                  double xStart, double xEnd, double* y, double* dxTry,
                  double* work, int* stats);

/* singleCellSolver + computeRR over nCells cells in SoA layout (see
 * gpufoam_amd.h).  nThreads <= 0 uses OpenMP's default.  Returns 0. */
int orc_solve_cells(const gfc_mech_view* m, const gfc_ode_inputs* ode,
                    double Treact, double deltaT, double deltaTChemMax,
                    int nCells, int ldY, const double* rho, const double* Y0,
                    double* Ynew, double* RR, double* deltaTChem,
                    double* deltaTMin, int* cellStats, int nThreads);

/* --- chemistry queries at the current state (gpuChemistryModel.C:146-466)
 * Y: SoA [(nS+2)*ldY], T at row nS, p at row nS+1; rho [nCells], the
 * thermo's density; c_i = rho Y_i / W_i, not clipped. */
/* OpenFOAM's Reaction::omega(p, T, c, li, omegaf, omegar) split:
 * out[0] = kf Cf, out[1] = kr Cr (rates at the clipped T) */
void orc_reaction_omega_fr(const gfc_mech_view* m, int r, double p, double T,
                           const double* c, double out[2]);
/* calculate(), :146-183: RR[i*ldY+cell] = W_i dNdtByV_i */
void orc_calculate(const gfc_mech_view* m, int nCells, int ldY,
                   const double* rho, const double* Y, double* RR);
/* reactionRR(r), :186-240: out[cell] = omega_r */
void orc_reaction_rr(const gfc_mech_view* m, int r, int nCells, int ldY,
                     const double* rho, const double* Y, double* out);
/* specieReactionRR(r), :242-317: RR[i*ldY+cell] = W_i nu_ir omega_r */
void orc_specie_reaction_rr(const gfc_mech_view* m, int r, int nCells,
                            int ldY, const double* rho, const double* Y,
                            double* RR);
/* tc(), :366-443 */
void orc_tc(const gfc_mech_view* m, int nCells, int ldY, const double* rho,
            const double* Y, double* tc);
/* Qdot(), :445-466: Qdot[cell] = -sum_i hf_i RR_i */
void orc_qdot(const gfc_mech_view* m, int nCells, int ldY, const double* RR,
              double* Qdot);

#ifdef __cplusplus
}
#endif
#endif
