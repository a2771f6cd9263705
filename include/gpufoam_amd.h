/*
 * gpufoam_amd — C ABI of the MI355X (gfx950) batched per-cell chemistry
 * integrator.
 *
 * The library replaces the C++ seam FoamGpu::GpuKernelEvaluator and the
 * host-side mechanism converters of gpuFoam with plain C entry points:
 * opaque handles, caller-owned buffers, plain pointers and sizes, integer
 * status codes instead of exit()/assert().  No torch or HIP types appear in
 * the signatures; a stream is passed as `void*` (a hipStream_t, NULL for the
 * default stream).
 *
 * Cell state layout (structure of arrays, one column per cell):
 *   Y[k*ldY + cell]   k = 0..nSpecie-1  mass fractions
 *   Y[nSpecie*ldY + cell]               temperature [K]
 *   Y[(nSpecie+1)*ldY + cell]           pressure [Pa]
 * so OpenFOAM's per-specie volScalarFields map onto rows without a repack.
 * The reference packs the same numbers cell-major (AoS):
 *   gpu_chemistry/src/gpuChemistryModels/gpuChemistryModel/gpuChemistryModel.C:34-53.
 *
 * All entry points return GFC_OK (0) on success; on failure they return a
 * non-zero gfc_status and gfc_last_error() returns a message (thread local).
 */
#ifndef GPUFOAM_AMD_H
#define GPUFOAM_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFC_MAX_REACTION_WIDTH 10 /* gpuReaction.H:12 MAX_REACTION_WIDTH */
#define GFC_MAX_SPECIES 300       /* gpuThirdBodyEfficiencies.H:8 effArray */
#define GFC_NO_INT_EXPONENT 214748364 /* gpuLabelMax = INT_MAX/10 */

typedef enum gfc_status {
    GFC_OK = 0,
    GFC_ERR_INVALID_ARGUMENT = 1,
    GFC_ERR_PARSE = 2,           /* malformed dictionary text */
    GFC_ERR_UNSUPPORTED = 3,     /* reaction type gpuFoam rejects */
    GFC_ERR_MASS_IMBALANCE = 4,  /* makeGpuReactions.H:48-54 */
    GFC_ERR_TOO_LONG = 5,        /* reaction side > 10 species */
    GFC_ERR_TOO_MANY_SPECIES = 6,/* > 300 species */
    GFC_ERR_EFFICIENCIES = 7,    /* efficiency count != nSpecie */
    GFC_ERR_SOLVER_NAME = 8,     /* makeGpuOdeSolver.H:28 */
    GFC_ERR_CAPACITY = 9,        /* nCells > maxCells of the evaluator */
    GFC_ERR_HIP = 10,            /* HIP runtime error */
    GFC_ERR_NOT_BUILT = 11,      /* no kernel for this mechanism size */
    GFC_ERR_LTS = 12,            /* local time stepping is not supported */
    GFC_ERR_BOUNDS = 13          /* a bounds-check build caught an index out of range */
} gfc_status;

typedef enum gfc_rate_type {
    GFC_RATE_ARRHENIUS = 0,          /* gpuArrheniusReactionRate */
    GFC_RATE_THIRD_BODY = 1,         /* gpuThirdBodyArrheniusReactionRate */
    GFC_RATE_LINDEMANN = 2,          /* FallOff<Arrhenius, Lindemann> */
    GFC_RATE_TROE = 3,               /* FallOff<Arrhenius, Troe> */
    /* Rate types OpenFOAM supports and gpuFoam rejects
     * (makeMoreGpuChemistryModels.C:59-95); parameters beyond A, beta, Ta
     * are in gfc_mech_view.extra. */
    GFC_RATE_SRI = 4,                /* FallOff<Arrhenius, SRI>: extra a b c d e */
    GFC_RATE_LINDEMANN_CA = 5,       /* ChemicallyActivated<Arrhenius, Lindemann> */
    GFC_RATE_TROE_CA = 6,            /* ChemicallyActivated<Arrhenius, Troe> */
    GFC_RATE_SRI_CA = 7,             /* ChemicallyActivated<Arrhenius, SRI>: extra a b c d e */
    GFC_RATE_LANDAU_TELLER = 8,      /* LandauTeller: extra B C */
    GFC_RATE_JANEV = 9,              /* Janev: extra b0 .. b8 */
    GFC_RATE_POWER_SERIES = 10       /* powerSeries: extra c1 .. c4 */
} gfc_rate_type;

#define GFC_EXTRA_PARAMS 9 /* extra rate parameters per reaction */

typedef enum gfc_solver {
    GFC_ROSENBROCK12 = 12,
    GFC_ROSENBROCK23 = 23,
    GFC_ROSENBROCK34 = 34
} gfc_solver;

/* Per-cell status bits written to the optional cellStatus array. */
#define GFC_CELL_MAX_STEPS 0x1     /* odeSolve ran out of maxSteps (silent in the reference) */
#define GFC_CELL_STEP_UNDERFLOW 0x2/* dx < vSmall (reference: printf + assert) */
#define GFC_CELL_SKIPPED 0x4       /* T <= Treact, chemistry not solved */

/* Flags of gfc_solve. */
#define GFC_HOST_POINTERS 0x1 /* buffers are host memory; staged through
                                 device buffers owned by the evaluator */
#define GFC_BATCHED 0x2       /* persistent work-stealing launch */
#define GFC_GRAPH 0x4         /* replay the launch as a hipGraph (batched) */
#define GFC_LANE_PER_CELL 0x8 /* without GFC_BATCHED: one lane per cell, the
                                 reference's mapping (for_each_index.H:24-48),
                                 in place of the wavefront kernels */

/* odeCoeffs of chemistryProperties (readGpuOdeInputs.H:8-26). */
typedef struct gfc_ode_inputs {
    int solver; /* gfc_solver */
    double absTol;
    double relTol;
    int maxSteps;
    double safeScale;
    double alphaInc;
    double alphaDec;
    double minScale;
    double maxScale;
} gfc_ode_inputs;

/* Read-only view of a compiled mechanism (structure of arrays).  The
 * species thermo coefficients are already multiplied by R = RR/W
 * (makeGpuThermo.H:41-44); the per-reaction composite thermo (used for Kc)
 * is the one built by set_reaction_thermo (makeGpuReactions.H:18-59). */
typedef struct gfc_mech_view {
    int n_species;
    int n_reactions;
    double Tlow;  /* reaction temperature limits (reactions dict Tlow/Thigh) */
    double Thigh;
    /* species [n_species] */
    const double* W;
    const double* sTlow;
    const double* sThigh;
    const double* sTcommon;
    const double* high; /* [n_species*7] */
    const double* low;  /* [n_species*7] */
    /* reactions [n_reactions] */
    const int* type;         /* gfc_rate_type */
    const int* irreversible; /* 0/1 */
    const double* k;         /* [nr*3] A, beta, Ta: the rate (Arrhenius,
                                third body) or k0 (fall-off) */
    const double* kInf;      /* [nr*3] fall-off kInf */
    const double* troe;      /* [nr*4] alpha, Tsss, Ts, Tss */
    const double* eff;       /* [nr*n_species] third-body efficiencies */
    const int* n_lhs;
    const int* n_rhs;
    const int* lhs_index;       /* [nr*GFC_MAX_REACTION_WIDTH] */
    const double* lhs_stoich;   /* idem */
    const double* lhs_exp;      /* exponent as a scalar */
    const int* lhs_exp_int;     /* integer exponent or GFC_NO_INT_EXPONENT */
    const int* rhs_index;
    const double* rhs_stoich;
    const double* rhs_exp;
    const int* rhs_exp_int;
    /* composite reaction thermo [n_reactions] */
    const double* rY;
    const double* rW;
    const double* rTcommon;
    const double* rhigh; /* [nr*7] */
    const double* rlow;  /* [nr*7] */
    /* [nr*GFC_EXTRA_PARAMS] parameters of the extended rate types (zero
     * for gpuFoam's four) */
    const double* extra;
} gfc_mech_view;

typedef struct gfc_mechanism gfc_mechanism;
typedef struct gfc_evaluator gfc_evaluator;

/* Last error message of this thread ("" if none). */
const char* gfc_last_error(void);

/* 1 in the debug bounds-check build of the library (built with
 * GFC_BOUNDS_CHECK: indices read from the tables, work-queue cells and the
 * adjoint's recorded steps are checked before they address memory; a
 * solve or adjoint call that catches one returns GFC_ERR_BOUNDS after
 * waiting for its stream), 0 in the default build.  No counterpart in
 * gpuFoam: a debugging aid for the batched kernels' shared-memory and
 * work-stealing indexing. */
int gfc_bounds_checked(void);

/* ---- mechanism compiler (host) ----------------------------------------
 * Replaces makeGpuThermos (makeGpuThermo.H:16-66), makeGpuReactions and
 * make_reaction_rate (makeGpuReactions.H:18-142, makeGpuReactionRate.H:9-135).
 * thermo_text: an OpenFOAM physicalProperties/thermo dictionary holding
 *   `species N ( ... );` and one sub-dictionary per specie.
 * reactions_text: an OpenFOAM reactions dictionary (`reactions { ... }`,
 *   optional Tlow/Thigh). */
int gfc_mechanism_from_foam(const char* thermo_text, const char* reactions_text,
                            gfc_mechanism** out);
/* CHEMKIN-II input, in place of OpenFOAM's chemkinToFoam
 * (tutorials/h2/Allrun:10-12): chem.inp (ELEMENTS, SPECIES, optional
 * THERMO, REACTIONS with CAL/MOLE, KCAL/MOLE, JOULES/MOLE, KJOULES/MOLE or
 * KELVINS) and a NASA-7 thermo database (therm.dat).  Elementary, +M
 * third-body and (+M)/(+specie) fall-off reactions with LOW/TROE/SRI,
 * HIGH (chemically activated) and REV map to the OpenFOAM reaction types
 * chemkinToFoam writes; rate constants are converted to kmol, m^3, s and
 * Ta = Ea/R.  gfc_chemkin_to_foam writes the two OpenFOAM dictionaries
 * (call with NULL buffers to get the lengths, NUL included);
 * gfc_mechanism_from_chemkin compiles them as gfc_mechanism_from_foam
 * does. */
int gfc_chemkin_to_foam(const char* chem_inp, const char* therm_dat,
                        char* thermo_out, size_t thermo_cap, size_t* thermo_len,
                        char* reactions_out, size_t reactions_cap,
                        size_t* reactions_len);
int gfc_mechanism_from_chemkin(const char* chem_inp, const char* therm_dat,
                               gfc_mechanism** out);
void gfc_mechanism_destroy(gfc_mechanism* mech);
int gfc_mechanism_view(const gfc_mechanism* mech, gfc_mech_view* out);
/* Name of specie i (pointer valid while the mechanism lives). */
const char* gfc_mechanism_species_name(const gfc_mechanism* mech, int i);
/* Reaction string of reaction r as written ("2H + O2 = ..."). */
const char* gfc_mechanism_reaction_equation(const gfc_mechanism* mech, int r);
/* Raw (unscaled, per-mole/R) NASA-7 coefficients as read from the
 * dictionary: high[7], low[7]. */
int gfc_mechanism_raw_thermo(const gfc_mechanism* mech, int i, double* high7,
                             double* low7);
/* Index of a specie by name, -1 if absent. */
int gfc_mechanism_species_index(const gfc_mechanism* mech, const char* name);

/* ---- chemistryProperties ---------------------------------------------
 * Reads odeCoeffs { solver ...; absTol ...; ... } and Treact from a
 * chemistryProperties dictionary text with the reference's defaults
 * (readGpuOdeInputs.H:12-22; gpuChemistryModel.C:69 Treact default 0). */
int gfc_ode_inputs_from_foam(const char* chemistry_properties_text,
                             gfc_ode_inputs* out, double* Treact);
/* The switches OpenFOAM's chemistry-model base class reads from the same
 * dictionary, which gpuChemistryModel uses as chemistry_ and
 * deltaTChemMax_ (gpuChemistryModel.C:147,338,345): `chemistry on|off`
 * (required), `initialChemicalTimeStep` (required, the first deltaTChem of
 * every cell) and `maxChemicalTimeStep` (default great = 1/DBL_EPSILON).
 * Outputs may be NULL. */
int gfc_chemistry_switches_from_foam(const char* chemistry_properties_text,
                                     int* chemistry, double* initialChemicalTimeStep,
                                     double* maxChemicalTimeStep);
/* "Rosenbrock12|23|34" -> gfc_solver (makeGpuOdeSolver.H:13-31). */
int gfc_solver_from_name(const char* name, int* solver);
/* Defaults of readGpuOdeInputs.H:12-22 for the given solver. */
void gfc_ode_inputs_default(int solver, gfc_ode_inputs* out);

/* ---- evaluator (device) -----------------------------------------------
 * Replaces GpuKernelEvaluator's constructor (gpuKernelEvaluator.cu:14-33):
 * uploads the mechanism and pre-allocates every device buffer for up to
 * maxCells cells, so that gfc_solve allocates nothing.  The device is the
 * current HIP device at creation.  One evaluator per host thread. */
int gfc_evaluator_create(const gfc_mechanism* mech, const gfc_ode_inputs* ode,
                         double Treact, int maxCells, gfc_evaluator** out);
void gfc_evaluator_destroy(gfc_evaluator* ev);

/* One CFD time step of chemistry over nCells cells.
 * Replaces GpuKernelEvaluator::computeRR / computeYNew
 * (gpuKernelEvaluator.cu:35-97) and singleCellSolver (singleCellSolver.H:27-51).
 *   deltaT         CFD time step
 *   deltaTChemMax  upper clamp of the returned chemical time step
 *   ldY            leading dimension (>= nCells) of Y0, Ynew, RR
 *   rho    [nCells]            density (only read when RR != NULL)
 *   Y0     [(nS+2)*ldY]        initial state (SoA, see above)
 *   Ynew   [(nS+2)*ldY]        final state; may alias Y0; may be NULL
 *   RR     [nS*ldY]            rho*(Ynew-Y0)/deltaT; may be NULL
 *   deltaTChem [nCells]        in: chemical time step guess; out: new step
 *                              clamped to deltaTChemMax
 *   deltaTMin  [1]             out: min over cells of the new step before
 *                              clamping (computeRR: :91-94); may be NULL
 *   cellStats  [3*ldY]         optional: status bits, step attempts,
 *                              accepted steps per cell
 * Pointers are device pointers unless GFC_HOST_POINTERS is set.  Without
 * GFC_HOST_POINTERS the call is asynchronous on `stream` and allocates
 * nothing.
 * Kernels: without GFC_BATCHED one launch covers the cells (the drop-in
 * form): one wavefront per cell (nSpecie+1 <= 64, at most 3 species per
 * reaction side and 128 distinct third-body efficiency sets; four cells per
 * wavefront where the batched shape chooses that), one 512-thread workgroup per cell
 * (64 < nSpecie+1 <= 208) on a resident grid dealing cells round-robin, or
 * with GFC_LANE_PER_CELL one lane per cell.  GFC_BATCHED runs the same
 * per-cell kernels over a persistent work-stealing grid; both forms give
 * bitwise-identical results. */
int gfc_solve(gfc_evaluator* ev, double deltaT, double deltaTChemMax,
              int nCells, int ldY, const double* rho, const double* Y0,
              double* Ynew, double* RR, double* deltaTChem, double* deltaTMin,
              int* cellStats, void* stream, unsigned flags);

/* Allocation-free variant of gfc_solve's persistent batched form.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory at the call (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Device pointers only, the same
 * arguments and results as gfc_solve(..., GFC_BATCHED) without a graph.
 * The persistent kernel's work-queue counter and the deltaTMin reduction
 * word live in a library-static device array of 256 bytes per GPU
 * (resolved when an evaluator is created on that GPU, held until the
 * process ends) instead of the evaluator's buffers, and the call allocates
 * nothing.  Limits: the wavefront kernels only (nSpecie+1 <= 64, or four
 * cells per wavefront where the batched shape selects it); mechanisms that
 * need the workgroup-per-cell or one-lane kernels (per-cell scratch beyond
 * the array), or a batched shape of one lane per cell, return
 * GFC_ERR_CAPACITY, as does nCells > maxCells.  Calls on one GPU share its
 * array: calls that overlap in time on the same GPU race; serialise them
 * (one stream, or events). */
int gfc_solve_noalloc(gfc_evaluator* ev, double deltaT, double deltaTChemMax,
                      int nCells, int ldY, const double* rho, const double* Y0,
                      double* Ynew, double* RR, double* deltaTChem,
                      double* deltaTMin, int* cellStats, void* stream);

/* Launch-shape control for the persistent batched form (the autotuner's
 * search space, gpufoam_amd/autotune.py).  cellsPerWave: 0 = automatic
 * (by mechanism size: four cells per wavefront for nSpecie+1 <= 16 from
 * 16384 cells on, one wavefront per cell up to 64 (at most 128 efficiency
 * sets), one 512-thread workgroup
 * per cell up to 208, else one lane per cell), 1 = one wavefront per cell,
 * 4 = four 16-lane cells per wavefront (nSpecie+1 <= 16, at most 16
 * third-body efficiency sets, at most 3 species per reaction side; other
 * mechanisms return GFC_ERR_INVALID_ARGUMENT), 64 = one lane per cell,
 * 2 = one cell per workgroup of two wavefronts (32 < nSpecie+1 <= 64, at
 * most 16 efficiency sets, at most 3 species per reaction side; others
 * return GFC_ERR_INVALID_ARGUMENT).
 * wavesPerBlock: resident workgroups per compute unit of the persistent
 * grid (0 = as many as fit).  chunk: cells per work-queue ticket of the
 * wavefront kernel: 0 = automatic (16 consecutive cells, one 128-byte line
 * of every state row, per ticket, except the last 16 per wavefront, which
 * go out one at a time), 1 = one cell per ticket. */
int gfc_evaluator_set_batched_shape(gfc_evaluator* ev, int cellsPerWave,
                                    int wavesPerBlock, int chunk);
int gfc_evaluator_get_batched_shape(const gfc_evaluator* ev, int* cellsPerWave,
                                    int* wavesPerBlock, int* chunk);

/* Factorisation of the workgroup-per-cell kernel (64 < nSpecie+1 <= 208),
 * which factorises the Rosenbrock matrix with the row-scaled partial
 * pivoting of gpuFoam's LUDecompose (ludecompose.H:14-95):
 * GFC_LU_COLUMN eliminates column by column, one workgroup barrier per
 * column; GFC_LU_PANEL factorises four columns at a time in one wavefront
 * (one barrier per four columns) and applies the rank-4 trailing update with
 * vector FMAs (same arithmetic, factors bit-identical to GFC_LU_COLUMN);
 * GFC_LU_PANEL_MFMA applies that update with v_mfma_f64_16x16x4f64, which
 * accumulates its four products in the same order (also bit-identical).
 * The default is the measured fastest for the kernel's register block:
 * GFC_LU_PANEL_MFMA up to nSpecie+1 = 128, GFC_LU_COLUMN above.  Other
 * kernels ignore the setting. */
#define GFC_LU_COLUMN 0
#define GFC_LU_PANEL 1
#define GFC_LU_PANEL_MFMA 2
int gfc_evaluator_set_factorization(gfc_evaluator* ev, int variant);
int gfc_evaluator_get_factorization(const gfc_evaluator* ev, int* variant);

/* ---- device-level evaluations ---------------------------------------
 * The physics and the LU of one batched kernel family run on their own,
 * one cell per wavefront / 16-lane team / workgroup, as the solve runs them:
 * the counterparts of the reference's device kernels that its unit tests
 * call (gpuODESystem::derivatives and ::jacobian, gpuODESystem.H:164-281;
 * LUDecompose + LUBacksubstitute, ludecompose.H:9-140; tested in
 * unittest/tests/Test-gpuChemistry.C:521-691).  shape: 0 = the family the
 * batched solve uses for this evaluator, GFC_SHAPE_WAVE (nSpecie+1 <= 64),
 * GFC_SHAPE_PAIR (two wavefronts per cell, 32 < nSpecie+1 <= 64),
 * GFC_SHAPE_TEAM16 (nSpecie+1 <= 16), GFC_SHAPE_WORKGROUP
 * (64 < nSpecie+1 <= 208); a family that cannot take the mechanism, or a
 * mechanism with rate types beyond gpuFoam's four, returns
 * GFC_ERR_UNSUPPORTED.  Device pointers, asynchronous on `stream`.
 * Y is the SoA state of gfc_solve ([nSpecie+2][ldY]); dYdt receives
 * [n][ldY], n = nSpecie+1 (species, then T); J receives per cell the
 * row-major n x n Jacobian d(dYdt)/d(Y, T) at J + cell*n*n (the pressure
 * row and column of the reference's (nSpecie+2)^2 matrix are zero and
 * dropped); gfc_eval_jacobian's dYdt is the derivative the Jacobian pass
 * computes alongside (the reference's dfdx).  gfc_eval_lu_solve factorises
 * each cell's row-major n x n matrix M + cell*n*n with the family's LU (for
 * the workgroup family, the evaluator's factorisation variant) and solves
 * for b ([n][ldB]) into x ([n][ldB]). */
#define GFC_SHAPE_WAVE 1
#define GFC_SHAPE_PAIR 2
#define GFC_SHAPE_TEAM16 4
#define GFC_SHAPE_WORKGROUP 512
int gfc_eval_derivatives(gfc_evaluator* ev, int shape, int nCells, int ldY, const double* Y,
                         double* dYdt, void* stream);
int gfc_eval_jacobian(gfc_evaluator* ev, int shape, int nCells, int ldY, const double* Y,
                      double* dYdt, double* J, void* stream);
int gfc_eval_lu_solve(gfc_evaluator* ev, int shape, int nCells, int ldB, const double* M,
                      const double* b, double* x, void* stream);

/* ---- chemistry queries at the current state ----------------------------
 * The non-ODE queries of gpuFoam's chemistry model, which the reference
 * evaluates on the CPU with OpenFOAM's Reaction objects
 * (gpuChemistryModel.C:146-466).  Y is the SoA state of gfc_solve (T in
 * row nSpecie, p in row nSpecie+1), rho the caller's density field, and
 * c_i = rho Y_i / W_i (not clipped).  Rates are gpuReaction::omega's:
 * constants at T clipped to the reaction limits, kr = kf/Kc.  Device
 * pointers, asynchronous on `stream`, unless GFC_HOST_POINTERS is set.
 * nCells <= the evaluator's maxCells; reaction indices are in reaction
 * order of the dictionary.
 *
 * gfc_calculate: gpuChemistryModel::calculate (:146-183),
 *   RR [nSpecie*ldY]: W_i sum_r nu_ir omega_r [kg/m^3/s].
 * gfc_reaction_rr: reactionRR(reaction) (:186-240), rr [nCells]: omega_r
 *   [kmol/m^3/s].
 * gfc_specie_reaction_rr: specieReactionRR(reaction) (:242-317),
 *   RR [nSpecie*ldY]: W_i nu_ir omega_r.
 * gfc_tc: tc() (:366-443), tc [nCells]: sum(c) sum_r(wf_r + wr_r) /
 *   sum_r(wf_r^2 + wr_r^2) with wf_r = omegaf_r sum_rhs nu and
 *   wr_r = omegar_r sum_lhs nu, or DBL_MAX/10 when every rate is zero.
 * gfc_qdot: Qdot() (:445-466), Qdot [nCells]: -sum_i hf_i RR_i for a
 *   given RR [nSpecie*ldY] (e.g. the last gfc_solve's or gfc_calculate's),
 *   hf_i the specie's enthalpy of formation per unit mass. */
int gfc_calculate(gfc_evaluator* ev, int nCells, int ldY, const double* rho,
                  const double* Y, double* RR, void* stream, unsigned flags);
int gfc_reaction_rr(gfc_evaluator* ev, int reaction, int nCells, int ldY,
                    const double* rho, const double* Y, double* rr, void* stream,
                    unsigned flags);
int gfc_specie_reaction_rr(gfc_evaluator* ev, int reaction, int nCells, int ldY,
                           const double* rho, const double* Y, double* RR,
                           void* stream, unsigned flags);
int gfc_tc(gfc_evaluator* ev, int nCells, int ldY, const double* rho,
           const double* Y, double* tc, void* stream, unsigned flags);
int gfc_qdot(gfc_evaluator* ev, int nCells, int ldY, const double* RR,
             double* Qdot, void* stream, unsigned flags);

/* ---- rate parameters ----------------------------------------------------
 * No counterpart in the reference (its rates are fixed at construction,
 * makeGpuReactions.H:90-142).  Six parameters per reaction, in reaction
 * order: ln|A|, beta, Ta of the rate expression (k0 for fall-off rates),
 * then ln|A|, beta, Ta of kInf (fall-off rates; 0 and ignored otherwise),
 * with k = A T^beta exp(-Ta/T) and Ta = Ea/R (makeGpuReactionRate.H:9-14).
 * Setting keeps the sign of A and re-uploads the tables (synchronises the
 * device). */
int gfc_evaluator_param_count(const gfc_evaluator* ev);
int gfc_evaluator_get_rate_parameters(const gfc_evaluator* ev, double* params);
int gfc_evaluator_set_rate_parameters(gfc_evaluator* ev, const double* params);

/* ---- discrete adjoint -----------------------------------------------------
 * No counterpart in the reference.  Gradients of the solve's final state
 * with respect to the initial state (Y, T) and the rate parameters, for the
 * discrete map with the accepted step sizes of the forward solve held fixed.
 *
 * gfc_adjoint_forward: gfc_solve's per-cell solve (device pointers, no RR,
 * no deltaTMin, deltaTChem not clamped) that records the accepted steps.
 * For nSpecie+1 <= 16 it takes the batched form's choice of four cells per
 * wavefront (gfc_evaluator_set_batched_shape) and then returns the batched
 * solve's results; for 16 < nSpecie+1 <= 64 it runs one wavefront per cell
 * (the batched wavefront kernel with the recorder compiled in: the batched
 * solve's results) unless the batched shape asks for one lane per cell
 * (cellsPerWave 64); otherwise it returns the one-lane solve's
 * (GFC_LANE_PER_CELL).  The workspace
 * (step sizes, checkpoints every K accepted steps, per-cell scratch) is
 * allocated by this call when it grows.
 *
 * gfc_adjoint_backward uses the form its forward recorded with.  The
 * gradients of the two forms differ only by rounding.
 *
 * gfc_adjoint_backward: for the last recorded forward,
 *   Ybar     [(nS+1)*ldb]  dL/d(final Y_0..Y_nS-1, T)
 *   Y0bar    [(nS+1)*ldb]  out: dL/d(initial Y_0..Y_nS-1, T)
 *   paramBar [6*nR*ldb]    out (may be NULL): dL/dparams per cell, layout
 *                          of gfc_evaluator_get_rate_parameters
 * dL/dEa = dL/dTa / R.
 *
 * gfc_adjoint_replay: re-integrates the recorded accepted steps from Y0
 * (same cells, same layout as the forward) with the current parameters:
 * the map whose derivative gfc_adjoint_backward returns (one lane per
 * cell; after a four-cells-per-wavefront forward it reproduces that
 * forward's states to rounding).
 *
 * All three are asynchronous on `stream`. */
int gfc_adjoint_set_checkpoint_interval(gfc_evaluator* ev, int K); /* 0: auto */
int gfc_adjoint_checkpoint_interval(const gfc_evaluator* ev);
/* Workspace planning (host only, no device call; no counterpart in gpuFoam):
 * the bytes gfc_adjoint_forward allocates for nCells cells of an nSpecie
 * mechanism at checkpoint interval K, and the interval the automatic choice
 * takes for a budget: the smallest K whose whole workspace fits, searched up
 * to ceil(sqrt(maxSteps)), where the workspace is smallest (returned when
 * nothing fits).  The automatic choice uses 80 % of the free device memory
 * (plus the workspace already held) as the budget. */
double gfc_adjoint_workspace_bytes(int nCells, int nSpecie, int maxSteps, int K);
int gfc_adjoint_plan_interval(double budgetBytes, int nCells, int nSpecie, int maxSteps);
int gfc_adjoint_forward(gfc_evaluator* ev, double deltaT, int nCells, int ldY,
                        const double* Y0, double* deltaTChem, double* Ynew,
                        int* cellStats, void* stream);
int gfc_adjoint_backward(gfc_evaluator* ev, const double* Ybar, int ldb,
                         double* Y0bar, double* paramBar, void* stream);
int gfc_adjoint_replay(gfc_evaluator* ev, const double* Y0, int ldY,
                       double* Ynew, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GPUFOAM_AMD_H */
