/*
 * spb6_imex.h -- C ABI of the MI355X (gfx950) semi-implicit IMEX-RK time step of the SPB 6+1-equation two-phase model.
 *
 * Drop-in boundary for diffusionIMEX's SPB6_PC_IMEX stepper. The reference has no FFI; its "interface" is the set of
 * free functions, globals and grid types of src/SPB6_PC_IMEX.cpp ("PC") and src/spb6_physics.h ("PHY"). Each entry point
 * below names the reference function or global it replaces (file:line). Plain C types only: pointers and sizes.
 *
 * Ownership: the caller owns host arrays; the context owns every device buffer (allocated once in spb6_create, the
 * analogue of PC_omp's reused static buffers, src/SPB6_PC_IMEX_omp.cpp:802, :912-918). No allocation per stage or step.
 * Threading: one context per host thread; calls on one context are not re-entrant. Results are run-to-run deterministic
 * (fixed-order reductions).
 * Layout at the boundary: the reference's AoS cell order U[(i + j*nx)*8 + k], interior cells only (PC:104, PC:124);
 * on the device the state is structure-of-arrays with two ghost layers.
 */
#ifndef SPB6_IMEX_H
#define SPB6_IMEX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Conserved variables per cell (PHY:25-26): alpha1, m1, m2, rho u, rho v, eps1, eps2, E */
#define SPB6_NUM_VARS 8

/* boundary kinds per side (left, right, bottom, top) -- PHY:906 BC_TRANS / BC_WALL / BC_PERIODIC */
enum spb6_bc { SPB6_BC_TRANSMISSIVE = 0, SPB6_BC_WALL = 1, SPB6_BC_PERIODIC = 2 };

/* status codes: the reference returns bool false for every one of these and halves dt (PC:1287-1290) */
enum spb6_status {
    SPB6_OK = 0,
    SPB6_RHO_NONPOSITIVE = 1,     /* rho <= 0 in Qe (PC:257) */
    SPB6_STAGE_INADMISSIBLE = 2,  /* sum alpha* <= 0 or rho* <= 0 (PC:862) */
    SPB6_CLOSURE_FAIL = 3,        /* inadmissible closure input/output in the final state (PHY:787-790, PHY:490) */
    SPB6_NEWTON_FAIL = 4,         /* no descent and no round-off floor, or no convergence (PC:842, PC:846) */
    SPB6_KRYLOV_FAIL = 5,         /* BiCGStab breakdown or no convergence (PC:743, PC:757) */
    SPB6_DT_UNDERFLOW = 6,        /* dt < 1e-12 t_end (PC:1240) */
    SPB6_STEP_FAILED = 7,         /* 30 halvings without an admissible step (PC:1290) */
    SPB6_BAD_ARG = 8,
    SPB6_HIP_ERROR = 9
};

/* IMEX variants: the reference's sibling programs, the same time step with another explicit update and cell residual
 * (src/SPB6_<name>_IMEX.cpp; docs/devlog.md section 23.6 tabulates them) */
enum spb6_variant {
    SPB6_VARIANT_PC = 0,       /* SPB6_PC_IMEX.cpp: the paper scheme */
    SPB6_VARIANT_TV = 1,       /* SPB6_TV_IMEX.cpp: alpha moved by its flux and the increments, s = 1 - alpha_1 - alpha_2 */
    SPB6_VARIANT_TVPURE = 2,   /* SPB6_TVpure_IMEX.cpp: TV split of the energy (kinetic explicit, enthalpy flux implicit) */
    SPB6_VARIANT_TVINC = 3,    /* SPB6_TVinc_IMEX.cpp: TVpure + mass / momentum increments */
    SPB6_VARIANT_TVSPLIT = 4,  /* SPB6_TVsplit_IMEX.cpp: TVinc + explicit phase-energy dissipation */
    SPB6_VARIANT_TVBB = 5,     /* SPB6_TVbb_IMEX.cpp: TVpure + explicit phase-energy dissipation */
    SPB6_VARIANT_BB = 6        /* SPB6_BB_IMEX.cpp: TVpure with the face states moved to the explicit new alpha_1 */
};

/* Phase EOS parameters (PHY:49-54 Phase{g, pi, q, qp, cv, b}); stiffened gas when b = 0 */
typedef struct spb6_phase {
    double gamma, pi_inf, q, qp, cv, b;
} spb6_phase_t;

/* Replaces the reference's configuration globals (PC:57-80, PHY:33-37, PHY:906-913). */
typedef struct spb6_config {
    int nx, ny;                      /* interior cells (ny = 1: 1D, PHY:952-956) */
    double x0, x1, y0, y1;           /* domain (TestSetup, PHY:898) */
    int bc[4];                       /* BCS[4] (PHY:908) */
    double bcu[4];                   /* BCU[4]: wall normal velocity (PHY:908) */
    spb6_phase_t ph[2];              /* PH[2] (PHY:54) */
    int rk;                          /* RKs: 1 backward Euler, 2 LSDIRK2, 3 SSP2(3,3,2) (PC:57, PC:1195-1218) */
    double cfl, cfl_ac;              /* CFL, CFL_ac (PC:59-60) */
    int ramp;                        /* RAMP (PC:63) */
    double dtfix;                    /* DTFIX (PC:65) */
    int acdt;                        /* caltimestep == Fullydt (PC:50) */
    double newton_tol; int newton_iter;   /* NewtonTol, NewtonIter (PC:67-70) */
    double krylov_tol; int krylov_iter;   /* GMREStol, GMRESiter (PC:71-72) */
    int limiter;                     /* LIMITER (PHY:913) */
    int order;                       /* SpaceOrder (PC:61) */
    int rus_local;                   /* RUS_LOCAL (PC:80); only 1 is supported on the device */
    /* additions without a reference counterpart */
    int krylov_fp32;                 /* 1: BiCGStab vectors/matrix in FP32 inside the FP64 Newton (defect correction) */
    int learned_guess;               /* 1: Newton initial guess from the MLP (spb6_mlp_*), plain guess as fallback */
    int analytic_jacobian;           /* 1: Newton Jacobian from the closure's implicit derivative at the residual's root
                                        instead of five finite-difference closure solves per row (PC:790-804); with
                                        krylov_fp32 the last Newton step as a chord step (the solve's previous
                                        Jacobian) when the residual is within 100x of the tolerance after a full step
                                        that contracted 1000x */
    int jreuse;                      /* JREUSE (src/SPB6_PC_IMEX_omp.cpp:83, :796-888): modified Newton; the Jacobian (and
                                        its scale s) is rebuilt at a stage's first iteration and after a damped step or a
                                        step that contracted the residual less than 4x, and a line search that fails with
                                        a reused Jacobian rebuilds it and retakes the step. Single-grid stepper only (the
                                        batch rejects it) */
    int variant;                     /* spb6_variant: which of the reference's IMEX programs the step follows (single-grid
                                        stepper, finite-difference Jacobian and plain guess only; the batch rejects it) */
    int relax;                       /* RELAX_LEVEL (PHY:37): 0 p (mechanical equilibrium); in the cells with
                                        eps_i <= alpha_1 <= 1 - eps_i also 1 p-pT (thermal equilibrium: the pT closure,
                                        PHY:660-672), 2 p-pT-pTg (chemical equilibrium where the pT state exceeds the
                                        saturation temperature, else thermal), 3 p-pTg (chemical where a phase at the
                                        mechanical equilibrium exceeds it, else mechanical): the pTg closure on the
                                        saturation line T_sat(p), PHY:111-163 and PHY:756-778, and the step's
                                        RelaxationActivation, PC:537-560, with the reference's instantaneous rates.
                                        Finite-difference Jacobian and plain guess only; the batch and the explicit
                                        HLLC scheme reject levels 1-3 */
    double eps_i;                    /* EPS_I (PHY:38); <= 0: the reference's 1e-4 */
    int graph;                       /* 1: each stage's fixed head (stage operands, explicit part, preparation, shock budget)
                                        and each Krylov chunk (BiCGStab passes; the FP32 form's defect, sweep close and
                                        trial residual) launched as a captured hipGraph; the same kernels, the same bits */
    int shock_fused;                 /* 1: the strong-shock profile integrations (PHY:302-341) in fused form: the ODE's two
                                        phase terms over one denominator (one division per right-hand side instead of
                                        ten) and FMA stage sums; the same equation and step control, rounded differently
                                        (the shares agree with the reference arithmetic's to ~1e-15 relative). Grid
                                        stepper only (the batch rejects it) */
    double theta_rate, nu_rate;      /* theta_rate, nu_rate (PHY:33): finite thermal / chemical relaxation rates at
                                        relaxation levels >= 1 / 2-3 (the theta closure PHY:679-700, the nu closure
                                        PHY:716-753, backward-Euler form); <= 0: the level's default, instantaneous
                                        (PC:1148-1149) */
    int eos_kind[2];                 /* Phase::kind (PHY:50-53): 0 NASG / SG with ph[k]; 1 Cochran-Chan, 2 JWL: a
                                        Mie-Grueneisen phase p = G rho (e - er(rho)) + pr(rho) with mg[k] (ph[k] unused).
                                        Relaxation level 0 (the reference's MG phases relax mechanically only, PHY:790),
                                        the reference scheme, the finite-difference Jacobian, the FP64 Krylov solve, the
                                        plain guess and the exact profile form only; the batch and the explicit HLLC
                                        scheme reject it */
    double mg[2][7];                 /* a Mie-Grueneisen phase's G, c1, c2, e1, e2, rr (reference density), e0 (JWL's
                                        energy shift): Cochran-Chan pr = c1 x^e1 - c2 x^e2 with x = rho / rr, JWL
                                        pr = c1 exp(-e1 rr / rho) + c2 exp(-e2 rr / rho) (PHY:58-66) */
    double mu_rate;                  /* mu_rate (PHY:33): a finite mechanical relaxation rate (the mu closure, PHY:575-656,
                                        backward-Euler form, with the stage's frozen impedances PC:956-958) in the cells of
                                        mechanical type; <= 0: instantaneous (inf, the default). The reference scheme, the
                                        finite-difference Jacobian, the plain guess and NASG phases only; the batch and the
                                        explicit HLLC scheme reject it */
    int nu_law;                      /* NU_LAW (PHY:34-36, the command line's nulaw=1): at levels 2-3 the chemical cells take
                                        the nu closure with Pelanti's state-dependent rate nu = nu_cr alpha_v^0.6
                                        ((psat(T) - p) / (nu_pc - psat(T)))^1.76 (PHY:172-177), frozen per cell at each
                                        attempt's activation from the cell's pT equilibrium (PC:546-556) */
    double nu_cr, nu_pc;             /* NU_CR, NU_PC (PHY:36); <= 0: 1 and 22.064e6 */
    int relax_form;                  /* RELAX_FORM (PHY:39-41): 0 backward Euler over the stage's step (the default); 1 the
                                        exponential solution of the finite-rate closures (PHY:685, PHY:721-745, PHY:644-647:
                                        T2 - T1 decays by exp(-tau theta S); Y and the volume transfer move towards their
                                        equilibria along secants) */
} spb6_config_t;

/* run statistics: the counters of the reference's stats line (PC:1337-1339) plus Krylov counts */
typedef struct spb6_stats {
    long steps, rk1, rejects, floor_acc, stages, newton_it, krylov_it, linsolves;
    double maxres, t, dt_last;
    int status;
    long guess_used, guess_fallback;  /* learned-guess stages, and stages where the plain guess was kept */
    double guess_log10_gain;          /* sum over stages of log10(plain residual norm / chosen guess's residual norm) */
    long guess_skipped;               /* stages of guess_fallback where the network was not evaluated: after two
                                         stages in a row where the plain guess won, 1, 2, 4, ... 32 stages are skipped
                                         before the network is tried again (a win resets the back-off) */
    double tmax[2];                   /* the phases' largest temperatures in the stages' final states where the phase is
                                         present (alpha > 1e-9), the reference's st_Tmax (PC:986-989) */
    long closure_fails;               /* stages whose final state failed in the closure (an attempt rejected for it) */
    long jac_builds;                  /* Newton Jacobians built (finite-difference or analytic; jreuse and chord steps
                                         reuse one) */
} spb6_stats_t;

typedef struct spb6_ctx spb6_ctx;

/* Create a context on HIP device `device`: allocates all device buffers for the grid (main's Grid/Array/PField
 * allocations, PC:1157-1163, and StageData, PC:565-582). Returns NULL on error (spb6_last_error tells why). */
spb6_ctx* spb6_create(const spb6_config_t* cfg, int device);
void spb6_destroy(spb6_ctx* ctx);
const char* spb6_last_error(void);

/* Upload interior AoS state (u_initialization + bound_cond_u, PC:450-475) and initialise the pressure unknown
 * P = pbar - P_REF with P_REF = min pbar (PC:1222-1227). Resets time and counters. */
int spb6_set_state(spb6_ctx* ctx, const double* U_aos);
/* Download interior AoS state. */
int spb6_get_state(spb6_ctx* ctx, double* U_aos);
/* Pressure unknown P (interior, nx*ny, PField PC:155) and the scalar P_REF (PC:68). */
int spb6_get_pressure(spb6_ctx* ctx, double* P, double* p_ref);
int spb6_set_pressure(spb6_ctx* ctx, const double* P, double p_ref);

/* caldt (PC:477-488): returns 1 / (max(|u|+c)/dx + max(|v|+c)/dy) in *inv_rate (c = frozen speed if acoustic). */
int spb6_caldt(spb6_ctx* ctx, int acoustic, double* inv_rate);
/* bound_cond_u on the state (PHY:926-957). */
int spb6_bc_fill(spb6_ctx* ctx);

/* Explicit part of the first stage of an attempt (Qe = Qs = U) with step adt: TVDReconstruction (PC:349) + Update
 * (PC:251); a later stage's operands go through spb6_stage_combine / spb6_stage_update below. */
int spb6_explicit_stage(spb6_ctx* ctx, double adt);
/* Implicit part of that first stage (Qs = U, stage derivative Hs[0]): PressureSubsystem (PC:851-993) = stage prep, shock
 * budget, Newton_Solver (PC:761) with Thomas (PC:807-812) or BiCGStab (PC:725-758), final state. The stage's Newton
 * guess is the current pressure. Any stage: spb6_stage_pressure below. */
int spb6_implicit_stage(spb6_ctx* ctx, double adt, int* newton_iters, double* residual);
/* Only the stage preparation and shock budget of PressureSubsystem (PC:859-969), for per-kernel checks. */
int spb6_stage_prep(spb6_ctx* ctx, double adt);
/* One attempt of a step of size dt with an rk-stage table (body of the attempt loop, PC:1243-1285); on success the
 * state and pressure advance. */
int spb6_step(spb6_ctx* ctx, double dt, int rk);
/* ---- one attempt of the reference's step, stage by stage (the attempt body PC:1243-1285), for a caller that keeps its
 * own time loop (PC:1230-1296). Ae, Ai are the caller's Butcher tables (set_tab, PC:1195-1218: rk 1, 2 or 3 stages):
 *
 *   spb6_attempt_begin(ctx);                                     u_con_n = u_con, Pst = P             (PC:1247-1250)
 *   for (rks = 0; rks < rk; rks++) {
 *       spb6_stage_combine(ctx, rks, dt, Ae[rks], Ai[rks]);       Qe, Qs from the stage derivatives     (PC:1252-1266)
 *       ok = spb6_stage_update(ctx, Ai[rks][rks] * dt) == 0;      TVDReconstruction(Qe); Update(Ust, Qe, Qs)   (PC:1268-1271)
 *       ok = ok && spb6_stage_pressure(ctx, Ai[rks][rks] * dt, rks, &it, &res) == 0;
 *                                                                  PressureSubsystem(Ust, Qs, Qe, Pst, adt, rks) and
 *                                                                  Hs[rks] = (Ust - Qs) / adt          (PC:1273-1284)
 *   }
 *   ok ? spb6_attempt_accept(ctx, dt) : halve dt and retry         u_con = Ust, P = Pst, t += dt       (PC:1286-1291)
 *
 * Each call waits for its own kernels and reports its own failure, so a replay gives the bits of spb6_run. */
/* u_con_n = u_con and Pst = P: the stage operands start from the state of the step (PC:1247-1250). */
int spb6_attempt_begin(spb6_ctx* ctx);
/* Qe = U + dt sum_{r<rks} ae[r] Hs[r], Qs = U + dt sum_{r<rks} ai[r] Hs[r], then bound_cond_u on both (PC:1252-1266);
 * rks = 0: Qe = Qs = U. rks <= 2; Hs[2] exists only in a context created with rk = 3. */
int spb6_stage_combine(spb6_ctx* ctx, int rks, double dt, const double* ae, const double* ai);
/* TVDReconstruction(Qe) (PC:349) + Update(Ust, Qe, Qs) (PC:251) with step adt into X*; SPB6_RHO_NONPOSITIVE where the
 * reference's Update returns false (PC:255-258). */
int spb6_stage_update(spb6_ctx* ctx, double adt);
/* PressureSubsystem(Ust, Qs, Qe, Pst, adt, rks) (PC:851-993): stage data, shock budget, Newton from the current Pst,
 * final state into Ust, and the stage derivative Hs[rks] = (Ust - Qs) / adt (PC:1276-1284). */
int spb6_stage_pressure(spb6_ctx* ctx, double adt, int rks, int* newton_iters, double* residual);
/* u_con = Ust, P = Pst, t += dt, steps + 1 (PC:1286, PC:1291). */
int spb6_attempt_accept(spb6_ctx* ctx, double dt);
/* The time loop (PC:1230-1296): dt from caldt with the start-up ramp, reject-and-halve, until t_end or max_steps
 * accepted steps (max_steps < 0: no cap). */
int spb6_run(spb6_ctx* ctx, double t_end, long max_steps);
/* RelaxationActivation (PC:537-560) on the state: at relaxation levels 1-3 each attempt starts with it; the reference's
 * main applies it once more before its output (PC:1300), which a caller mirrors by calling this before
 * spb6_get_state / spb6_write_profile. No-op at level 0. */
int spb6_relax_activate(spb6_ctx* ctx);
/* ---- the reference's fully explicit scheme, src/SPB6_explicit.cpp ("EX"), at relaxation level 0, on the same context
 * and state (an independent cross-check of the IMEX path on shock problems; it leaves the pressure unknown alone):
 * primitive-variable minmod reconstruction (EX:228-256; cfg.order, cfg.limiter), HLLC fluxes with the phase Hugoniot
 * states (EX:145-192), the Godunov update (EX:194-225), the mechanical relaxation by the p closure with the shock
 * budget's defect split (EX:305-361), Heun's two stages (EX:546-556). The first call allocates the scheme's work arrays
 * (cell primitives and face values: 27 doubles per ghosted cell in 2D). */
/* One step of size dt (EX:546-556): SPB6_OK, SPB6_STAGE_INADMISSIBLE (a cell prim_of rejects, or alpha_1, m_1, m_2 out
 * of range after an update) or SPB6_CLOSURE_FAIL (the relaxation's closure; counted in stats.closure_fails). The state
 * changes only on success. */
int spb6_hllc_step(spb6_ctx* ctx, double dt);
/* Its time loop (EX:537-566): dt = min(cfl * caldt(acoustic), t_end - t) (the reference's cfl is 0.4), failed steps
 * halved up to 30 times (stats.rejects), SPB6_DT_UNDERFLOW below 1e-12 t_end, SPB6_STEP_FAILED after 30 halvings. */
int spb6_hllc_run(spb6_ctx* ctx, double t_end, long max_steps, double cfl);
int spb6_get_stats(spb6_ctx* ctx, spb6_stats_t* stats);
double spb6_time(spb6_ctx* ctx);

/* Residual G_i(P) = pbar_cl,i - P_i of every interior cell at the given P (residual_at, PC:699-705) after
 * spb6_stage_prep; returns the number of cells whose closure failed (-1 on error). */
int spb6_residual(spb6_ctx* ctx, const double* P, double* G);
/* Finite-difference Jacobian rows (Newton_Solver's SpRow, PC:790-804): val[5*n] = (self, W, E, S, N), scale sc[n]. */
int spb6_jacobian(spb6_ctx* ctx, const double* P, const double* G, double* val, double* sc);
/* Linear solve with the 5-point (2D) / 3-point (1D) rows val[5*n]: BiCGStab (use_thomas = 0, PC:725) or Thomas
 * (PC:807-812). */
int spb6_linsolve(spb6_ctx* ctx, const double* val, const double* b, double* x, int use_thomas, int* iters);

/* No counterpart in the reference: an allocation-free form of spb6_linsolve for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The caller passes device
 * buffers: J[5 * n] the five diagonals (self, W, E, S, N; J[w * n + r]) of the 5-point (2D) / 3-point (1D) rows as
 * spb6_jacobian gives them, b[n], x[n] the solution, on an nx x ny grid with boundary kinds bc[4] (periodic sides wrap).
 * BiCGStab (use_thomas = 0: tolerance tol |b|, at most maxit iterations, the reference's Jacobi + coarse-space
 * preconditioner, GMRES_Solver PC:725-758) or Thomas (use_thomas = 1, ny = 1, PC:807-812); the same kernels and sums as
 * spb6_linsolve, so the same bits. Work vectors live in a fixed device array of the library (about 12.7 MB per device,
 * held until the process ends): at most SPB6_LINSOLVE_DEVICE_MAX_ROWS rows (SPB6_BAD_ARG above it, no work done). No
 * hipMalloc, hipFree or other allocation. stream: a hipStream_t (NULL: the null stream); the call waits for it. Calls
 * on one device share the array: calls that overlap on one device (other threads, other streams) race. */
#define SPB6_LINSOLVE_DEVICE_MAX_ROWS 262144
int spb6_linsolve_device(int nx, int ny, const int* bc, const double* J, const double* b, double* x, double tol, int maxit,
                         int use_thomas, int* iters, void* stream);

/* Copy a named device field to the host (ghosted fields: (nx+4)*(ny+4) values, grids 8x that in AoS order).
 * Names: U Ust Qe Qs wxL wxR wyL wyR phx phy a2s ubx uby ust vst cz Kw ubsx kpx wLx pvx ubsy kpy wLy pvy Qph P Pst.
 * The face states wxL wxR wyL wyR (TVDReconstruction's, PC:392-445) are rebuilt from the last explicit stage's Qe for
 * the call (the stage itself keeps only the face donors the implicit stage reads); the call allocates their scratch.
 * Returns the number of doubles written, or -1. */
long spb6_get_field(spb6_ctx* ctx, const char* name, double* out, long capacity);

/* ---- output (the reference's write_profile, conserved_totals and stats line) */
/* write_profile's columns (PC:1002-1013) of every interior cell of the current state, out[(i + j nx) * 17 + k]:
 * x y p u v rho alpha2 Y2 rho1 rho2 c_frozen c_wood T1 T2 p1 p2 g1-g2 (cell_diag, PHY:1105-1116, computed on the device;
 * NaN where prim_of fails; g1 - g2 takes the device log, every other column is the reference's arithmetic). */
int spb6_get_profile(spb6_ctx* ctx, double* out);
/* write_profile (PC:1002-1013): the file with the reference's header line and columns, prec significant decimals after
 * the point for the state columns (OUTPREC, default 16 in the reference; 8 for c_frozen, c_wood, g1-g2; 10 for T1, T2). */
int spb6_write_profile(spb6_ctx* ctx, const char* path, int prec);
/* conserved_totals (PC:1104-1107): sum m1, sum m2, sum E over the interior cells in the reference's cell order. */
int spb6_conserved_totals(spb6_ctx* ctx, double M[3]);
/* The reference's run statistics line (PC:1337-1339) into buf: case id and wall time from the caller, the counters of
 * spb6_get_stats, and dM1/M, dM2/M, dE/E against the totals of the state set last. */
int spb6_stats_line(spb6_ctx* ctx, int case_id, double wall_s, char* buf, long len);

/* ---- per-launch timing of the dominant kernels (HIP events on the context's stream) */
typedef struct spb6_kernel_times {
    double explicit_ms, prep_ms, shock_ms, residual_ms, jacobian_ms, krylov_ms, final_ms, other_ms;
    long explicit_n, prep_n, shock_n, residual_n, jacobian_n, krylov_n, final_n;
    long krylov_iters;
    double krylov_aux_ms;        /* BiCGStab setup, FP32 conversions and FP64 defects (krylov_ms: iteration passes only) */
    double guess_ms;             /* learned-guess features and inference */
    long krylov_launched;        /* BiCGStab iterations launched (a pass launched after convergence returns at once) */
    double krylov_idle_ms;       /* FP32 form: time of the iterations launched past a solve's end (krylov_ms: taken ones) */
    long shock_rk_steps;         /* RK45 profile steps taken by the shock budget's strong-shock passes (its work count) */
    long graph_launches;         /* graph mode: captured graphs launched, and graphs instantiated anew (a shape that */
    long graph_rebuilds;         /* hipGraphExecUpdate could not apply to the graph of its slot) */
} spb6_kernel_times_t;
int spb6_enable_timing(spb6_ctx* ctx, int on);
/* categories timed while timing is on (default all): bit k = the k-th category in the order explicit, prep, shock,
 * residual, jacobian, krylov, final, other, krylov_aux, guess, krylov_idle; the others record no events, so the step
 * runs with fewer of them (no counterpart in the reference: instrumentation of this library) */
int spb6_set_timing_mask(spb6_ctx* ctx, unsigned mask);
int spb6_get_kernel_times(spb6_ctx* ctx, spb6_kernel_times_t* t);

/* ---- learned Newton guess (no counterpart in the reference; the reference's guess is the previous stage's pressure,
 * PC:1250, PC:1286). A 175-128-128-1 GELU network maps the 5x5 stencil of 7 stage fields to the stage's pressure change;
 * with cfg.learned_guess the stepper evaluates the Newton residual at the plain and at the learned guess and starts
 * Newton from the one with the smaller residual norm (the plain guess is the fallback and the judge). */
typedef struct spb6_mlp spb6_mlp;
spb6_mlp* spb6_mlp_create(int device);
void spb6_mlp_destroy(spb6_mlp* m);
int spb6_mlp_nparams(void);
/* params: FP32 parameter vector (W1[175x128], b1, W2[128x128], b2, w3[128], b3); norm: feature means[175], feature
 * standard deviations[175], target mean, target standard deviation (NULL keeps the current ones) */
int spb6_mlp_set_params(spb6_mlp* m, const float* params, const float* norm);
int spb6_mlp_get_params(spb6_mlp* m, float* params);
/* training set of standardised features X[n x 175] and targets Y[n], copied to the device */
int spb6_mlp_set_data(spb6_mlp* m, const float* X, const float* Y, long n);
/* Adam epochs on the device data (hand-written FP32 forward/backward kernels); loss_out[epochs] */
int spb6_mlp_train(spb6_mlp* m, int epochs, int batch, float lr, unsigned long long seed, float* loss_out);
/* FP32 forward, FP32 gradients of mean((y - t)^2), and the BF16 MFMA inference path on host features */
int spb6_mlp_forward(spb6_mlp* m, const float* X, int n, float* y);
int spb6_mlp_gradients(spb6_mlp* m, const float* X, const float* Y, int n, float* grad, float* loss);
int spb6_mlp_infer_bf16(spb6_mlp* m, const float* X, int n, float* y);
/* stepper side: use this network for the learned guess (the network must outlive the context's use of it) */
int spb6_set_guess_model(spb6_ctx* ctx, spb6_mlp* m);
/* learned guess at plain guess P after spb6_stage_prep (no counterpart in the reference: the check of the library's
 * learned guess): the guess as a stage computes it (features of the listed tiles' stencils only) into Pg and the same
 * guess from the features of every cell into Pg_full; the guess's residual as a stage computes it -- evaluated only on
 * the cells whose residual inputs it changes, the plain residual elsewhere -- into G, the same residual evaluated on
 * every cell into G_full, and the max-norms of the two (PC:770) into norms[2]; returns the number of cells evaluated
 * (-1 on error, or without a network) */
int spb6_guess_residual(spb6_ctx* ctx, const double* P, double* Pg, double* G, double* G_full, double* norms, double* Pg_full);
/* record raw (features, target) pairs of every stage's cells q = off, off + stride, ... into a device buffer of
 * `capacity` rows (capacity 0 disables); spb6_dump_fetch copies up to max_rows rows to the host and empties it */
int spb6_dump_enable(spb6_ctx* ctx, long capacity, int stride);
long spb6_dump_fetch(spb6_ctx* ctx, float* X, float* Y, long max_rows);

/* ---- batched 1D problems: many independent runs of the reference's 1D path (main's time loop PC:1230-1296 with
 * PressureSubsystem PC:851-993 and its Thomas solve PC:807-812, one run per problem, each with its own length, domain,
 * boundaries and phases) advanced together on one GPU. Each problem keeps its own dt, start-up ramp, first-order ramp
 * steps and reject-and-halve loop and gives the bits of a single-problem run of the same configuration. Scheme settings
 * (rk, CFL, ramp, tolerances, limiter, order) are shared and taken from a spb6_config_t whose grid, boundary and phase
 * fields are ignored. Periodic boundaries are not accepted (the reference solves periodic 1D rows with BiCGStab). */
typedef struct spb6_batch spb6_batch;
typedef struct spb6_batch_problem {
    int nx;                      /* interior cells */
    double x0, x1;               /* domain */
    int bc[2];                   /* left, right: SPB6_BC_TRANSMISSIVE or SPB6_BC_WALL (BCS[0..1], PHY:908) */
    double bcu[2];               /* wall normal velocity (BCU[0..1]) */
    spb6_phase_t ph[2];          /* the problem's phases (PH[2], PHY:54) */
} spb6_batch_problem_t;
/* launch shapes tried by spb6_batch_autotune and the one kept (no counterpart in the reference) */
typedef struct spb6_batch_tune {
    int threads, per_cu;         /* workgroup size (64, 128, 256) and persistent workgroups per compute unit */
    double best_ms;              /* wall time of the trial at the kept shape */
    int ntried; int tried_threads[16], tried_per_cu[16]; double tried_ms[16];
} spb6_batch_tune_t;

/* Allocate the batch (every device buffer, once); problems keep their index order at the boundary. */
spb6_batch* spb6_batch_create(const spb6_batch_problem_t* problems, int nproblems, const spb6_config_t* scheme, int device);
void spb6_batch_destroy(spb6_batch* batch);
const char* spb6_batch_last_error(spb6_batch* batch);
/* total interior cells (sum of nx) */
long spb6_batch_cells(spb6_batch* batch);
/* U: the problems' AoS states back to back (sum nx * 8 doubles); per problem P_REF = min pbar and P = pbar - P_REF
 * (PC:1222-1227), t = 0, start-up ramp armed, counters cleared */
int spb6_batch_set_state(spb6_batch* batch, const double* U);
int spb6_batch_get_state(spb6_batch* batch, double* U);
/* P: sum nx doubles; pref: one per problem (may be NULL) */
int spb6_batch_get_pressure(spb6_batch* batch, double* P, double* pref);
/* Advance every problem by `steps` accepted steps or to t_end, whichever comes first (spb6_run per problem). Returns
 * SPB6_STEP_FAILED if any problem failed (its status is in its stats; the others still ran). */
int spb6_batch_run(spb6_batch* batch, double t_end, long steps);
/* per-problem counters (array of nproblems, may be NULL) and their totals (may be NULL) */
int spb6_batch_get_stats(spb6_batch* batch, spb6_stats_t* per_problem, spb6_stats_t* total);
/* launch shape: threads per workgroup (64, 128, 256) and persistent workgroups per compute unit (1..16); results do
 * not depend on it (no counterpart in the reference) */
int spb6_batch_set_launch(spb6_batch* batch, int threads, int per_cu);
/* time `steps` steps at each candidate shape on a snapshot of the current state, restore the state, keep the fastest
 * shape (no counterpart in the reference) */
int spb6_batch_autotune(spb6_batch* batch, long steps, spb6_batch_tune_t* out);

#ifdef __cplusplus
}
#endif
#endif
