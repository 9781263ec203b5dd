// spb6_kernels.h -- device-side structures and kernel launchers of the SPB 6+1 IMEX step (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include "spb6_device.h"

namespace spb {

// error / event flags written by kernels (int slots in device memory): the flags of one evaluation first, then the
// stage's admissibility flags, which a stage keeps until its first residual is read back, then the closure failure of
// the final state of an attempt's first stage, which the next stage reads back with its first residual
enum { FLAG_RES = 0, FLAG_JAC = 1, FLAG_FLOOR = 2, FLAG_FINAL = 3, FLAG_RES2 = 4, FLAG_RHO = 5, FLAG_STAGE = 6, FLAG_FINAL_PREV = 7, NFLAGS = 8 };

struct Geo {
	int nx, ny, NXT, NYT, NC, n;   // interior sizes, ghosted sizes, ghosted cell count, interior cell count
	int twoD, xper, yper;
	int bc[4]; double bcu[4];
	double dx, dy;
};

// BiCGStab loop scalars kept on the device between kernels (rho, alpha, omega double-buffered by iteration parity)
struct KS {
	double Einv, cc_p, cc_s, bn, rho[2], alpha[2], omega[2];
	int done, iters, early, status;
};

// all device pointers of one stage (passed by value to the residual-type kernels)
struct Dev {
	Geo g; Phys ph;
	double P_REF, rx, ry, tau;
	const double *Ust_c;
	double *Ust, *a2s, *ubx, *uby, *phx, *phy, *vsc, *esc, *vu, *vv;
	double *wxL, *wxR, *wyL, *wyR;   // full face states (the batch's update)
	double *donx, *dony;             // face donors per direction: for each of m1, m2, e1, e2 an array of (left, right) pairs
	double *ust, *vst, *cz, *Kw, *Pold;
	double *ubsx, *kpx, *wLx, *pvx, *ubsy, *kpy, *wLy, *pvy, *Qph;
	int* flags;
	// IMEX variants (VAR_*): the TV-type variants' explicit phase energies and total energy, BB's alpha_1 step, and TV's
	// alpha_1 face donors per direction as (left, right) pairs; null for the reference's PC scheme
	int variant;
	double *eTV1, *eTV2, *ETV, *dA, *donax, *donay;
	const int* ctype;   // relaxation levels 1-3: each interior row's closure type of the step (CL_P / CL_PT / CL_PTG); null at level 0
	RelaxDev rel;         // the saturation table (relaxation levels 2-3; n = 0 otherwise) and the finite rates
	double* Z;            // finite mu: the phases' acoustic impedances of the stage's CL_MU cells (2 NC; null otherwise)
};
// the IMEX variants of the reference (its sibling programs SPB6_<name>_IMEX.cpp); VAR_PC is SPB6_PC_IMEX.cpp
enum { VAR_PC = 0, VAR_TV = 1, VAR_TVPURE = 2, VAR_TVINC = 3, VAR_TVSPLIT = 4, VAR_TVBB = 5, VAR_BB = 6 };
// the implicit stage takes the general cell assembly (cell_assemble_var) for an IMEX variant or relaxation level 1
inline bool general_path(const Dev& D) { return D.variant != VAR_PC || D.ctype != nullptr; }

template <typename T> struct KrylovBufsT {
	T *x, *r, *r0, *p, *v, *s, *t;
	const T* J;
	double* part; int stride; KS* ks;
	T *ph = nullptr, *sh = nullptr, *dg = nullptr;   // red-black sweep preconditioner: M^-1 p, M^-1 s, 1 / diagonal
	T* p2 = nullptr;                                 // second buffer of p (the red-black form alternates p between p, p2)
};
using KrylovBufs = KrylovBufsT<double>;
using KrylovBufs32 = KrylovBufsT<float>;
constexpr int KRYLOV_SLOTS = 11;   // partial-sum slots of the Krylov passes
int krylov_items(int n);

void launch_bc_grid(const Geo& g, double* u, hipStream_t st);
// zero n ints with a kernel (capturable in a graph)
void launch_zero_ints(int* p, long n, hipStream_t st);
void launch_bc_array(const Geo& g, double* a, hipStream_t st);
void launch_combine(const Geo& g, const double* Un, const double* H0, const double* H1, int nr, double ae0, double ae1, double ai0, double ai1, double* Qe, double* Qs, hipStream_t st);
void launch_explicit(const Dev& D, int limiter, int order, const double* Qe, const double* Qs, double adt, hipStream_t st);
// full 8-field face states of Qe (TVDReconstruction) into the given grids, for inspection: the explicit stage itself
// keeps only the face donors
void launch_recon_full(const Dev& D, int limiter, int order, const double* Qe, double* wxL, double* wxR, double* wyL, double* wyR, hipStream_t st);
void launch_prep(const Dev& D, int limiter, int order, double tau, const double* Pst, hipStream_t st);
// shock budget in two passes; list holds n ints, count two ints (the list's length and the strong pass's cursor), rec
// SHOCK_REC * n doubles (the strong cells' profile problems), all device scratch; rk_steps accumulates the strong
// pass's RK45 steps (its work count); strong_blocks: persistent 256-thread blocks of the strong-shock pass (as many as
// are resident at once)
void launch_shock(const Dev& D, int* list, int* count, double* rec, unsigned long long* rk_steps, int strong_blocks, int fused, hipStream_t st);
// 256-thread blocks of the strong-shock pass resident on one compute unit at once
int shock_strong_blocks_per_cu();
// (gate: launched speculatively, the evaluation runs only where *gate is set)
void launch_residual(const Dev& D, const double* P, const double* dP, double lam, double* Pout, double* G, unsigned long long* nrm, int flag_slot, hipStream_t st,
                     const int* gate = nullptr);
// close of an FP32 Krylov sweep behind a chunk: dP += x once the solve is done, the defect test into *spec, the trial
// residual's norm slots (nrm[0..63]) and flag cleared
// the residual on the rows list[0, *count) only (blocks: the grid, each thread striding over the list)
void launch_residual_list(const Dev& D, const double* P, const int* list, const int* count, double* G, unsigned long long* nrm, int flag_slot,
                          int blocks, hipStream_t st);
void launch_spec_commit(long n, const float* x, double* dP, const KS* ks, const double* dsum, double eta, int* spec, unsigned long long* nrm, int* flags,
                        hipStream_t st);
// clear (optional): nclear ints zeroed by the step's first pass (a Newton step's norm slots and flags)
void launch_jacobian(const Dev& D, const double* P, const double* G, double* sc, double* J, double* rhs, bool analytic, hipStream_t st,
                     int* clear = nullptr, int nclear = 0);
// FP64 defect b - J (dP + x) and the step criterion of dP + x for a running FP32 solve x: (r.r, b.b, criterion) into sums[3]
void launch_defect_x(const Geo& g, const double* J, const double* b, const double* dP, const float* x, double* r, double* part, int stride,
                     double* sums, const double* P, const double* sc, hipStream_t st);
void launch_chord_rhs(const Dev& D, const double* P, const double* G, double* sc, double* rhs, hipStream_t st, int* clear = nullptr, int nclear = 0);
void launch_stepcrit(int n, const double* dP, const double* P, const double* sc, unsigned long long* nrm, hipStream_t st);
void launch_floor(const Dev& D, const double* P, double tol, hipStream_t st);
// tmax: 128 slots, the running maxima of the two phase temperatures (bit patterns of non-negative doubles, 64 striped
// slots each)
void launch_final(const Dev& D, const double* P, const double* Qs, double adt, double* H, unsigned long long* tmax, hipStream_t st,
                  int flag_slot = FLAG_FINAL);
void launch_caldt(const Geo& g, const Phys& P, const double* U, int acoustic, unsigned long long* nrm, hipStream_t st);
// cell_diag's 15 columns of every interior cell of U into out[q * 15 + k]
void launch_cell_diag(const Geo& g, const Phys& P, const double* U, double* out, hipStream_t st);
void launch_thomas(int n, const double* J, const double* rhs, double* x, double* cp, double* dp, hipStream_t st);
void launch_krylov_setup(const Geo& g, const KrylovBufs& B, const double* b, hipStream_t st);
void launch_krylov_iter(const Geo& g, const KrylovBufs& B, hipStream_t st, int it, double tol, int maxit);
void launch_add_f32(long n, const float* d, double* x, hipStream_t st);
void launch_to_f32(long n, const double* a, float* b, hipStream_t st);
void launch_krylov_setup32(const Geo& g, const KrylovBufs32& B, const float* b, hipStream_t st);
// the red-black FP32 setup of a Newton step's first sweep straight from the FP64 system: b32 = (float) b64, dP = 0, and
// with J64 (a new Jacobian) J32 = (float) J64, in the setup's first pass
void launch_krylov_setup32_cvt(const Geo& g, const KrylovBufs32& B, const double* J64, float* J32, const double* b64, float* b32, double* dP,
                               hipStream_t st);
void launch_krylov_iter32(const Geo& g, const KrylovBufs32& B, hipStream_t st, int it, double tol, int maxit);

// ---- the IMEX variants (spb6_variants.hip): the explicit stage from the full face states wxL.. (launch_recon_full),
// then the variant's update; the implicit-stage kernels of a Dev on the general path (an IMEX variant or relaxation
// level 1: residual, finite-difference Jacobian, floor test, final state) take the general cell residual.
// launch_explicit routes a variant's Dev here, and launch_residual, launch_jacobian, launch_floor and launch_final a
// Dev on the general path.
void launch_explicit_var(const Dev& D, int limiter, int order, const double* Qe, const double* Qs, double adt, hipStream_t st);
void launch_residual_var(const Dev& D, const double* P, const double* dP, double lam, double* Pout, double* G, unsigned long long* nrm, int flag_slot,
                         hipStream_t st, const int* gate);
void launch_jacobian_var(const Dev& D, const double* P, const double* G, const double* sc, double* J, double* rhs, hipStream_t st);
void launch_floor_var(const Dev& D, const double* P, double tol, hipStream_t st);
void launch_final_var(const Dev& D, const double* P, const double* Qs, double adt, double* H, unsigned long long* tmax, hipStream_t st, int flag_slot);
// relaxation level 1 (RelaxationActivation, PC:537-560): each interior cell's closure type into ctype, the cells of
// type pT projected onto the pT equilibrium in U (a failed projection keeps the cell mechanical), then U's ghosts
void launch_relax_activate(const Geo& g, const Phys& P, const RelaxDev& S, int level, double* U, double eps_i, int* ctype, hipStream_t st);

// ---- the explicit HLLC scheme (spb6_hllc.hip): cell primitives (hllc_prim_fields arrays over the ghosted grid), face
// values (hllc_face_fields arrays per direction, at the face's low cell); every pass returns at once once flags
// FLAG_STAGE or FLAG_FINAL is set
int hllc_prim_fields();
int hllc_face_fields();
void launch_hllc_prim(const Geo& g, const Phys& P, const double* U, double* W, int* flags, hipStream_t st);
void launch_hllc_flux(const Geo& g, const Phys& P, const double* W, int order, int limiter, double* FX, double* FY, const int* flags, hipStream_t st);
void launch_hllc_update(const Geo& g, const double* U, const double* W, const double* FX, const double* FY, double rx, double ry, double* Uo,
                        int* flags, hipStream_t st);
// fails: closure failures counted (one int)
void launch_hllc_relax(const Geo& g, const Phys& P, double* U, const double* Qph, int* flags, int* fails, hipStream_t st);
void launch_hllc_average(long n, double* V, const double* W, const int* flags, hipStream_t st);

}   // namespace spb
