// spb6_batch.hip -- batched 1D stepper: thousands of independent 1D two-phase problems of uneven lengths, each with its
// own stiffened-gas phases, time step, start-up ramp and reject-and-halve loop, advanced together on one GPU.
//
// Every problem runs the reference's 1D path (the SPB6_PC_IMEX time loop, PC:1230-1296, with PressureSubsystem and its
// Thomas solve, PC:851-993 / PC:807-812) exactly: each problem's arithmetic is the grid path's, cell for cell, so a
// problem gives the same bits as a single-problem run. Only the execution differs:
//  - one persistent launch per IMEX stage for the whole batch: each workgroup takes problems from a longest-first queue
//    (one atomic ticket per problem) and runs the problem's whole stage -- explicit part, preparation, shock budget,
//    Newton with finite-difference Jacobian, Thomas and line search, final state -- with workgroup barriers between
//    the phases and no grid-wide synchronisation (problems are independent);
//  - the per-problem time loop (dt, ramp, first-order ramp steps, reject and halve) lives on the device: stage 1 of a
//    step attempt computes the problem's dt when the step is new, stage 2 accepts or rejects;
//  - the two launches of a step are captured once in a hipGraph per launch shape and replayed step after step;
//  - the launch shape (workgroup size, persistent workgroups per CU) is chosen by an autotuner that times graph
//    replays on a snapshot of the state; results do not depend on the shape (all reductions are maxima).
// Layout: structure of arrays over the concatenation of the problems' ghosted rows (L_p + 4 cells each, field stride
// NT = sum of them); interior vectors (P, G, Jacobian rows, ...) over the concatenated interiors (NI = sum L_p).
#include "spb6_device.h"
#include "spb6_kernels.h"
#include "spb6_cell.h"
#include "../../include/spb6_imex.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace spb {
namespace {

struct BProb {                 // immutable description of one problem
	int L, off, poff;          // cells; first ghosted cell in the field arrays; first interior row in the vectors
	int bc[2]; double bcu[2];  // left / right boundary
	double dx;
	Phys ph;
};
struct BState {                // the problem's time loop (device resident)
	double t, dt, dtprev, P_REF, maxres, dt_last;
	long steps, target, rejects, newton_it, linsolves, floor_acc, stages, rk1;
	long nt0;                  // newton_it when the current step began
	double cost;               // work of the last completed step: (Newton iterations + 2 stages) * cells
	int phase;                 // 0: the next attempt starts a new step (compute dt), 1: an attempt of this step is pending
	int ramping, startup, rk, att, par, ppar, stage_ok, status, done;
};
struct BArgs {
	const BProb* prob; BState* st; const int* order; int np; int* qctr; int* nfin;
	int NT;
	double *U0, *U1, *H0, *H1, *Qe, *Qs, *wxL, *wxR, *donx;
	double *vsc, *esc, *phx, *a2s, *ubx, *ust, *vst, *cz, *Kw, *Pold, *ubsx, *kpx, *wLx, *pvx, *Qph;
	double *P0, *P1, *Pn, *G, *Gn, *dP, *sc, *rhs, *J, *cpw, *dpw;
	int* slist;                // strong-shock cells of each problem (sum L ints)
	// scheme (shared by all problems)
	int limiter, order_, rk, ramp, acdt, newton_iter;
	double cfl, cfl_ac, dtfix, tol, t_end;
	double ai00_rk1, ai00, ae10, ai10, ai11;   // LSDIRK2 / backward Euler tableaux (PC:1195-1218)
};

__device__ __forceinline__ double bmax_wave(double v) {
	#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) { const double o = __shfl_xor(v, off, 64); v = smax(v, o); }
	return v;
}
// workgroup maximum of a non-negative value (every thread gets it). Both reductions start with a workgroup barrier,
// which also orders the preceding phase's memory writes before the next phase reads them (a single-wavefront
// workgroup included: lanes of one wavefront are not ordered without it).
template <int T>
__device__ __forceinline__ double wg_max(double v, double* red) {
	__syncthreads();
	v = bmax_wave(v);
	if constexpr (T == 64) return v;
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
	__syncthreads();
	double m = red[0];
	#pragma unroll
	for (int w = 1; w < T / 64; w++) m = smax(m, red[w]);
	return m;
}
// workgroup OR of a flag (every thread gets it)
template <int T>
__device__ __forceinline__ int wg_any(int f, int* red) {
	__syncthreads();
	const int w = __ballot(f) != 0ull ? 1 : 0;
	if constexpr (T == 64) return w;
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
	__syncthreads();
	int m = 0;
	#pragma unroll
	for (int q = 0; q < T / 64; q++) m |= red[q];
	return m;
}

// per-problem geometry and device view: pointers shifted so that the grid path's ghosted index c = i + GH * NXT
// (interior row j = GH of a one-row grid) addresses the problem's cells; field arrays carry a front pad so the shift
// stays inside the allocation
struct View { Geo g; Dev D; double *U, *Ust, *H0, *H1, *Qe, *Qs; double *P, *Pst, *Pn, *G, *Gn, *dP, *sc, *rhs, *J, *cpw, *dpw; int* slist; };
__device__ __forceinline__ View make_view(const BArgs& A, const BProb& B, const BState& S) {
	View v;
	Geo& g = v.g;
	g.nx = B.L; g.ny = 1; g.NXT = B.L + 2 * GH; g.NYT = 1 + 2 * GH; g.NC = A.NT; g.n = B.L;
	g.twoD = 0; g.xper = 0; g.yper = 0;
	g.bc[0] = B.bc[0]; g.bc[1] = B.bc[1]; g.bc[2] = BC_TRANS; g.bc[3] = BC_TRANS;
	g.bcu[0] = B.bcu[0]; g.bcu[1] = B.bcu[1]; g.bcu[2] = 0.0; g.bcu[3] = 0.0;
	g.dx = B.dx; g.dy = 1.0;
	const long sh = (long)B.off - (long)GH * g.NXT;
	auto f = [&](double* p) { return p + sh; };
	v.U = f(S.par ? A.U1 : A.U0); v.Ust = f(S.par ? A.U0 : A.U1);
	v.H0 = f(A.H0); v.H1 = f(A.H1); v.Qe = f(A.Qe); v.Qs = f(A.Qs);
	Dev& D = v.D;
	D.g = g; D.ph = B.ph; D.P_REF = S.P_REF; D.rx = 0; D.ry = 0; D.tau = 0;
	D.Ust_c = v.Ust; D.Ust = v.Ust; D.a2s = f(A.a2s); D.ubx = f(A.ubx); D.uby = D.ubx; D.phx = f(A.phx); D.phy = D.phx;
	D.vsc = f(A.vsc); D.esc = f(A.esc); D.vu = nullptr; D.vv = nullptr; D.wxL = f(A.wxL); D.wxR = f(A.wxR); D.wyL = D.wxL; D.wyR = D.wxR;
	D.donx = A.donx + 2 * sh; D.dony = D.donx;   // pairs: a shift of sh cells is 2 sh doubles
	D.ust = f(A.ust); D.vst = f(A.vst); D.cz = f(A.cz); D.Kw = f(A.Kw); D.Pold = f(A.Pold);
	D.ubsx = f(A.ubsx); D.kpx = f(A.kpx); D.wLx = f(A.wLx); D.pvx = f(A.pvx);
	D.ubsy = D.ubsx; D.kpy = D.kpx; D.wLy = D.wLx; D.pvy = D.pvx; D.Qph = f(A.Qph); D.flags = nullptr;
	const long po = B.poff;
	v.P = (S.ppar ? A.P1 : A.P0) + po; v.Pst = (S.ppar ? A.P0 : A.P1) + po;
	v.Pn = A.Pn + po; v.G = A.G + po; v.Gn = A.Gn + po; v.dP = A.dP + po; v.sc = A.sc + po; v.rhs = A.rhs + po;
	v.J = A.J + 3 * po; v.cpw = A.cpw + po; v.dpw = A.dpw + po; v.slist = A.slist + po;
	return v;
}

// x ghost fill of a conserved row (bound_cond_u / bc_apply in 1D, PHY:926-957)
template <int T>
__device__ __forceinline__ void wg_bc_grid(const Geo& g, double* u) {
	__syncthreads();
	if (threadIdx.x < 2 * GH) bc_x_at(g, u, GH, threadIdx.x >> 1, threadIdx.x & 1);
	__syncthreads();
}

// Newton_Solver (PC:761-847) on the problem's pressure, iterating in the buffers (Pst, Pn) and (G, Gn); the converged
// iterate ends in v.Pst. Returns an SPB6 status.
template <int T>
__device__ int wg_newton(const BArgs& A, View& v, BState& S, double* red, int* ired, double& resfinal) {
	const Dev& D = v.D; const int n = v.g.n; const double tol = A.tol;
	double* Pst = v.Pst; double* Pn = v.Pn; double* G = v.G; double* Gn = v.Gn;
	auto eval = [&](const double* P, const double* dP, double lam, double* Pout, double* Gout, bool& ok) -> double {
		double m = 0.0; int bad = 0;
		for (int q = threadIdx.x; q < n; q += T) { double nv; if (!residual_at(D, P, dP, lam, Pout, Gout, q, nv)) bad = 1; m = smax(m, nv); }
		ok = !wg_any<T>(bad, ired);
		return wg_max<T>(m, red);
	};
	int rc = SPB6_NEWTON_FAIL;
	bool ok; double res = eval(Pst, nullptr, 0.0, nullptr, G, ok);
	resfinal = res;
	if (!ok) return SPB6_NEWTON_FAIL;
	int it = 0;
	for (; it < A.newton_iter; it++) {
		resfinal = res;
		if (res < tol) { rc = SPB6_OK; break; }
		for (int q = threadIdx.x; q < n; q += T) fdscale_at(D, Pst, v.sc, q);
		__syncthreads();
		int badj = 0;
		for (int q = threadIdx.x; q < n; q += T) if (!jacobian_at(D, Pst, G, v.sc, v.J, v.rhs, q)) badj = 1;
		__syncthreads();
		if (threadIdx.x == 0) { S.linsolves++; thomas_serial(n, v.J, v.rhs, v.dP, v.cpw, v.dpw); }
		__syncthreads();
		double sm = 0.0;
		for (int q = threadIdx.x; q < n; q += T) sm = smax(sm, fabs(v.dP[q]) / (fabs(Pst[q]) + v.sc[q]));
		const double sn = wg_max<T>(sm, red);
		if (wg_any<T>(badj, ired)) { rc = SPB6_NEWTON_FAIL; goto out; }
		if (sn < tol) {   // step criterion (PC:819-826)
			res = eval(Pst, v.dP, 1.0, Pn, Gn, ok);
			{ double* t = Pst; Pst = Pn; Pn = t; t = G; G = Gn; Gn = t; }
			if (threadIdx.x == 0) S.newton_it++;
			rc = ok ? SPB6_OK : SPB6_NEWTON_FAIL; resfinal = res; goto out;
		}
		{
			double lam = 1.0; bool acc = false;
			for (int ls = 0; ls < 40; ls++) {   // Armijo line search (PC:827-832)
				const double rn = eval(Pst, v.dP, lam, Pn, Gn, ok);
				if (ok && (rn < res * (1.0 - 1e-4 * lam) || rn < tol)) { acc = true; res = rn; break; }
				lam *= 0.5;
			}
			if (!acc) {   // round-off floor acceptance (PC:834-842)
				resfinal = res;
				int badf = 0;
				for (int q = threadIdx.x; q < n; q += T) if (!floor_ok_at(D, Pst, tol, q)) badf = 1;
				if (wg_any<T>(badf, ired)) { rc = SPB6_NEWTON_FAIL; goto out; }
				if (threadIdx.x == 0) S.floor_acc++;
				rc = SPB6_OK; goto out;
			}
		}
		{ double* t = Pst; Pst = Pn; Pn = t; t = G; G = Gn; Gn = t; }
		if (threadIdx.x == 0) S.newton_it++;
	}
	if (it == A.newton_iter) { resfinal = res; rc = res < tol ? SPB6_OK : SPB6_NEWTON_FAIL; }
out:
	// the converged iterate is the caller's Pst buffer
	if (Pst != v.Pst) {
		__syncthreads();
		for (int q = threadIdx.x; q < n; q += T) v.Pst[q] = Pst[q];
	}
	__syncthreads();
	return rc;
}

// one IMEX stage of one problem (stage 0: Qe = Qs = U; stage 1: the LSDIRK2 combination); returns an SPB6 status
template <int T>
__device__ int wg_stage(const BArgs& A, View& v, BState& S, int s, double dt, double* red, int* ired) {
	const Geo& g = v.g; const int NXT = g.NXT, n = g.n, NC = g.NC;
	const double* Qe = v.U; const double* Qs = v.U;
	double* H = s == 0 ? v.H0 : v.H1;
	if (s == 1) {   // stage combination (PC:1255-1264)
		const double ae0 = A.ae10 * dt, ai0 = A.ai10 * dt;
		for (int q = threadIdx.x; q < n; q += T) {
			const int c = q + GH + GH * NXT;
			for (int k = 0; k < NV; k++) {
				const int o = k * NC + c;
				double e = v.U[o], z = v.U[o];
				e = e + ae0 * v.H0[o]; z = z + ai0 * v.H0[o];
				v.Qe[o] = e; v.Qs[o] = z;
			}
		}
		wg_bc_grid<T>(g, v.Qe); wg_bc_grid<T>(g, v.Qs);
		Qe = v.Qe; Qs = v.Qs;
	}
	const double adt = s == 0 ? (S.rk == 1 ? A.ai00_rk1 : A.ai00) : A.ai11;
	const double tau = adt * dt;
	Dev& D = v.D; D.tau = tau; D.rx = tau / g.dx; D.ry = tau / g.dy;
	// explicit part: TVDReconstruction + Update (PC:349, PC:251)
	int bad = 0;
	for (int i = threadIdx.x; i < NXT; i += T) if (!scales_at(g, D.ph, Qe, D.vsc, D.esc, i + GH * NXT)) bad = 1;
	__syncthreads();
	for (int i = threadIdx.x; i < NXT; i += T) recon_x_at(g, D.ph, A.limiter, A.order_, Qe, D.vsc, D.esc, D.wxL, D.wxR, D.phx, i + GH * NXT, D.donx);
	__syncthreads();
	for (int q = threadIdx.x; q < n; q += T) update_at(g, tau / g.dx, tau / g.dy, Qe, Qs, D.wxL, D.wxR, D.wyL, D.wyR, D.Ust, D.a2s, D.ubx, D.uby, q);
	__syncthreads();
	if (threadIdx.x < 2 * GH) bc_x_at(g, D.Ust, GH, threadIdx.x >> 1, threadIdx.x & 1);
	if (threadIdx.x >= 2 * GH && threadIdx.x < 3 * GH) bcarr_x_at(g, D.a2s, GH, threadIdx.x - 2 * GH);
	if (wg_any<T>(bad, ired)) return SPB6_RHO_NONPOSITIVE;
	// stage preparation and shock budget (PC:859-969)
	bad = 0;
	for (int i = threadIdx.x; i < NXT; i += T) if (!prep_cell_at(g, D.ph, D.Ust, D.a2s, v.Pst, D.ust, D.vst, D.cz, D.Kw, D.Pold, i + GH * NXT)) bad = 1;
	__syncthreads();
	for (int i = threadIdx.x; i < NXT; i += T)
		prep_face_at(g, A.order_, A.limiter, tau, D.Ust, D.ust, D.vst, D.cz, D.Pold, D.phx, D.phy, D.ubsx, D.kpx, D.wLx, D.pvx,
		             D.ubsy, D.kpy, D.wLy, D.pvy, i + GH * NXT);
	// shock budget in two passes: cells without a strong shock are settled at once, the strong-shock cells are gathered
	// and then run with every lane busy
	__shared__ int scnt;
	if (threadIdx.x == 0) scnt = 0;
	__syncthreads();
	for (int q = threadIdx.x; q < n; q += T) if (shock_classify_at(g, D.ph, D.Ust, D.Qph, q)) v.slist[atomicAdd(&scnt, 1)] = q;
	__syncthreads();
	const int nstrong = scnt;
	for (int k = threadIdx.x; k < nstrong; k += T) shock_strong_at(g, D.ph, D.Ust, D.Qph, v.slist[k]);
	if (wg_any<T>(bad, ired)) return SPB6_STAGE_INADMISSIBLE;   // (barrier inside)
	double res;
	const int rc = wg_newton<T>(A, v, S, red, ired, res);
	if (rc != SPB6_OK) return rc;
	if (threadIdx.x == 0) S.maxres = smax(S.maxres, res);
	// final state at the converged pressure (PC:975-992)
	bad = 0;
	for (int q = threadIdx.x; q < n; q += T) if (!final_at(D, v.Pst, Qs, tau, H, q)) bad = 1;
	const int fb = wg_any<T>(bad, ired);
	wg_bc_grid<T>(g, D.Ust);
	if (fb) return SPB6_CLOSURE_FAIL;
	if (threadIdx.x == 0) S.stages++;
	return SPB6_OK;
}

// caldt (PC:477-488) of one problem: 1 / (max |u| + c) * dx, infinite when every cell is at rest
template <int T>
__device__ double wg_caldt(const BArgs& A, const View& v, bool acoustic, double* red) {
	double m = 0.0;
	for (int q = threadIdx.x; q < v.g.n; q += T) { double ux, uy; caldt_at(v.g, v.D.ph, v.U, acoustic ? 1 : 0, q, ux, uy); m = smax(m, ux); }
	const double ux = wg_max<T>(m, red);
	const double rate = ux / v.g.dx;
	return rate > 0 ? 1.0 / rate : dinf();
}

// one problem's share of launch s (0: stage 1 of an attempt, 1: stage 2); the step logic follows spb6_run
template <int T>
__device__ void wg_problem(const BArgs& A, int s, int p, double* red, int* ired) {
	BState& Sg = A.st[p];
	__shared__ BState S;
	if (threadIdx.x == 0) S = Sg;
	__syncthreads();
	if (S.done || S.status != SPB6_OK || (s == 1 && !S.stage_ok)) return;
	const BProb B = A.prob[p];
	View v = make_view(A, B, S);
	if (s == 0) {
		if (S.phase == 0) {   // a new step: dt with the start-up ramp (PC:1233-1241)
			const double ramp_factor = 1.2;
			double dt; bool ramping = false; double dtprev = S.dtprev;
			if (A.dtfix > 0) dt = A.dtfix;
			else {
				dt = A.cfl * wg_caldt<T>(A, v, A.acdt != 0, red);
				if (A.ramp) {
					if (dtprev < 0) dtprev = A.cfl_ac * wg_caldt<T>(A, v, true, red) / ramp_factor;
					if (dt > ramp_factor * dtprev) { dt = ramp_factor * dtprev; ramping = true; }
				}
				if (!isfin(dt)) dt = A.cfl_ac * wg_caldt<T>(A, v, true, red);
			}
			dt = ((A.t_end - S.t) < dt) ? (A.t_end - S.t) : dt;
			__syncthreads();
			if (threadIdx.x == 0) {
				S.dtprev = dtprev; S.dt = dt; S.ramping = ramping; S.att = 0; S.phase = 1; S.nt0 = S.newton_it;
				S.rk = (A.ramp && ramping && S.startup) ? 1 : A.rk;
				if (!(dt > 1e-12 * A.t_end)) S.status = SPB6_DT_UNDERFLOW;
			}
			__syncthreads();
			if (S.status != SPB6_OK) { if (threadIdx.x == 0) { Sg = S; atomicAdd(A.nfin, 1); } return; }
		}
	}
	// Attempts (PC:1243-1290). A rejected attempt is retried at once in this launch with dt / 2 -- from stage 1, and on
	// through stage 2 when this is the stage-2 launch -- so each launch pair completes one accepted step per problem.
	int st = s;
	for (;;) {
		if (st == 0) {
			if (threadIdx.x == 0 && S.rk == 1) S.rk1++;
			for (int q = threadIdx.x; q < v.g.n; q += T) v.Pst[q] = v.P[q];   // the attempt starts from the step's P
			__syncthreads();
		}
		const int rc = wg_stage<T>(A, v, S, st, S.dt, red, ired);
		__syncthreads();
		if (threadIdx.x == 0) {
			if (rc != SPB6_OK) {   // reject and halve (PC:1286-1290)
				S.rejects++; S.dt *= 0.5; S.att++; S.stage_ok = 0;
				if (S.att >= 30) { S.status = SPB6_STEP_FAILED; atomicAdd(A.nfin, 1); }
			} else if (st == 0 && S.rk == 2) S.stage_ok = 1;
			else {                 // accept (PC:1285-1296)
				S.par ^= 1; S.ppar ^= 1; S.stage_ok = 0; S.phase = 0;
				S.t = S.t + S.dt; S.dtprev = S.dt; S.dt_last = S.dt; S.steps++;
				S.cost = (double)(S.newton_it - S.nt0 + 2) * (double)B.L;
				if (!S.ramping) S.startup = 0;
				if (S.steps >= S.target || !(S.t < A.t_end * (1 - 1e-12))) { S.done = 1; atomicAdd(A.nfin, 1); }
			}
			Sg = S;
		}
		__syncthreads();
		if (S.status != SPB6_OK || S.phase == 0) break;   // failed, or accepted
		if (S.stage_ok) { if (s == 0) break; st = 1; }     // stage 1 done: stage 2 belongs to the stage-2 launch
		else st = 0;                                       // rejected: the attempt restarts at stage 1
	}
}

template <int T>
__global__ void __launch_bounds__(T) k_batch_stage(BArgs A, int s) {
	__shared__ double red[T / 64];
	__shared__ int ired[T / 64];
	__shared__ int tk;
	if (blockIdx.x == 0 && threadIdx.x == 0) A.qctr[1 - s] = 0;   // the other launch's queue, idle now
	for (;;) {
		if (threadIdx.x == 0) tk = atomicAdd(&A.qctr[s], 1);
		__syncthreads();
		const int t = tk;
		__syncthreads();
		if (t >= A.np) break;
		const int p = __builtin_amdgcn_readfirstlane(A.order[t]);
		wg_problem<T>(A, s, p, red, ired);
	}
}

// ghost fill of every problem's row of U (after the state is set)
__global__ void __launch_bounds__(64) k_batch_bc(BArgs A, int which) {
	const int p = blockIdx.x;
	if (p >= A.np) return;
	const BProb B = A.prob[p];
	BState S = A.st[p];
	S.par = which;
	const View v = make_view(A, B, S);
	if (threadIdx.x < 2 * GH) bc_x_at(v.g, v.U, GH, threadIdx.x >> 1, threadIdx.x & 1);
}

}   // namespace
}   // namespace spb

using namespace spb;

// ================================================================================================ C ABI
struct spb6_batch {
	int device = 0; hipStream_t st = nullptr; int cus = 256;
	spb6_config_t cfg{};
	std::vector<BProb> hprob; std::vector<int> horder;
	std::vector<void*> allocs;
	BArgs A{};
	long NT = 0, NI = 0, pad = 0;
	int* hnfin = nullptr;     // pinned
	struct Shape { int threads, per_cu; hipGraphExec_t exec; };
	std::vector<Shape> shapes; int cur = -1;
	spb6_batch_tune_t tune{};
	double graph_t_end = 1.0;   // t_end held by the captured graphs' kernel arguments
	std::string err;
};

namespace {
int bfail(spb6_batch* b, const std::string& what) { if (b) b->err = what; return SPB6_HIP_ERROR; }
#define BCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return bfail(b, std::string(#x) + ": " + hipGetErrorString(e_)); } while (0)
int balloc(spb6_batch* b, double** p, long count, long front = 0) {
	void* q = nullptr;
	BCHK(hipMalloc(&q, (size_t)(count + front) * sizeof(double)));
	b->allocs.push_back(q);
	BCHK(hipMemset(q, 0, (size_t)(count + front) * sizeof(double)));
	*p = (double*)q + front;
	return SPB6_OK;
}
template <int T>
void launch_stage(const BArgs& A, int blocks, int s, hipStream_t st) { hipLaunchKernelGGL(k_batch_stage<T>, dim3(blocks), dim3(T), 0, st, A, s); }
void launch_stage_t(int threads, const BArgs& A, int blocks, int s, hipStream_t st) {
	if (threads == 64) launch_stage<64>(A, blocks, s, st);
	else if (threads == 128) launch_stage<128>(A, blocks, s, st);
	else launch_stage<256>(A, blocks, s, st);
}
// the two launches of a step, captured once per launch shape
int shape_graph(spb6_batch* b, int threads, int per_cu, hipGraphExec_t& exec) {
	hipGraph_t gr = nullptr;
	BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
	const int blocks = std::max(1, std::min(b->cus * per_cu, (int)b->hprob.size()));
	launch_stage_t(threads, b->A, blocks, 0, b->st);
	launch_stage_t(threads, b->A, blocks, 1, b->st);
	BCHK(hipStreamEndCapture(b->st, &gr));
	BCHK(hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0));
	BCHK(hipGraphDestroy(gr));
	return SPB6_OK;
}
int use_shape(spb6_batch* b, int threads, int per_cu) {
	b->tune.threads = threads; b->tune.per_cu = per_cu;
	for (size_t k = 0; k < b->shapes.size(); k++)
		if (b->shapes[k].threads == threads && b->shapes[k].per_cu == per_cu) { b->cur = (int)k; return SPB6_OK; }
	hipGraphExec_t ex = nullptr;
	int rc = shape_graph(b, threads, per_cu, ex); if (rc) return rc;
	b->shapes.push_back({threads, per_cu, ex});
	b->cur = (int)b->shapes.size() - 1;
	b->tune.threads = threads; b->tune.per_cu = per_cu;
	return SPB6_OK;
}
std::vector<BState> read_states(spb6_batch* b) {
	std::vector<BState> s(b->hprob.size());
	(void)hipMemcpy(s.data(), b->A.st, s.size() * sizeof(BState), hipMemcpyDeviceToHost);
	return s;
}
// replay the current shape's step graph until every problem has finished (step target, t_end or failure)
int run_steps(spb6_batch* b, double t_end, long steps) {
	std::vector<BState> s = read_states(b);
	int fin = 0;
	for (auto& x : s) {
		x.target = x.steps + steps;
		const bool over = !(x.t < t_end * (1 - 1e-12));
		x.done = (steps <= 0 || over) ? 1 : 0;
		if (x.done || x.status != SPB6_OK) fin++;
	}
	BCHK(hipMemcpy(b->A.st, s.data(), s.size() * sizeof(BState), hipMemcpyHostToDevice));
	BCHK(hipMemcpy(b->A.nfin, &fin, sizeof(int), hipMemcpyHostToDevice));
	b->A.t_end = t_end;
	// the captured graphs hold the kernel arguments, t_end included: rebuild them when t_end changes
	if (!b->shapes.empty() && b->graph_t_end != t_end) {
		const int th = b->shapes[b->cur].threads, pc = b->shapes[b->cur].per_cu;
		for (auto& sh : b->shapes) (void)hipGraphExecDestroy(sh.exec);
		b->shapes.clear();
		b->graph_t_end = t_end;
		int rc = use_shape(b, th, pc); if (rc) return rc;
	}
	b->graph_t_end = t_end;
	const int np = (int)b->hprob.size();
	const int per_check = 4;
	long replays = 0;
	std::vector<std::pair<double, int>> key(np);
	while (fin < np) {
		for (int k = 0; k < per_check; k++) BCHK(hipGraphLaunch(b->shapes[b->cur].exec, b->st));
		replays += per_check;
		BCHK(hipMemcpyAsync(b->hnfin, b->A.nfin, sizeof(int), hipMemcpyDeviceToHost, b->st));
		BCHK(hipStreamSynchronize(b->st));
		fin = *b->hnfin;
		if (fin >= np) break;
		// longest-processing-time-first on the measured cost of each problem's last step (the queue order is read by
		// the captured launches from device memory, so no re-capture is needed)
		s = read_states(b);
		for (int p = 0; p < np; p++) key[p] = {(s[p].done || s[p].status != SPB6_OK) ? -1.0 : (s[p].cost > 0 ? s[p].cost : (double)b->hprob[p].L), p};
		std::stable_sort(key.begin(), key.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& c) { return a.first > c.first; });
		for (int p = 0; p < np; p++) b->horder[p] = key[p].second;
		BCHK(hipMemcpyAsync((void*)b->A.order, b->horder.data(), np * sizeof(int), hipMemcpyHostToDevice, b->st));
		if (replays > 64 * (steps + 64)) return bfail(b, "batch run did not finish");   // every step takes at most 30 attempts
	}
	return SPB6_OK;
}
}   // namespace

extern "C" {

const char* spb6_batch_last_error(spb6_batch* b) { return b ? b->err.c_str() : "null batch"; }

spb6_batch* spb6_batch_create(const spb6_batch_problem_t* probs, int nprob, const spb6_config_t* cfg, int device) {
	if (!probs || nprob < 1 || !cfg) return nullptr;
	// the batch runs the reference scheme (PC) with full Newton at relaxation level 0 and the reference's profile
	// arithmetic: no jreuse, variant, relaxation, fused shock integration or Mie-Grueneisen phase
	if (cfg->rus_local != 1 || cfg->rk < 1 || cfg->rk > 2 || cfg->jreuse || cfg->variant || cfg->relax || cfg->shock_fused ||
	    cfg->eos_kind[0] || cfg->eos_kind[1] || cfg->mu_rate > 0) return nullptr;
	spb6_batch* b = new spb6_batch();
	b->device = device; b->cfg = *cfg;
	auto fail = [&]() -> spb6_batch* { spb6_batch_destroy(b); return nullptr; };
	if (hipSetDevice(device) != hipSuccess) return fail();
	if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) return fail();
	(void)hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device);
	long off = 0, poff = 0; int Lmax = 0;
	b->hprob.resize(nprob);
	for (int p = 0; p < nprob; p++) {
		const spb6_batch_problem_t& q = probs[p];
		if (q.nx < 2 || q.bc[0] == SPB6_BC_PERIODIC || q.bc[1] == SPB6_BC_PERIODIC || !(q.x1 > q.x0)) return fail();
		BProb& B = b->hprob[p];
		B.L = q.nx; B.off = (int)off; B.poff = (int)poff;
		B.bc[0] = q.bc[0]; B.bc[1] = q.bc[1]; B.bcu[0] = q.bcu[0]; B.bcu[1] = q.bcu[1];
		B.dx = (q.x1 - q.x0) / q.nx;
		B.ph = Phys{};
		for (int k = 0; k < 2; k++) { B.ph.g[k] = q.ph[k].gamma; B.ph.pi[k] = q.ph[k].pi_inf; B.ph.q[k] = q.ph[k].q; B.ph.b[k] = q.ph[k].b; B.ph.cv[k] = q.ph[k].cv; B.ph.qp[k] = q.ph[k].qp; }
		off += q.nx + 2 * GH; poff += q.nx; Lmax = std::max(Lmax, q.nx);
	}
	if (off >= (1L << 31) / NV) return fail();
	b->NT = off; b->NI = poff; b->pad = (long)GH * (Lmax + 2 * GH);
	// longest problems first: the persistent workgroups then finish together
	b->horder.resize(nprob);
	std::iota(b->horder.begin(), b->horder.end(), 0);
	std::stable_sort(b->horder.begin(), b->horder.end(), [&](int a, int c) { return b->hprob[a].L > b->hprob[c].L; });
	BArgs& A = b->A;
	A.np = nprob; A.NT = (int)b->NT;
	auto grid = [&](double** p, int fields) { return balloc(b, p, fields * b->NT, b->pad); };
	if (grid(&A.U0, NV) || grid(&A.U1, NV) || grid(&A.H0, NV) || grid(&A.H1, NV) || grid(&A.Qe, NV) || grid(&A.Qs, NV) ||
	    grid(&A.wxL, NV) || grid(&A.wxR, NV)) return fail();
	// donor pairs: 4 fields x 2 per cell; a view's shift of sh cells moves pair addresses by 2 sh
	if (balloc(b, &A.donx, 8L * b->NT, 2L * b->pad)) return fail();
	double** arrs[] = {&A.vsc, &A.esc, &A.phx, &A.a2s, &A.ubx, &A.ust, &A.vst, &A.cz, &A.Kw, &A.Pold, &A.ubsx, &A.kpx, &A.wLx, &A.pvx};
	for (auto p : arrs) if (grid(p, 1)) return fail();
	if (grid(&A.Qph, 2)) return fail();
	double** vecs[] = {&A.P0, &A.P1, &A.Pn, &A.G, &A.Gn, &A.dP, &A.sc, &A.rhs, &A.cpw, &A.dpw};
	for (auto p : vecs) if (balloc(b, p, b->NI)) return fail();
	if (balloc(b, &A.J, 3 * b->NI)) return fail();
	{ void* sl = nullptr; if (hipMalloc(&sl, b->NI * sizeof(int)) != hipSuccess) return fail(); b->allocs.push_back(sl); A.slist = (int*)sl; }
	void* q = nullptr;
	if (hipMalloc(&q, nprob * sizeof(BProb)) != hipSuccess) return fail(); b->allocs.push_back(q); A.prob = (const BProb*)q;
	if (hipMemcpy(q, b->hprob.data(), nprob * sizeof(BProb), hipMemcpyHostToDevice) != hipSuccess) return fail();
	if (hipMalloc(&q, nprob * sizeof(BState)) != hipSuccess) return fail(); b->allocs.push_back(q); A.st = (BState*)q;
	if (hipMemset(q, 0, nprob * sizeof(BState)) != hipSuccess) return fail();
	if (hipMalloc(&q, nprob * sizeof(int)) != hipSuccess) return fail(); b->allocs.push_back(q); A.order = (const int*)q;
	if (hipMemcpy(q, b->horder.data(), nprob * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return fail();
	if (hipMalloc(&q, 4 * sizeof(int)) != hipSuccess) return fail(); b->allocs.push_back(q);
	if (hipMemset(q, 0, 4 * sizeof(int)) != hipSuccess) return fail();
	A.qctr = (int*)q; A.nfin = (int*)q + 2;
	if (hipHostMalloc((void**)&b->hnfin, sizeof(int), hipHostMallocDefault) != hipSuccess) return fail();
	A.limiter = cfg->limiter; A.order_ = cfg->order; A.rk = cfg->rk; A.ramp = cfg->ramp; A.acdt = cfg->acdt; A.newton_iter = cfg->newton_iter;
	A.cfl = cfg->cfl; A.cfl_ac = cfg->cfl_ac; A.dtfix = cfg->dtfix; A.tol = cfg->newton_tol; A.t_end = 1.0;
	const double bb = 1.0 - 1.0 / std::sqrt(2.0), cc = 1.0 / (2.0 * bb);
	A.ai00_rk1 = 1.0; A.ai00 = cfg->rk == 2 ? bb : 1.0; A.ae10 = cc; A.ai10 = 1 - bb; A.ai11 = bb;
	b->graph_t_end = A.t_end;
	if (use_shape(b, 64, 4)) return fail();
	if (hipDeviceSynchronize() != hipSuccess) return fail();
	return b;
}

void spb6_batch_destroy(spb6_batch* b) {
	if (!b) return;
	if (b->st) (void)hipStreamSynchronize(b->st);
	for (auto& s : b->shapes) if (s.exec) (void)hipGraphExecDestroy(s.exec);
	for (void* p : b->allocs) (void)hipFree(p);
	if (b->hnfin) (void)hipHostFree(b->hnfin);
	if (b->st) (void)hipStreamDestroy(b->st);
	delete b;
}

long spb6_batch_cells(spb6_batch* b) { return b ? b->NI : -1; }

int spb6_batch_set_state(spb6_batch* b, const double* U) {
	if (!b || !U) return SPB6_BAD_ARG;
	(void)hipSetDevice(b->device);
	const int np = (int)b->hprob.size();
	std::vector<double> h((size_t)NV * b->NT, 0.0), P(b->NI);
	std::vector<BState> s(np);
	for (int p = 0; p < np; p++) {
		const BProb& B = b->hprob[p];
		const double* u = U + (size_t)NV * B.poff;
		double pref = INFINITY;
		auto pbar = [&](const double* Q) {
			const double a1 = Q[IA1];
			return a1 * p_ame(B.ph, 0, a1, Q[IM1], Q[IE1]) + (1 - a1) * p_ame(B.ph, 1, 1 - a1, Q[IM2], Q[IE2]);
		};
		for (int i = 0; i < B.L; i++) {
			for (int k = 0; k < NV; k++) h[(size_t)k * b->NT + B.off + GH + i] = u[(size_t)i * NV + k];
			const double pb = pbar(&u[(size_t)i * NV]); pref = (pb < pref) ? pb : pref;
		}
		if (!std::isfinite(pref)) pref = 0.0;
		for (int i = 0; i < B.L; i++) P[B.poff + i] = pbar(&u[(size_t)i * NV]) - pref;   // PC:1222-1227
		BState& x = s[p];
		std::memset(&x, 0, sizeof x);
		x.P_REF = pref; x.dtprev = -1.0; x.startup = 1; x.status = SPB6_OK;
	}
	BCHK(hipMemcpy(b->A.U0, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
	BCHK(hipMemcpy(b->A.P0, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice));
	BCHK(hipMemcpy(b->A.st, s.data(), s.size() * sizeof(BState), hipMemcpyHostToDevice));
	hipLaunchKernelGGL(k_batch_bc, dim3(np), dim3(64), 0, b->st, b->A, 0);
	BCHK(hipGetLastError());
	BCHK(hipStreamSynchronize(b->st));
	return SPB6_OK;
}

int spb6_batch_get_state(spb6_batch* b, double* U) {
	if (!b || !U) return SPB6_BAD_ARG;
	(void)hipSetDevice(b->device);
	std::vector<BState> s = read_states(b);
	std::vector<double> h0((size_t)NV * b->NT), h1((size_t)NV * b->NT);
	BCHK(hipMemcpy(h0.data(), b->A.U0, h0.size() * sizeof(double), hipMemcpyDeviceToHost));
	BCHK(hipMemcpy(h1.data(), b->A.U1, h1.size() * sizeof(double), hipMemcpyDeviceToHost));
	for (size_t p = 0; p < b->hprob.size(); p++) {
		const BProb& B = b->hprob[p];
		const std::vector<double>& h = s[p].par ? h1 : h0;
		for (int i = 0; i < B.L; i++) for (int k = 0; k < NV; k++) U[((size_t)B.poff + i) * NV + k] = h[(size_t)k * b->NT + B.off + GH + i];
	}
	return SPB6_OK;
}

int spb6_batch_get_pressure(spb6_batch* b, double* P, double* pref) {
	if (!b || !P) return SPB6_BAD_ARG;
	(void)hipSetDevice(b->device);
	std::vector<BState> s = read_states(b);
	std::vector<double> p0(b->NI), p1(b->NI);
	BCHK(hipMemcpy(p0.data(), b->A.P0, p0.size() * sizeof(double), hipMemcpyDeviceToHost));
	BCHK(hipMemcpy(p1.data(), b->A.P1, p1.size() * sizeof(double), hipMemcpyDeviceToHost));
	for (size_t p = 0; p < b->hprob.size(); p++) {
		const BProb& B = b->hprob[p];
		const std::vector<double>& h = s[p].ppar ? p1 : p0;
		for (int i = 0; i < B.L; i++) P[B.poff + i] = h[B.poff + i];
		if (pref) pref[p] = s[p].P_REF;
	}
	return SPB6_OK;
}

int spb6_batch_run(spb6_batch* b, double t_end, long steps) {
	if (!b || !(t_end > 0)) return SPB6_BAD_ARG;
	(void)hipSetDevice(b->device);
	int rc = run_steps(b, t_end, steps); if (rc) return rc;
	for (const BState& x : read_states(b)) if (x.status != SPB6_OK) return SPB6_STEP_FAILED;
	return SPB6_OK;
}

int spb6_batch_get_stats(spb6_batch* b, spb6_stats_t* per_problem, spb6_stats_t* total) {
	if (!b) return SPB6_BAD_ARG;
	(void)hipSetDevice(b->device);
	std::vector<BState> s = read_states(b);
	spb6_stats_t tot; std::memset(&tot, 0, sizeof tot);
	for (size_t p = 0; p < s.size(); p++) {
		spb6_stats_t o; std::memset(&o, 0, sizeof o);
		const BState& x = s[p];
		o.steps = x.steps; o.rejects = x.rejects; o.stages = x.stages; o.newton_it = x.newton_it; o.linsolves = x.linsolves;
		o.floor_acc = x.floor_acc; o.rk1 = x.rk1; o.maxres = x.maxres; o.dt_last = x.dt_last; o.t = x.t; o.status = x.status;
		if (per_problem) per_problem[p] = o;
		tot.steps += o.steps; tot.rejects += o.rejects; tot.stages += o.stages; tot.newton_it += o.newton_it; tot.linsolves += o.linsolves;
		tot.floor_acc += o.floor_acc; tot.rk1 += o.rk1; tot.maxres = std::max(tot.maxres, o.maxres);
		if (o.status != SPB6_OK && tot.status == SPB6_OK) tot.status = o.status;
	}
	if (total) *total = tot;
	return SPB6_OK;
}

int spb6_batch_set_launch(spb6_batch* b, int threads, int per_cu) {
	if (!b || !(threads == 64 || threads == 128 || threads == 256) || per_cu < 1 || per_cu > 16) return SPB6_BAD_ARG;
	(void)hipSetDevice(b->device);
	return use_shape(b, threads, per_cu);
}

int spb6_batch_autotune(spb6_batch* b, long steps, spb6_batch_tune_t* out) {
	if (!b || steps < 1) return SPB6_BAD_ARG;
	(void)hipSetDevice(b->device);
	BArgs& A = b->A;
	const size_t gbytes = (size_t)NV * b->NT * sizeof(double), pbytes = (size_t)b->NI * sizeof(double);
	const int np = (int)b->hprob.size();
	// snapshot of everything a step reads at its start
	std::vector<double*> snap;
	auto save = [&](double* src, size_t bytes) -> int {
		void* q = nullptr; BCHK(hipMalloc(&q, bytes)); snap.push_back((double*)q);
		BCHK(hipMemcpy(q, src, bytes, hipMemcpyDeviceToDevice)); return SPB6_OK;
	};
	int rc = save(A.U0, gbytes) || save(A.U1, gbytes) || save(A.P0, pbytes) || save(A.P1, pbytes) || save((double*)A.st, np * sizeof(BState));
	auto release = [&]() { for (double* q : snap) (void)hipFree(q); };
	if (rc) { release(); return rc; }
	auto restore = [&]() -> int {
		BCHK(hipMemcpy(A.U0, snap[0], gbytes, hipMemcpyDeviceToDevice)); BCHK(hipMemcpy(A.U1, snap[1], gbytes, hipMemcpyDeviceToDevice));
		BCHK(hipMemcpy(A.P0, snap[2], pbytes, hipMemcpyDeviceToDevice)); BCHK(hipMemcpy(A.P1, snap[3], pbytes, hipMemcpyDeviceToDevice));
		BCHK(hipMemcpy(A.st, snap[4], np * sizeof(BState), hipMemcpyDeviceToDevice)); return SPB6_OK;
	};
	const int threads[3] = {64, 128, 256}, per_cu[4] = {1, 2, 4, 8};
	double best = INFINITY; int bt = 64, bp = 4, k = 0;
	const double t_end = b->graph_t_end;
	for (int ti = 0; ti < 3; ti++) for (int pi = 0; pi < 4; pi++) {
		if (threads[ti] * per_cu[pi] > 1024) continue;
		rc = restore(); if (rc) { release(); return rc; }
		rc = use_shape(b, threads[ti], per_cu[pi]); if (rc) { release(); return rc; }
		BCHK(hipDeviceSynchronize());
		const auto t0 = std::chrono::steady_clock::now();
		rc = run_steps(b, t_end, steps); if (rc) { release(); return rc; }
		const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		if (k < 16) { b->tune.tried_threads[k] = threads[ti]; b->tune.tried_per_cu[k] = per_cu[pi]; b->tune.tried_ms[k] = ms; k++; }
		if (ms < best) { best = ms; bt = threads[ti]; bp = per_cu[pi]; }
	}
	b->tune.ntried = k;
	rc = restore(); release(); if (rc) return rc;
	rc = use_shape(b, bt, bp); if (rc) return rc;
	b->tune.best_ms = best;
	if (out) *out = b->tune;
	return SPB6_OK;
}

}   // extern "C"
