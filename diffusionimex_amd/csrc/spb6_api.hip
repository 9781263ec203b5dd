// spb6_api.hip -- host driver and C ABI of the gfx950 SPB 6+1 IMEX step (include/spb6_imex.h).
//
// The driver follows the reference's control flow exactly (time loop PC:1230-1296, attempt loop PC:1243-1289,
// PressureSubsystem PC:851-993, Newton_Solver PC:761-847 with PCO:799-904 at jreuse=0, GMRES_Solver = BiCGStab PC:725-758,
// Thomas PC:807-812); every data-parallel loop runs as a HIP kernel on the context's stream, and the host only takes
// the scalar decisions the reference takes (norms, step criterion, line-search acceptance, reject-and-halve).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>
#include "spb6_kernels.h"
#include "spb6_mlp.h"
#include "../../include/spb6_imex.h"

using namespace spb;

static thread_local std::string g_err;
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { g_err = std::string(#x) + ": " + hipGetErrorString(e_); return SPB6_HIP_ERROR; } } while (0)

namespace {

// the scalars the host reads back: norms (8 slots of 64 striped maxima), flags, the Krylov loop state and the FP32
// form's defect sums and step criterion. The device keeps the same struct in one allocation, so every read-back is one copy.
struct HostScratch { unsigned long long nrm[8 * 64]; int flags[NFLAGS]; KS ks; double dsum[3]; int spec; double gthr; int hxf; };

// T_KRYLOV: BiCGStab iteration passes only; T_KAUX: the solves' setup, precision conversions and FP64 defects;
// T_GUESS: the learned guess's features and inference
// T_KIDLE: BiCGStab iterations launched past a solve's end (in a chunk), which return at once, and speculative trial
// residuals whose gate was off (they return at once too)
enum TimerCat { T_EXPLICIT = 0, T_PREP, T_SHOCK, T_RESIDUAL, T_JACOBIAN, T_KRYLOV, T_FINAL, T_OTHER, T_KAUX, T_GUESS, T_KIDLE, T_NCAT };

struct Timer {
	bool on = false;
	unsigned mask = ~0u;   // categories timed while on (bit k: category k); the others record no events
	bool want(int cat) const { return on && ((mask >> cat) & 1u); }
	bool hold = false;   // events are held outside pend (the FP32 solve's per-iteration pairs): no resolve meanwhile
	std::vector<hipEvent_t> pool;
	struct Pend { int cat; hipEvent_t a, b; };
	std::vector<Pend> pend;
	double ms[T_NCAT] = {0}; long cnt[T_NCAT] = {0};
	size_t next = 0;
	hipEvent_t get() {
		if (next >= pool.size()) { hipEvent_t e; hipEventCreate(&e); pool.push_back(e); }
		return pool[next++];
	}
	void resolve() {
		if (pend.empty()) return;
		hipEventSynchronize(pend.back().b);
		for (auto& p : pend) { float t = 0; hipEventElapsedTime(&t, p.a, p.b); ms[p.cat] += t; cnt[p.cat]++; }
		pend.clear(); next = 0;
	}
	~Timer() { for (auto e : pool) hipEventDestroy(e); }
};

}   // namespace

struct spb6_ctx {
	spb6_config_t cfg;
	int device = 0;
	hipStream_t st = nullptr;
	Geo g; Phys ph;
	// ghosted grids (8 fields each) and arrays
	double *U = nullptr, *Ust = nullptr, *Qe = nullptr, *Qs = nullptr, *H[3] = {nullptr, nullptr, nullptr};   // H[2]: rk = 3 only
	const double *qe_cur = nullptr, *qs_cur = nullptr;   // the stage operands of the per-stage entry points
	double *donx, *dony, *phx, *phy, *a2s, *ubx, *uby, *vsc, *esc, *vu, *vv;   // face donors: (left, right) pairs of m1, m2, e1, e2
	const double* qe_last = nullptr;   // Qe of the last explicit stage
	bool final_pending = false;        // the previous stage's final-state flag (FLAG_FINAL_PREV) is not read back yet
	double *ust, *vst, *cz, *Kw, *Pold, *ubsx, *kpx, *wLx, *pvx, *ubsy, *kpy, *wLy, *pvy, *Qph;
	// pressure system (interior)
	double *P, *Pst, *Pn, *G, *Gn, *dP, *rhs, *sc, *J, *cpw, *dpw;
	double *kr, *kr0, *kp, *kv, *ks_, *kt;
	// FP32 Krylov form: matrix, right-hand side and vectors in single precision
	float *J32 = nullptr, *b32 = nullptr, *x32 = nullptr, *r32 = nullptr, *r032 = nullptr, *p32 = nullptr, *v32 = nullptr, *s32 = nullptr, *t32 = nullptr;
	float *ph32 = nullptr, *sh32 = nullptr, *dg32 = nullptr, *pp32 = nullptr;   // red-black preconditioner vectors of the FP32 solve
	double* part = nullptr; int part_stride = 0; KS* ksd = nullptr;
	double *part2 = nullptr, *dsum = nullptr;   // the FP32 form's defect partials and their two sums
	HostScratch* mail = nullptr;               // device side of hs: nrm, flags, ksd and dsum point into it
	unsigned long long* nrm = nullptr; int* flags = nullptr;
	HostScratch* hs = nullptr;
	std::vector<void*> allocs;
	// driver state
	double P_REF = 0.0, t = 0.0, dtprev = -1.0; bool startup = true;
	double Ae[3][3], Ai[3][3];
	spb6_stats_t stats;
	int last_krylov = 0;
	long krylov_launched = 0;          // BiCGStab iteration passes launched since timing was (re)enabled
	double tau_last = 0.0;
	// learned guess back-off: after two stages in a row where the plain guess judged the network's guess worse, the
	// network is skipped for 1, 2, 4, ... 32 stages (guess_skip) before it is tried again; a win resets the back-off
	int guess_skip = 0, guess_losses = 0, guess_backoff = 0;
	Timer timer;
	// learned guess and training dumps
	spb6_mlp* model = nullptr;
	float* F = nullptr;                 // 7 per-cell feature fields over the ghosted grid
	double* Pplain = nullptr;           // the plain guess of the current stage
	double* Gplain = nullptr;           // its residual
	int *slist = nullptr, *scount = nullptr;   // strong-shock cells of the shock budget's second pass (scount[1]: its cursor)
	double* srec = nullptr;                    // their profile problems (SHOCK_REC x n)
	unsigned long long* shock_steps = nullptr; // RK45 steps of the strong-shock passes since timing was (re)enabled
	unsigned long long* tmax = nullptr;        // running maxima of the phase temperatures (2 x 64 striped slots)
	double M0[3] = {0, 0, 0};                  // conserved totals of the state set last (PC:1166)
	int *rlist = nullptr, *rcount = nullptr;   // cells whose residual the learned guess can change
	int cus = 256;                      // compute units of the device
	int shock_blocks = 1024;            // resident 256-thread blocks of the strong-shock pass
	float *dX = nullptr, *dY = nullptr; unsigned long long* dcount = nullptr; long dcap = 0; int dstride = 0; unsigned dseq = 0;
	// the explicit HLLC scheme (allocated at its first call): cell primitives and the x / y face values
	double *hxp = nullptr, *hfx = nullptr, *hfy = nullptr;
	// an IMEX variant's arrays (cfg.variant != 0): the full face states, the TV-type variants' eTV_k and ETV, BB's
	// alpha_1 step, TV's alpha_1 face donors
	double *fwxL = nullptr, *fwxR = nullptr, *fwyL = nullptr, *fwyR = nullptr;
	double *eTV1 = nullptr, *eTV2 = nullptr, *ETV = nullptr, *dA = nullptr, *donax = nullptr, *donay = nullptr;
	int* ctype = nullptr;   // relaxation levels 1-3: the closure type of each interior cell for the step (STEP_TYPE)
	double eps_i = 1e-4;    // EPS_I (PHY:38)
	RelaxDev rel{nullptr, nullptr, 0, 0.0, 0.0, 0.0, 0.0, INFINITY, 0, 1.0, 22.064e6, nullptr, 0};   // the saturation table on the device (levels 2-3) and the rates
	double* Z = nullptr;    // finite mu: the stage's impedances of the CL_MU cells (2 NC)
	// graph mode (cfg.graph): a stage's fixed head (combination, explicit part, preparation, shock budget) and each
	// Krylov chunk (its BiCGStab passes, the FP32 form's defect, sweep close and trial residual) run as captured graphs,
	// re-captured per launch and applied to the instantiated graph of the same shape with hipGraphExecUpdate
	bool capturing = false;
	hipGraphExec_t g_head[3] = {nullptr, nullptr, nullptr};
	hipGraphExec_t g_chunk[2][33] = {};   // [trial residual behind the chunk][chunk length]
	long graph_launches = 0, graph_rebuilds = 0;
};

struct spb6_mlp {
	int device = 0; hipStream_t st = nullptr;
	float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr, *mean = nullptr, *istd = nullptr, *loss = nullptr;
	float *mub1 = nullptr;                      // per-field input shift and folded layer-1 bias of the BF16 inference
	__bf16 *W1s = nullptr, *W1f = nullptr, *W2p = nullptr;
	float mu_t = 0.f, sd_t = 1.f, b3 = 0.f;
	float *X = nullptr, *Y = nullptr; long ndata = 0;
	float *Xb = nullptr, *Yb = nullptr, *ws = nullptr, *yb = nullptr; int* idx = nullptr; int bcap = 0;
	int step = 0;
	std::vector<void*> allocs;
};

namespace {

int dalloc(spb6_ctx* c, double** p, size_t count) {
	HIPCHK(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(double)));
	c->allocs.push_back(*p);
	HIPCHK(hipMemset(*p, 0, std::max<size_t>(count, 1) * sizeof(double)));
	return SPB6_OK;
}

Dev make_dev(const spb6_ctx* c) {
	Dev D;
	D.g = c->g; D.ph = c->ph; D.P_REF = c->P_REF; D.rx = 0; D.ry = 0; D.tau = 0;
	D.Ust_c = c->Ust; D.Ust = c->Ust; D.a2s = c->a2s; D.ubx = c->ubx; D.uby = c->uby; D.phx = c->phx; D.phy = c->phy;
	D.vsc = c->vsc; D.esc = c->esc; D.vu = c->vu; D.vv = c->vv; D.wxL = D.wxR = D.wyL = D.wyR = nullptr; D.donx = c->donx; D.dony = c->dony;
	D.ust = c->ust; D.vst = c->vst; D.cz = c->cz; D.Kw = c->Kw; D.Pold = c->Pold;
	D.ubsx = c->ubsx; D.kpx = c->kpx; D.wLx = c->wLx; D.pvx = c->pvx; D.ubsy = c->ubsy; D.kpy = c->kpy; D.wLy = c->wLy; D.pvy = c->pvy;
	D.Qph = c->Qph; D.flags = c->flags;
	D.variant = c->cfg.variant;
	D.eTV1 = c->eTV1; D.eTV2 = c->eTV2; D.ETV = c->ETV; D.dA = c->dA; D.donax = c->donax; D.donay = c->donay;
	if (D.variant != VAR_PC) { D.wxL = c->fwxL; D.wxR = c->fwxR; D.wyL = c->fwyL; D.wyR = c->fwyR; }
	D.ctype = c->ctype;
	D.rel = c->rel;
	D.Z = c->Z;
	return D;
}
// the saturation table T_sat(ln p) (PHY:125-152), built on the host with the device's portable exp, log and pow: 1600
// points in ln p over [ln 0.1, ln 5e9]; the first root by a scan of T in [20, 5000] K and bisection, each next one by
// Newton from the previous; it ends where no root is found or where the liquid's volume reaches the vapour's (the
// critical point, refined by bisection in p)
void build_sat(const Phys& P, std::vector<double>& lpv, std::vector<double>& Tv, double& pcrit, double& vcrit) {
	lpv.clear(); Tv.clear(); pcrit = INFINITY; vcrit = 0.0;
	const int NP = 1600; const double l0 = plog(1e-1), l1 = plog(5e9);
	double Tprev = -1.0;
	for (int n = 0; n < NP; n++) {
		const double lp = l0 + (l1 - l0) * n / (NP - 1), p = pexp(lp);
		double Troot = -1.0;
		if (Tprev < 0) {
			double Ta = 20.0, fa = sat_f(P, p, Ta);
			for (int m = 1; m <= 4000; m++) {
				const double Tb = 20.0 * ppow(250.0, m / 4000.0), fb = sat_f(P, p, Tb);
				if (fa < 0 && fb >= 0) {
					double a = Ta, b = Tb;
					for (int it = 0; it < 200; it++) { const double cm = 0.5 * (a + b); if (sat_f(P, p, cm) < 0) a = cm; else b = cm; }
					Troot = 0.5 * (a + b); break;
				}
				Ta = Tb; fa = fb;
			}
		} else { bool ok; const double T = sat_newton(P, p, Tprev, ok); if (ok) Troot = T; }
		if (Troot < 0) { if (n > 0) pcrit = pexp(lpv.back()); break; }
		if (!(v_pT(P, 1, p, Troot) > v_pT(P, 0, p, Troot))) { pcrit = p; vcrit = v_pT(P, 0, p, Troot); break; }
		lpv.push_back(lp); Tv.push_back(Troot); Tprev = Troot;
	}
	if (vcrit == 0.0 && !Tv.empty()) vcrit = v_pT(P, 0, pexp(lpv.back()), Tv.back());
	if (std::isfinite(pcrit) && !lpv.empty()) {
		double a = pexp(lpv.back()), b = pcrit, Ta = Tv.back();
		for (int it = 0; it < 100; it++) {
			const double cm = std::sqrt(a * b); bool ok; const double T = sat_newton(P, cm, Ta, ok);
			if (ok && v_pT(P, 1, cm, T) > v_pT(P, 0, cm, T)) { a = cm; Ta = T; } else b = cm;
		}
		pcrit = a; vcrit = v_pT(P, 0, a, Ta);
	}
}

// timed launch region
struct TRegion {
	spb6_ctx* c; int cat; hipEvent_t a = nullptr;
	TRegion(spb6_ctx* c_, int cat_) : c(c_), cat(cat_) { if (c->timer.want(cat) && !c->capturing) { a = c->timer.get(); hipEventRecord(a, c->st); } }
	~TRegion() { if (a) { hipEvent_t b = c->timer.get(); hipEventRecord(b, c->st); c->timer.pend.push_back({cat, a, b}); if (c->timer.pend.size() > 512 && !c->timer.hold) c->timer.resolve(); } }
};

// clear the norms and the flags of one evaluation (the stage's admissibility flags stay)
int reset_scalars(spb6_ctx* c) {
	launch_zero_ints(reinterpret_cast<int*>(c->mail), (long)((offsetof(HostScratch, flags) + FLAG_RHO * sizeof(int)) / sizeof(int)), c->st);
	return SPB6_OK;
}
// clear the norms and the flags (the start of a stage); keep_prev keeps the previous stage's final-state flag
int reset_stage(spb6_ctx* c, bool keep_prev = false) {
	const size_t end = keep_prev ? offsetof(HostScratch, flags) + FLAG_FINAL_PREV * sizeof(int) : offsetof(HostScratch, ks);
	launch_zero_ints(reinterpret_cast<int*>(c->mail), (long)(end / sizeof(int)), c->st);   // nrm and flags (contiguous)
	return SPB6_OK;
}
// copy norms, flags and Krylov scalars to pinned host memory and wait
int fetch(spb6_ctx* c, bool with_ks = false) {
	(void)with_ks;
	HIPCHK(hipMemcpyAsync(c->hs, c->mail, sizeof(HostScratch), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return SPB6_OK;
}
// maximum of the 64 striped slots of norm k (non-negative doubles compare like their bit patterns)
inline double nrm_val(const spb6_ctx* c, int k) {
	unsigned long long m = 0;
	for (int q = 0; q < 64; q++) m = std::max(m, c->hs->nrm[64 * k + q]);
	double v; std::memcpy(&v, &m, sizeof v); return v;
}

// graph mode: the launches of f captured into a graph and run through the instantiated graph in slot (updated in place
// when the shape matches, instantiated anew when it does not); without graph mode f launches directly
template <class F>
int run_graph(spb6_ctx* c, hipGraphExec_t& slot, F f) {
	if (!c->cfg.graph) { f(); return SPB6_OK; }
	HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
	c->capturing = true;
	f();
	c->capturing = false;
	hipGraph_t gr = nullptr;
	HIPCHK(hipStreamEndCapture(c->st, &gr));
	if (slot) {
		hipGraphNode_t bad = nullptr; hipGraphExecUpdateResult res;
		if (hipGraphExecUpdate(slot, gr, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
			(void)hipGetLastError();
			(void)hipGraphExecDestroy(slot); slot = nullptr;
		}
	}
	if (!slot) { c->graph_rebuilds++; HIPCHK(hipGraphInstantiate(&slot, gr, nullptr, nullptr, 0)); }
	HIPCHK(hipGraphDestroy(gr));
	HIPCHK(hipGraphLaunch(slot, c->st));
	c->graph_launches++;
	return SPB6_OK;
}

// BiCGStab (PC:725-758) on the Jacobian c->J with right-hand side c->rhs; solution in c->dP. With krylov_fp32 the
// iteration runs on FP32 copies of the matrix and vectors (inner tolerance at least 1e-6 |b|, the FP32 floor) and the
// step is widened back to FP64; the Newton residual stays FP64 and corrects the inexact step (defect correction).
template <typename T>
int krylov_run(spb6_ctx* c, const KrylovBufsT<T>& B, const T* b, double tol, int& iters, bool& ok) {
	const Geo& g = c->g; const int maxit = c->cfg.krylov_iter;
	{
		TRegion tr(c, T_KAUX);
		if constexpr (sizeof(T) == 8) launch_krylov_setup(g, B, b, c->st); else launch_krylov_setup32(g, B, b, c->st);
	}
	// the loop top of iteration `launched` runs in its first pass, so maxit + 1 passes decide every outcome
	int chunk = std::max(4, c->last_krylov + 1), launched = 0;
	for (;;) {
		{
			TRegion tr(c, T_KRYLOV);
			const int nk = std::min(chunk, maxit + 1 - launched), l0 = launched;
			int rc = run_graph(c, c->g_chunk[0][std::min(nk, 32)], [&] {
				for (int k = 0; k < nk; k++) {
					if constexpr (sizeof(T) == 8) launch_krylov_iter(g, B, c->st, l0 + k, tol, maxit);
					else launch_krylov_iter32(g, B, c->st, l0 + k, tol, maxit);
				}
			});
			if (rc) return rc;
			launched += nk;
		}
		HIPCHK(hipMemcpyAsync(&c->hs->ks, c->ksd, sizeof(KS), hipMemcpyDeviceToHost, c->st));
		HIPCHK(hipStreamSynchronize(c->st));
		if (c->hs->ks.done || launched > maxit) break;
		chunk = 4;
	}
	iters = c->hs->ks.iters; ok = c->hs->ks.done && c->hs->ks.status == 0;
	c->last_krylov = iters;
	c->krylov_launched += launched;
	return SPB6_OK;
}
// FP32 Krylov form: inexact Newton step. The step solves J dP = b to a relative FP64 residual (forcing term)
// eta = max(0.1 tol / res, min(1e-6, res)) (res = the Newton residual norm, tol = the Newton tolerance). The next
// Newton residual is about C res^2 + eta res, so a step solved beyond eta ~ res buys nothing (Eisenstat and Walker's
// forcing terms with exponent 2), and eta = 0.1 tol / res already finishes Newton in this step; the cap 1e-6 keeps the
// smooth-mode error of the early steps small enough that P does not move across the upwind kink (u_f)^+ / (u_f)^- at
// u_f = 0 in quiescent cells (looser forcing, ~1e-3, leaves ~1e-10 errors there and Newton then converges only
// linearly). The solve runs as sweeps of FP32 BiCGStab on the FP64 defect r = b - J dP (iterative refinement), each
// sweep asking its FP32 solve for max(eta / 2, 5e-7) relative accuracy (the FP32 floor), so that one sweep meets
// eta >= 1e-6. A step whose defect misses eta but has decreased is still returned: Newton's residual test and line
// search judge it.
// Dspec (FP32 form): launch the Newton step's first trial residual (P + dP at lam 1, into Pn / Gn) behind each chunk,
// gated on the device's close of the sweep; trial is set when it ran and the read-back holds its norm and flag
int krylov(spb6_ctx* c, int& iters, bool& ok, double newton_res, bool new_J, bool& crit, double& sn, bool finish,
           const Dev* Dspec = nullptr, bool* trial = nullptr) {
	crit = false; sn = 0.0;
	if (trial) *trial = false;
	const long n = c->g.n;
	if (!c->cfg.krylov_fp32) {
		KrylovBufs B{c->dP, c->kr, c->kr0, c->kp, c->kv, c->ks_, c->kt, c->J, c->part, c->part_stride, c->ksd};
		return krylov_run<double>(c, B, c->rhs, c->cfg.krylov_tol, iters, ok);
	}
	const double res = std::max(newton_res, 1e-300);
	const double eta = finish ? 0.1 * c->cfg.newton_tol / res : std::max(0.1 * c->cfg.newton_tol / res, std::min(1e-6, res));
	KrylovBufs32 B{c->x32, c->r32, c->r032, c->p32, c->v32, c->s32, c->t32, c->J32, c->part, c->part_stride, c->ksd, c->ph32, c->sh32, c->dg32, c->pp32};
	iters = 0; ok = false;
	const double tin = std::max(0.5 * eta, 5e-7);
	const int maxit = c->cfg.krylov_iter;
	const Geo& g = c->g;
	for (int sweep = 0; sweep < 6; sweep++) {
		{
			TRegion tr(c, T_KAUX);
			// the first sweep converts the system in its setup's first pass: b and dP = 0 always, J when it is new (a
			// chord step keeps the previous FP32 copy)
			if (sweep == 0) launch_krylov_setup32_cvt(g, B, new_J ? c->J : nullptr, c->J32, c->rhs, c->b32, c->dP, c->st);
			else launch_krylov_setup32(g, B, c->b32, c->st);
		}
		// the sweep's BiCGStab passes in chunks sized by the previous solve; behind each chunk, in the same stream, the
		// FP64 defect and the step criterion of dP + x, so that one synchronisation reads the solve's state, the defect
		// and the criterion together (a solve that is not done yet continues; dP is only written once it is)
		int chunk = std::max(2, c->last_krylov + 1), launched = 0;
		// with timing on, each launched iteration gets its own event pair; once the solve's iteration count is read,
		// the iterations taken go to T_KRYLOV and those launched past the end to T_KIDLE
		// (and a speculative trial residual is filed once its gate is known: T_RESIDUAL when it ran, T_KIDLE otherwise)
		struct IterEv { int it; hipEvent_t a, b; };
		std::vector<IterEv> iev;
		const bool iter_ev = (c->timer.want(T_KRYLOV) || c->timer.want(T_KIDLE)) && !c->cfg.graph;
		const bool spec_ev = (c->timer.want(T_RESIDUAL) || c->timer.want(T_KIDLE)) && !c->cfg.graph;
		c->timer.hold = c->timer.on;
		for (;;) {
			hipEvent_t sa = nullptr, sb = nullptr;
			const int nk = std::min(chunk, maxit + 1 - launched), l0 = launched;
			int grc = run_graph(c, c->g_chunk[Dspec ? 1 : 0][std::min(nk, 32)], [&] {
				for (int k = 0; k < nk; k++) {
					hipEvent_t a = nullptr, b = nullptr;
					if (iter_ev) { a = c->timer.get(); hipEventRecord(a, c->st); }
					launch_krylov_iter32(g, B, c->st, l0 + k, tin, maxit);
					if (iter_ev) { b = c->timer.get(); hipEventRecord(b, c->st); iev.push_back({l0 + k, a, b}); }
				}
				{
					TRegion tr(c, T_KAUX);
					launch_defect_x(g, c->J, c->rhs, c->dP, c->x32, c->kr, c->part2, c->part_stride, c->dsum, c->Pst, c->sc, c->st);
				}
				if (Dspec) {
					{ TRegion tr(c, T_KAUX); launch_spec_commit(n, c->x32, c->dP, c->ksd, c->dsum, eta, &c->mail->spec, &c->nrm[0], c->flags, c->st); }
					if (spec_ev) { sa = c->timer.get(); hipEventRecord(sa, c->st); }
					launch_residual(*Dspec, c->Pst, c->dP, 1.0, c->Pn, c->Gn, &c->nrm[0], FLAG_RES, c->st, &c->mail->spec);
					if (spec_ev) { sb = c->timer.get(); hipEventRecord(sb, c->st); }
				}
			});
			if (grc) return grc;
			launched += nk;
			HIPCHK(hipMemcpyAsync(c->hs, c->mail, sizeof(HostScratch), hipMemcpyDeviceToHost, c->st));
			HIPCHK(hipStreamSynchronize(c->st));
			if (sa) {
				const int cat = c->hs->spec ? T_RESIDUAL : T_KIDLE;
				if (c->timer.want(cat)) c->timer.pend.push_back({cat, sa, sb});
			}
			if (c->hs->ks.done || launched > maxit) break;
			chunk = 4;
		}
		const int it = c->hs->ks.iters;
		const bool kok = c->hs->ks.done && c->hs->ks.status == 0;
		for (const IterEv& e : iev) {
			const int cat = e.it < it ? T_KRYLOV : T_KIDLE;
			if (c->timer.want(cat)) c->timer.pend.push_back({cat, e.a, e.b});
		}
		c->timer.hold = false;
		c->last_krylov = it; c->krylov_launched += launched; iters += it;
		const bool closed = Dspec && c->hs->ks.done;   // the device has added x and decided the defect test
		if (!closed) { TRegion tr(c, T_KAUX); launch_add_f32(n, c->x32, c->dP, c->st); }
		crit = true; sn = c->hs->dsum[2];   // the step criterion just read is that of dP as it now stands
		const double rr = c->hs->dsum[0], bb = c->hs->dsum[1];
		if (getenv("SPB6_VERBOSE") && atoi(getenv("SPB6_VERBOSE")) >= 3)
			fprintf(stderr, "        sweep %d krylov %d defect %.3e eta %.3e\n", sweep, it, std::sqrt(rr / bb), eta);
		if (closed ? c->hs->spec != 0 : (std::sqrt(rr) <= eta * std::sqrt(bb) || !(bb > 0))) {
			ok = true;
			if (closed && trial) *trial = true;
			break;
		}
		ok = std::isfinite(rr) && rr < bb;   // progress so far
		if (!kok && sweep > 0) break;
		{ TRegion tr(c, T_KAUX); launch_to_f32(n, c->kr, c->b32, c->st); }
	}
	return SPB6_OK;
}

// evalAll + norm (PC:764-770) at P + lam dP (dP may be null); trial pressure to Pout, residual to G
int eval_all(spb6_ctx* c, const Dev& D, const double* P, const double* dP, double lam, double* Pout, double* G, bool& ok, double& res) {
	int rc = reset_scalars(c); if (rc) return rc;
	{
		TRegion tr(c, T_RESIDUAL);
		launch_residual(D, P, dP, lam, Pout, G, &c->nrm[0], FLAG_RES, c->st);
	}
	rc = fetch(c); if (rc) return rc;
	ok = c->hs->flags[FLAG_RES] == 0; res = nrm_val(c, 0);
	return SPB6_OK;
}

// Newton_Solver (PC:761-847): unknown c->Pst; status SPB6_OK or SPB6_NEWTON_FAIL / SPB6_KRYLOV_FAIL
int newton(spb6_ctx* c, const Dev& D, int& iters, double& resfinal, bool have_G = false, bool ok0 = false, double res0 = 0.0) {
	const int n = c->g.n; const double tol = c->cfg.newton_tol;
	const bool tri = (c->g.ny == 1 && c->cfg.bc[0] != SPB6_BC_PERIODIC);
	bool ok = ok0; double res = res0;
	int rc = SPB6_OK;
	if (!have_G) { rc = eval_all(c, D, c->Pst, nullptr, 0.0, nullptr, c->G, ok, res); if (rc) return rc; }
	if (!ok) { iters = 0; resfinal = INFINITY; return SPB6_NEWTON_FAIL; }
	const int verbose = getenv("SPB6_VERBOSE") ? atoi(getenv("SPB6_VERBOSE")) : 0;
	// faster forms: the last step of a solve as a chord step (this solve's previous Jacobian): once the residual is
	// within 100x of the tolerance after a full step that contracted at least 1000x, the Jacobian has moved by about that
	// step and the chord step's extra error is ~res_prev x res. A chord step that does not contract 100x turns chord
	// steps off for the rest of the solve (near the cavitation threshold the Jacobian can move much more than the step
	// suggests; there chord steps from res ~1e-8 were measured to cost Newton iterations).
	// jreuse (modified Newton, PCO:796-888) takes the place of the chord step
	const bool jreuse = c->cfg.jreuse != 0;
	bool chord_ok = c->cfg.analytic_jacobian && c->cfg.krylov_fp32 && !tri && !jreuse;
	bool have_J = false, full_step = false;
	double res_prev = INFINITY;
	bool fresh = false; double ratio = 1.0, lamprev = 1.0;   // jreuse: the last step's contraction and damping (PCO:825)
	for (int it = 0; it < c->cfg.newton_iter; it++) {
		iters = it; resfinal = res;
		if (verbose >= 3) fprintf(stderr, "    newton it %d res %.3e\n", it, res);   // the reference's verbose >= 3 trace (PC:781)
		if (res < tol) return SPB6_OK;
		// the step's first pass clears its norm slots and flags (reset_scalars' bytes)
		int* const clr = reinterpret_cast<int*>(c->mail);
		const int ncl = (int)((offsetof(HostScratch, flags) + FLAG_RHO * sizeof(int)) / sizeof(int));
		const bool chord = chord_ok && have_J && full_step && res < 100.0 * tol && res <= 1e-3 * res_prev;
		// jreuse: rebuild at the stage's first iteration and after a damped step or one that contracted less than 4x
		// (fresh_ok, PCO:797, PCO:828); a reused Jacobian keeps the scale s of its build too
		const bool reuse = jreuse && it > 0 && lamprev == 1.0 && ratio <= 0.25;
		const double res_before = res;
		{
			TRegion tr(c, T_JACOBIAN);
			if (reuse) launch_chord_rhs(D, c->Pst, c->G, nullptr, c->rhs, c->st, clr, ncl);
			else if (chord) launch_chord_rhs(D, c->Pst, c->G, c->sc, c->rhs, c->st, clr, ncl);
			else launch_jacobian(D, c->Pst, c->G, c->sc, c->J, c->rhs, c->cfg.analytic_jacobian != 0, c->st, clr, ncl);
		}
		if (!reuse && !chord) c->stats.jac_builds++;
		fresh = !reuse;
		have_J = true;
		c->stats.linsolves++;
		bool kok = true, crit = false, trial = false; int kit = 0; double snk = 0.0;
		if (tri) {
			TRegion tr(c, T_OTHER);
			launch_thomas(n, c->J, c->rhs, c->dP, c->cpw, c->dpw, c->st);
		} else {
			// a step predicted to be the last (the observed contraction res / res_prev^2 puts the next residual below
			// tol / 2) is solved to 0.1 tol / res, so that Newton finishes with it instead of taking one more step
			const bool finish = chord_ok && full_step && res * res * res < 0.5 * tol * res_prev * res_prev;
			rc = krylov(c, kit, kok, res, !chord && !reuse, crit, snk, finish, &D, &trial); if (rc) return rc;
		}
		if (!crit) {   // the FP32 form reads the step criterion and the flags with its last defect
			launch_stepcrit(n, c->dP, c->Pst, c->sc, &c->nrm[64], c->st);
			rc = fetch(c); if (rc) return rc;
		}
		if (c->hs->flags[FLAG_JAC]) return SPB6_NEWTON_FAIL;   // the reference returns before its linear solve (PC:806)
		c->stats.krylov_it += kit;
		if (!kok) return SPB6_KRYLOV_FAIL;
		const double sn = crit ? snk : nrm_val(c, 1);
		if (verbose >= 3) fprintf(stderr, "      step %.3e krylov %d\n", sn, kit);
		// the first evaluation of both branches below is P + dP at lam 1: with FP32 Krylov it has already run behind the
		// solve's last chunk (trial), and its norm and flag came back with the solve's state
		auto first_eval = [&](bool& okv, double& rv) -> int {
			if (trial) { okv = c->hs->flags[FLAG_RES] == 0; rv = nrm_val(c, 0); trial = false; return SPB6_OK; }
			return eval_all(c, D, c->Pst, c->dP, 1.0, c->Pn, c->Gn, okv, rv);
		};
		if (sn < tol) {   // step criterion (PC:819-826): accept P + dP
			rc = first_eval(ok, res); if (rc) return rc;
			std::swap(c->Pst, c->Pn); std::swap(c->G, c->Gn);
			c->stats.newton_it++; iters = it + 1;
			if (!ok) return SPB6_NEWTON_FAIL;
			resfinal = res; return SPB6_OK;
		}
		double lam = 1.0; bool acc = false;
		for (int ls = 0; ls < 40; ls++) {   // Armijo line search (PC:827-832)
			double rn;
			rc = ls == 0 ? first_eval(ok, rn) : eval_all(c, D, c->Pst, c->dP, lam, c->Pn, c->Gn, ok, rn); if (rc) return rc;
			if (verbose >= 3) fprintf(stderr, "      line search lam %.3e res %.3e ok %d\n", lam, rn, (int)ok);
			if (ok && (rn < res * (1.0 - 1e-4 * lam) || rn < tol)) { acc = true; res = rn; break; }
			lam *= 0.5;
		}
		full_step = acc && lam == 1.0;
		if (chord && !(acc && res <= 1e-2 * res_before)) chord_ok = false;
		res_prev = res_before;
		if (jreuse && !acc && !fresh) { ratio = 1.0; lamprev = 0.0; continue; }   // rebuild, retake the step from P (PCO:833)
		ratio = res / res_before; lamprev = lam;
		if (!acc) {   // round-off floor acceptance (PC:834-842)
			resfinal = res;
			rc = reset_scalars(c); if (rc) return rc;
			launch_floor(D, c->Pst, tol, c->st);
			rc = fetch(c); if (rc) return rc;
			if (c->hs->flags[FLAG_FLOOR]) return SPB6_NEWTON_FAIL;
			c->stats.floor_acc++;
			return SPB6_OK;
		}
		std::swap(c->Pst, c->Pn); std::swap(c->G, c->Gn);
		c->stats.newton_it++;
	}
	iters = c->cfg.newton_iter; resfinal = res;
	return res < tol ? SPB6_OK : SPB6_NEWTON_FAIL;
}

Dev stage_dev(spb6_ctx* c, double adt) {
	Dev D = make_dev(c); D.tau = adt; D.rx = adt / c->g.dx; D.ry = adt / c->g.dy; return D;
}

// explicit part of a stage from (Qe, Qs) into Ust / a2s / face data (TVDReconstruction + Update)
int explicit_part(spb6_ctx* c, const double* Qe, const double* Qs, double adt) {
	Dev D = stage_dev(c, adt);
	TRegion tr(c, T_EXPLICIT);
	c->qe_last = Qe;
	launch_explicit(D, c->cfg.limiter, c->cfg.order, Qe, Qs, adt, c->st);
	return SPB6_OK;
}
int prep_part(spb6_ctx* c, double adt) {
	c->tau_last = adt;
	Dev D = stage_dev(c, adt);
	{ TRegion tr(c, T_PREP); launch_prep(D, c->cfg.limiter, c->cfg.order, adt, c->Pst, c->st); }
	{ TRegion tr(c, T_SHOCK); launch_shock(D, c->slist, c->scount, c->srec, c->shock_steps, c->shock_blocks, c->cfg.shock_fused, c->st); }
	return SPB6_OK;
}
// PressureSubsystem (PC:851-993) after the explicit part; writes the stage state into Ust and H_rks. defer_final
// (a stage with another after it in the same attempt): the final state's closure flag is not waited for here but
// read back with the next stage's first residual, which then reports the failure as this stage would have
int implicit_part(spb6_ctx* c, double adt, const double* Qs, double* Hk, int& iters, double& res, bool defer_final = false, bool prepped = false) {
	int rc = prepped ? SPB6_OK : prep_part(c, adt); if (rc) return rc;
	Dev D = stage_dev(c, adt);
	const long n = c->g.n;
	const bool dumping = c->dX && c->dcap > 0;
	bool learned = c->cfg.learned_guess && c->model;
	if (learned && !dumping && c->guess_skip > 0) {   // backed off: the plain guess, without evaluating the network
		c->guess_skip--; c->stats.guess_skipped++; c->stats.guess_fallback++;
		learned = false;
	}
	// the plain guess's residual (Newton's first, PC:764-770); the stage's admissibility flags, of the explicit part
	// (rho > 0, PC:255-258) and of the preparation (PC:862), are read back with it: one synchronisation instead of
	// three. What runs meanwhile on an inadmissible stage ends normally (every loop in the kernels is bounded) and is
	// discarded with the attempt.
	bool ok0 = false; double res0 = 0.0, resl = 0.0; bool okl = false;
	if (learned) {
		// the plain guess's residual, the learned guess built from it (its activity threshold computed on the device
		// from that residual's norm) and the learned guess's residual (norm slots 128.., FLAG_RES2), all read back in one
		// synchronisation with the stage's flags
		rc = reset_scalars(c); if (rc) return rc;
		// the plain residual and its pressure straight into Gplain / Pplain (G takes that buffer after the read-back); a
		// training dump, which reads Gplain after Newton, keeps copies instead
		if (dumping) {
			{ TRegion tr(c, T_RESIDUAL); launch_residual(D, c->Pst, nullptr, 0.0, nullptr, c->G, &c->nrm[0], FLAG_RES, c->st); }
			HIPCHK(hipMemcpyAsync(c->Pplain, c->Pst, n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
			HIPCHK(hipMemcpyAsync(c->Gplain, c->G, n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
		} else {
			TRegion tr(c, T_RESIDUAL);
			launch_residual(D, c->Pst, nullptr, 0.0, c->Pplain, c->Gplain, &c->nrm[0], FLAG_RES, c->st);
		}
		const spb6_mlp* M = c->model;
		MlpDev W{M->P, M->W1s, M->W1f, M->W2p, M->mub1, M->b3, M->mu_t, M->sd_t};
		{
			TRegion tr(c, T_GUESS);
			if (dumping) mlp_feat_cell(D, c->Pplain, c->F, c->st);   // the dump takes the features of every cell
			// the shock budget's list is free again by now: it carries the active tiles; without a dump the features are
			// computed in mlp_infer, on the listed tiles' stencils only
			mlp_infer(&D, c->F, adt, (int)n, nullptr, W, c->Pplain, c->Gplain, 0.0, c->Pn, nullptr, c->st, c->slist, c->scount,
			          &c->nrm[0], &c->flags[FLAG_RES], &c->mail->gthr, c->Gn, c->rlist, c->rcount, &c->nrm[128], &c->flags[FLAG_RES2],
			          !dumping);
			// the learned guess's residual: the plain one where no pressure it reads has changed (set above), evaluated
			// on the listed cells
			launch_residual_list(D, c->Pn, c->rlist, c->rcount, c->Gn, &c->nrm[128], FLAG_RES2, 4 * c->cus, c->st);
		}
		rc = fetch(c); if (rc) return rc;
		if (!dumping) std::swap(c->G, c->Gplain);   // G: the plain guess's residual, as eval_all leaves it
		ok0 = c->hs->flags[FLAG_RES] == 0; res0 = nrm_val(c, 0);
		okl = c->hs->flags[FLAG_RES2] == 0; resl = nrm_val(c, 2);
	} else {
		rc = eval_all(c, D, c->Pst, nullptr, 0.0, nullptr, c->G, ok0, res0); if (rc) return rc;
	}
	if (c->final_pending) {   // the previous stage's final state, whose failure comes first
		c->final_pending = false;
		if (c->hs->flags[FLAG_FINAL_PREV]) { c->stats.closure_fails++; return SPB6_CLOSURE_FAIL; }
		c->stats.stages++;
	}
	if (c->hs->flags[FLAG_RHO]) return SPB6_RHO_NONPOSITIVE;
	if (c->hs->flags[FLAG_STAGE]) return SPB6_STAGE_INADMISSIBLE;
	const bool have_G = true;
	if (dumping && !learned) {
		// the plain guess and its residual: features and the activity mask of the training dump
		HIPCHK(hipMemcpyAsync(c->Pplain, c->Pst, n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
		HIPCHK(hipMemcpyAsync(c->Gplain, c->G, n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
		TRegion tr(c, T_GUESS);
		mlp_feat_cell(D, c->Pplain, c->F, c->st);
	}
	const double thr = ok0 ? GUESS_ACTIVE * res0 : 0.0;
	if (learned && !(ok0 && res0 < c->cfg.newton_tol)) {
		// the plain guess judges the learned one: keep whichever has the smaller Newton residual norm
		if (okl && (!ok0 || resl < res0)) {
			if (ok0 && resl > 0) c->stats.guess_log10_gain += std::log10(res0 / resl);
			std::swap(c->Pst, c->Pn); std::swap(c->G, c->Gn); ok0 = true; res0 = resl; c->stats.guess_used++;
			c->guess_losses = 0; c->guess_backoff = 0;
		}
		else {
			c->stats.guess_fallback++;
			if (++c->guess_losses >= 2) { c->guess_skip = 1 << std::min(c->guess_backoff, 5); c->guess_backoff++; }
		}
	}
	else if (learned) c->stats.guess_fallback++;   // the plain guess has already converged
	rc = newton(c, D, iters, res, have_G, ok0, res0);
	if (rc != SPB6_OK) return rc;
	if (dumping) {
		const int off = (int)((c->dseq++ * 2654435761u) % (unsigned)c->dstride);
		mlp_dump(D, c->F, adt, c->Pst, c->Pplain, c->Gplain, thr, off, c->dstride, c->dcap, c->dX, c->dY, c->dcount, c->st);
	}
	c->stats.maxres = std::max(c->stats.maxres, res);
	if (defer_final) {
		{ TRegion tr(c, T_FINAL); launch_final(D, c->Pst, Qs, adt, Hk, c->tmax, c->st, FLAG_FINAL_PREV); }
		launch_bc_grid(c->g, c->Ust, c->st);
		c->final_pending = true;
		return SPB6_OK;
	}
	rc = reset_scalars(c); if (rc) return rc;
	{ TRegion tr(c, T_FINAL); launch_final(D, c->Pst, Qs, adt, Hk, c->tmax, c->st); }
	launch_bc_grid(c->g, c->Ust, c->st);
	rc = fetch(c); if (rc) return rc;
	if (c->hs->flags[FLAG_FINAL]) { c->stats.closure_fails++; return SPB6_CLOSURE_FAIL; }
	c->stats.stages++;
	return SPB6_OK;
}

void set_tab(spb6_ctx* c, int rk) {   // PC:1195-1218: 1 backward Euler, 2 LSDIRK2(2,2,2), 3 SSP2(3,3,2)
	for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) c->Ae[a][b] = c->Ai[a][b] = 0;
	if (rk == 1) c->Ai[0][0] = 1.0;
	else if (rk == 2) { const double bb = 1.0 - 1.0 / std::sqrt(2.0), cc = 1.0 / (2.0 * bb); c->Ae[1][0] = cc; c->Ai[0][0] = bb; c->Ai[1][0] = 1 - bb; c->Ai[1][1] = bb; }
	else if (rk == 3) {
		const double r12 = 0.5, r13 = 1.0 / 3.0, r14 = 1.0 / 4.0;
		c->Ae[1][0] = r12; c->Ae[2][0] = r12; c->Ae[2][1] = r12;
		c->Ai[0][0] = r14; c->Ai[1][1] = r14; c->Ai[2][0] = r13; c->Ai[2][1] = r13; c->Ai[2][2] = r13;
	}
}

// one attempt (PC:1243-1285): on success Ust / Pst hold the new state and pressure
int attempt(spb6_ctx* c, double dt, int rk) {
	set_tab(c, rk);
	if (c->ctype) launch_relax_activate(c->g, c->ph, c->rel, c->cfg.relax, c->U, c->eps_i, c->ctype, c->st);   // PC:1247
	const size_t n = c->g.n;
	HIPCHK(hipMemcpyAsync(c->Pst, c->P, n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
	c->final_pending = false;
	for (int rks = 0; rks < rk; rks++) {
		const double *Qe = rks > 0 ? c->Qe : c->U, *Qs = rks > 0 ? c->Qs : c->U;
		const double adt = c->Ai[rks][rks] * dt;
		// the stage's fixed head: the stage operands, the explicit part (its flag is read with the first residual), the
		// preparation and the shock budget -- one graph in graph mode
		int rc = run_graph(c, c->g_head[rks], [&] {
			if (rks > 0) {
				TRegion tr(c, T_OTHER);
				launch_combine(c->g, c->U, c->H[0], c->H[1], rks, c->Ae[rks][0] * dt, rks > 1 ? c->Ae[rks][1] * dt : 0.0,
				               c->Ai[rks][0] * dt, rks > 1 ? c->Ai[rks][1] * dt : 0.0, c->Qe, c->Qs, c->st);
				launch_bc_grid(c->g, c->Qe, c->st); launch_bc_grid(c->g, c->Qs, c->st);
			}
			reset_stage(c, rks > 0);
			explicit_part(c, Qe, Qs, adt);
			prep_part(c, adt);
		});
		if (rc) return rc;
		int iters; double res;
		rc = implicit_part(c, adt, Qs, c->H[rks], iters, res, rks + 1 < rk, true);
		if (rc != SPB6_OK) return rc;
	}
	return SPB6_OK;
}

int caldt(spb6_ctx* c, bool acoustic, double& inv_rate) {   // PC:477-488
	HIPCHK(hipMemsetAsync(c->nrm, 0, 2 * 64 * sizeof(unsigned long long), c->st));
	launch_caldt(c->g, c->ph, c->U, acoustic ? 1 : 0, c->nrm, c->st);
	int rc = fetch(c); if (rc) return rc;
	const double ux = nrm_val(c, 0), uy = nrm_val(c, 1);
	const double rate = ux / c->g.dx + (c->g.ny > 1 ? uy / c->g.dy : 0.0);
	inv_rate = rate > 0 ? 1.0 / rate : INFINITY;
	return SPB6_OK;
}

// ---- the reference's fully explicit scheme (src/SPB6_explicit.cpp, "EX") at relaxation level 0 on the context's grids:
// the attempt's working state in Ust (V), the update's output in Qe; the stage data of the IMEX path are free meanwhile
int hllc_alloc(spb6_ctx* c) {
	if (c->hxp) return SPB6_OK;
	const size_t NC = c->g.NC;
	if (dalloc(c, &c->hxp, NC * hllc_prim_fields()) || dalloc(c, &c->hfx, NC * hllc_face_fields())) return SPB6_HIP_ERROR;
	if (c->g.twoD) { if (dalloc(c, &c->hfy, NC * hllc_face_fields())) return SPB6_HIP_ERROR; }
	else c->hfy = c->hfx;
	return SPB6_OK;
}
// RelaxationSubsystem (EX:344-361): ghost cells, the shock budget on the +-1 cross stencil, the per-cell relaxation
void hllc_relax(spb6_ctx* c, double* U) {
	launch_bc_grid(c->g, U, c->st);
	Dev D = make_dev(c); D.Ust = U; D.Ust_c = U;
	{ TRegion tr(c, T_SHOCK); launch_shock(D, c->slist, c->scount, c->srec, c->shock_steps, c->shock_blocks, c->cfg.shock_fused, c->st); }
	{ TRegion tr(c, T_FINAL); launch_hllc_relax(c->g, c->ph, U, c->Qph, c->flags, &c->mail->hxf, c->st); }
	launch_bc_grid(c->g, U, c->st);
}
// explicit_stage (EX:363-372): reconstruction, HLLC fluxes and the Godunov update of V into N, then its relaxation; V and
// N trade places
void hllc_stage(spb6_ctx* c, double*& V, double*& N, double h) {
	const Geo& g = c->g;
	launch_bc_grid(g, V, c->st);
	{
		TRegion tr(c, T_EXPLICIT);
		launch_hllc_prim(g, c->ph, V, c->hxp, c->flags, c->st);
		launch_hllc_flux(g, c->ph, c->hxp, c->cfg.order, c->cfg.limiter, c->hfx, c->hfy, c->flags, c->st);
		launch_hllc_update(g, V, c->hxp, c->hfx, c->hfy, h / g.dx, h / g.dy, N, c->flags, c->st);
	}
	hllc_relax(c, N);
	std::swap(V, N);
}
// one attempt of a step (EX:546-556): first order one stage; second order Heun, U2 = R(U1 + dt L(U1)) after U1 =
// R(U + dt L(U)), then (U + U2) / 2 relaxed. The passes stop at the first failure, as the reference does; one
// synchronisation reads the outcome. On success the new state becomes U.
int hllc_attempt(spb6_ctx* c, double dt) {
	int rc = reset_stage(c); if (rc) return rc;
	HIPCHK(hipMemsetAsync(&c->mail->hxf, 0, sizeof(int), c->st));
	const size_t bytes = (size_t)NV * c->g.NC * sizeof(double);
	double *V = c->Ust, *N = c->Qe;
	HIPCHK(hipMemcpyAsync(V, c->U, bytes, hipMemcpyDeviceToDevice, c->st));
	if (c->cfg.order < 2) hllc_stage(c, V, N, dt);
	else {
		hllc_stage(c, V, N, dt);
		hllc_stage(c, V, N, dt);
		launch_hllc_average((long)NV * c->g.NC, V, c->U, c->flags, c->st);
		hllc_relax(c, V);
	}
	rc = fetch(c); if (rc) return rc;
	c->stats.closure_fails += c->hs->hxf;
	if (c->hs->flags[FLAG_STAGE]) return SPB6_STAGE_INADMISSIBLE;
	if (c->hs->flags[FLAG_FINAL]) return SPB6_CLOSURE_FAIL;
	double* old = c->U;
	c->U = V; c->Ust = old; c->Qe = N;   // the three grids stay distinct
	return SPB6_OK;
}

// pbar_of (PHY:855-857) on the host
double pbar_host(const Phys& P, const double* Q) {
	const double a1 = Q[IA1];
	return a1 * p_ame(P, 0, a1, Q[IM1], Q[IE1]) + (1 - a1) * p_ame(P, 1, 1 - a1, Q[IM2], Q[IE2]);
}

// ---- the allocation-free linear solve (spb6_linsolve_device): its work vectors live in a file-scope device array
// of a fixed size, resolved once per device; see include/spb6_imex.h
constexpr long NA_ROWS = SPB6_LINSOLVE_DEVICE_MAX_ROWS;
constexpr long NA_STRIDE = NA_ROWS / 256 + 2;             // partial-sum slots of the 256-row blocks, as spb6_create sizes them
constexpr long NA_KS = 6 * NA_ROWS + KRYLOV_SLOTS * NA_STRIDE;   // offset of the loop scalars (doubles)
constexpr long NA_DOUBLES = NA_KS + (long)((sizeof(KS) + sizeof(double) - 1) / sizeof(double));

}   // namespace

// six BiCGStab work vectors (r, r0, p, v, s, t; Thomas uses the first two), the partial sums and the loop scalars:
// about 12.7 MB per device, held by every device the library runs on until the process ends
static __device__ __align__(256) double scratch[NA_DOUBLES];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[64] = {};
// the device address of scratch on the current device (hipGetSymbolAddress once per device, then cached); the array's
// name in host code is only a host-side stand-in and must never be passed to a kernel
double* scratch_ptr(int& dev) {
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) { g_err = "spb6_linsolve_device: no current device"; return nullptr; }
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) { g_err = std::string("spb6_linsolve_device: hipGetSymbolAddress: ") + hipGetErrorString(e); return nullptr; }
		g_scratch_ptr[dev] = static_cast<double*>(p);
	}
	return g_scratch_ptr[dev];
}
}   // namespace

// ================================================================================================ C ABI
extern "C" {

const char* spb6_last_error(void) { return g_err.c_str(); }

// the Mie-Grueneisen parameters of the phases (PHY:49-53) into the kernels' phase block
static void set_mg(Phys& P, const spb6_config_t* cfg) {
#if !SPB6_MG
	(void)P; (void)cfg;
#else
	P.mg = 0;
	for (int k = 0; k < 2; k++) {
		const double* m = cfg->mg[k];
		P.kind[k] = cfg->eos_kind[k];
		if (P.kind[k]) P.mg = 1;
		P.G[k] = m[0]; P.c1[k] = m[1]; P.c2[k] = m[2]; P.e1[k] = m[3]; P.e2[k] = m[4]; P.rr[k] = m[5]; P.e0[k] = m[6];
	}
#endif
}
spb6_ctx* spb6_create(const spb6_config_t* cfg, int device) {
	if (!cfg || cfg->nx < 2 || cfg->ny < 1) { g_err = "bad config"; return nullptr; }
	if (cfg->rus_local != 1) { g_err = "only the local Rusanov speed (rus_local = 1) is supported"; return nullptr; }
	if (cfg->rk < 1 || cfg->rk > 3) { g_err = "rk must be 1, 2 or 3"; return nullptr; }
	if (cfg->variant < SPB6_VARIANT_PC || cfg->variant > SPB6_VARIANT_BB) { g_err = "variant out of range"; return nullptr; }
	if (cfg->variant != SPB6_VARIANT_PC && (cfg->analytic_jacobian || cfg->learned_guess)) {
		g_err = "the IMEX variants run with the finite-difference Jacobian and the plain Newton guess"; return nullptr; }
	if (cfg->relax < 0 || cfg->relax > 3) { g_err = "relaxation level must be 0 (p), 1 (p-pT), 2 (p-pT-pTg) or 3 (p-pTg)"; return nullptr; }
	if (cfg->shock_fused < 0 || cfg->shock_fused > 1) { g_err = "shock_fused must be 0 or 1"; return nullptr; }
	if (cfg->relax > 0 && (cfg->analytic_jacobian || cfg->learned_guess)) {
		g_err = "relaxation levels 1-3 run with the finite-difference Jacobian and the plain Newton guess"; return nullptr; }
	for (int k = 0; k < 2; k++) {
		if (cfg->eos_kind[k] < 0 || cfg->eos_kind[k] > 2) { g_err = "eos_kind must be 0 (NASG), 1 (Cochran-Chan) or 2 (JWL)"; return nullptr; }
		if (cfg->eos_kind[k] && !(cfg->mg[k][0] > 0 && cfg->mg[k][5] > 0)) { g_err = "a Mie-Grueneisen phase needs G > 0 and rr > 0"; return nullptr; }
	}
	if (cfg->relax_form < 0 || cfg->relax_form > 1) {
		g_err = "relax_form is 0 (backward Euler) or 1 (the exponential form of the finite-rate closures)"; return nullptr; }
	if (cfg->nu_law && !(cfg->relax == 2 || cfg->relax == 3)) { g_err = "the nu law needs relaxation level 2 or 3"; return nullptr; }
	if (cfg->mu_rate > 0 && (cfg->analytic_jacobian || cfg->learned_guess || cfg->variant != SPB6_VARIANT_PC || cfg->eos_kind[0] ||
	                         cfg->eos_kind[1])) {
		g_err = "a finite mu runs with the reference scheme, the finite-difference Jacobian, the plain guess and NASG phases"; return nullptr; }
	if ((cfg->eos_kind[0] || cfg->eos_kind[1]) && !SPB6_MG) {
		g_err = "Mie-Grueneisen phases run in libspb6_hip_mg.so (the same sources built with SPB6_MG=1)"; return nullptr; }
	if ((cfg->eos_kind[0] || cfg->eos_kind[1]) && (cfg->relax || cfg->variant != SPB6_VARIANT_PC || cfg->shock_fused ||
	                                               cfg->analytic_jacobian || cfg->learned_guess || cfg->krylov_fp32)) {
		g_err = "a Mie-Grueneisen phase runs at relaxation level 0 with the reference scheme, the finite-difference Jacobian, "
		        "the FP64 Krylov solve, the plain guess and the exact profile form"; return nullptr; }
	spb6_ctx* c = new spb6_ctx();
	c->cfg = *cfg; c->device = device;
	std::memset(&c->stats, 0, sizeof c->stats);
	auto fail = [&](const char* what) -> spb6_ctx* { if (g_err.empty()) g_err = what; spb6_destroy(c); return nullptr; };
	if (hipSetDevice(device) != hipSuccess) return fail("hipSetDevice failed");
	if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) return fail("stream");
	Geo& g = c->g;
	g.nx = cfg->nx; g.ny = cfg->ny; g.NXT = cfg->nx + 2 * GH; g.NYT = cfg->ny + 2 * GH; g.NC = g.NXT * g.NYT; g.n = cfg->nx * cfg->ny;
	g.twoD = cfg->ny > 1; g.xper = cfg->bc[0] == SPB6_BC_PERIODIC; g.yper = cfg->bc[2] == SPB6_BC_PERIODIC;
	for (int q = 0; q < 4; q++) { g.bc[q] = cfg->bc[q]; g.bcu[q] = cfg->bcu[q]; }
	g.dx = (cfg->x1 - cfg->x0) / cfg->nx; g.dy = (cfg->y1 - cfg->y0) / cfg->ny;
	for (int k = 0; k < 2; k++) { c->ph.g[k] = cfg->ph[k].gamma; c->ph.pi[k] = cfg->ph[k].pi_inf; c->ph.q[k] = cfg->ph[k].q; c->ph.b[k] = cfg->ph[k].b; c->ph.cv[k] = cfg->ph[k].cv; c->ph.qp[k] = cfg->ph[k].qp; }
	set_mg(c->ph, cfg);
	const size_t NC = g.NC, n = g.n;
	g_err.clear();
	double** grids[] = {&c->U, &c->Ust, &c->Qe, &c->Qs, &c->H[0], &c->H[1], &c->donx};   // donors: 4 fields x 2 = NV per cell
	for (auto p : grids) if (dalloc(c, p, NC * NV)) return fail("alloc");
	if (cfg->rk == 3 && dalloc(c, &c->H[2], NC * NV)) return fail("alloc");
	if (g.twoD) { if (dalloc(c, &c->dony, NC * NV)) return fail("alloc"); }
	else c->dony = c->donx;
	if (cfg->variant != SPB6_VARIANT_PC) {
		double** vg[] = {&c->fwxL, &c->fwxR, &c->fwyL, &c->fwyR};
		for (auto p : vg) if (dalloc(c, p, NC * NV)) return fail("alloc");
		double** va[] = {&c->eTV1, &c->eTV2, &c->ETV, &c->dA};
		for (auto p : va) if (dalloc(c, p, NC)) return fail("alloc");
		if (dalloc(c, &c->donax, 2 * NC) || dalloc(c, &c->donay, 2 * NC)) return fail("alloc");
	}
	c->rel.mu = cfg->mu_rate > 0 ? cfg->mu_rate : INFINITY;   // PHY:33 (inf: instantaneous)
	c->rel.form = cfg->relax_form;
	if (cfg->relax > 0 || cfg->mu_rate > 0) {
		// the cells' closure types: at level 0 with a finite mu every cell is CL_MU (PHY:815), so the stage takes the
		// general per-cell-type path; each attempt's activation rewrites them
		if (hipMalloc((void**)&c->ctype, n * sizeof(int)) != hipSuccess) return fail("alloc");
		c->allocs.push_back(c->ctype);
		std::vector<int> t0(n, cfg->mu_rate > 0 ? CL_MU : CL_P);
		if (hipMemcpy(c->ctype, t0.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return fail("upload");
		if (cfg->eps_i > 0) c->eps_i = cfg->eps_i;
	}
	if (cfg->mu_rate > 0 && dalloc(c, &c->Z, 2 * NC)) return fail("alloc");
	c->rel.theta = cfg->theta_rate > 0 ? cfg->theta_rate : (cfg->relax >= 1 ? INFINITY : 0.0);                 // PC:1148-1149
	c->rel.nu = cfg->nu_rate > 0 ? cfg->nu_rate : ((cfg->relax == 2 || cfg->relax == 3) ? INFINITY : 0.0);
	if (cfg->nu_law) {   // PC:1150: the chemical cells take the nu closure, each with the rate of the law
		c->rel.nu_law = 1; c->rel.nu = 0.0;
		c->rel.nu_cr = cfg->nu_cr > 0 ? cfg->nu_cr : 1.0; c->rel.nu_pc = cfg->nu_pc > 0 ? cfg->nu_pc : 22.064e6;
		if (dalloc(c, &c->rel.cnu, n)) return fail("alloc");
	}
	if (cfg->relax >= 2) {   // the saturation line before the run (PC:1151)
		std::vector<double> lpv, Tv;
		build_sat(c->ph, lpv, Tv, c->rel.pcrit, c->rel.vcrit);
		if (lpv.size() < 2) return fail("the phases have no saturation line (relaxation levels 2-3 need one)");
		double *dl = nullptr, *dT = nullptr;
		if (dalloc(c, &dl, lpv.size()) || dalloc(c, &dT, Tv.size())) return fail("alloc");
		if (hipMemcpy(dl, lpv.data(), lpv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
		    hipMemcpy(dT, Tv.data(), Tv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return fail("upload");
		c->rel.lp = dl; c->rel.T = dT; c->rel.n = (int)lpv.size();
	}
	double** arrs[] = {&c->phx, &c->phy, &c->a2s, &c->ubx, &c->uby, &c->vsc, &c->esc, &c->vu, &c->vv, &c->ust, &c->vst, &c->cz, &c->Kw, &c->Pold,
	                   &c->ubsx, &c->kpx, &c->wLx, &c->pvx, &c->ubsy, &c->kpy, &c->wLy, &c->pvy};
	for (auto p : arrs) if (dalloc(c, p, NC)) return fail("alloc");
	if (dalloc(c, &c->Qph, 2 * NC)) return fail("alloc");
	double** vecs[] = {&c->P, &c->Pst, &c->Pn, &c->G, &c->Gn, &c->dP, &c->rhs, &c->sc, &c->cpw, &c->dpw, &c->kr, &c->kr0, &c->kp, &c->kv, &c->ks_, &c->kt};
	for (auto p : vecs) if (dalloc(c, p, n)) return fail("alloc");
	if (dalloc(c, &c->J, 5 * n)) return fail("alloc");
	if (cfg->krylov_fp32) {
		float** fv[] = {&c->b32, &c->x32, &c->r32, &c->r032, &c->p32, &c->v32, &c->s32, &c->t32, &c->ph32, &c->sh32, &c->dg32, &c->pp32};
		for (auto p : fv) { if (hipMalloc((void**)p, n * sizeof(float)) != hipSuccess) return fail("alloc"); c->allocs.push_back(*p); }
		if (hipMalloc((void**)&c->J32, 5 * n * sizeof(float)) != hipSuccess) return fail("alloc"); c->allocs.push_back(c->J32);
	}
	c->part_stride = (int)((n + 255) / 256) + 1;   // partial-sum slots sized for 256-row blocks
	if (dalloc(c, &c->part, (size_t)KRYLOV_SLOTS * c->part_stride)) return fail("alloc");
	if (dalloc(c, &c->part2, 3 * (size_t)c->part_stride)) return fail("alloc");
	if (hipMalloc((void**)&c->mail, sizeof(HostScratch)) != hipSuccess) return fail("alloc");
	c->allocs.push_back(c->mail);
	if (hipMemset(c->mail, 0, sizeof(HostScratch)) != hipSuccess) return fail("alloc");
	c->nrm = c->mail->nrm; c->flags = c->mail->flags; c->ksd = &c->mail->ks; c->dsum = c->mail->dsum;
	if (hipHostMalloc((void**)&c->hs, sizeof(HostScratch), hipHostMallocDefault) != hipSuccess) return fail("host alloc");
	if (dalloc(c, &c->Pplain, n) || dalloc(c, &c->Gplain, n)) return fail("alloc");
	if (hipMalloc((void**)&c->slist, (n + 2) * sizeof(int)) != hipSuccess) return fail("alloc"); c->allocs.push_back(c->slist);
	if (dalloc(c, &c->srec, (size_t)SHOCK_REC * n)) return fail("alloc");
	if (hipMalloc((void**)&c->tmax, 128 * sizeof(unsigned long long)) != hipSuccess) return fail("alloc");
	c->allocs.push_back(c->tmax);
	if (hipMalloc((void**)&c->shock_steps, sizeof(unsigned long long)) != hipSuccess) return fail("alloc");
	c->allocs.push_back(c->shock_steps);
	if (hipMemset(c->shock_steps, 0, sizeof(unsigned long long)) != hipSuccess) return fail("alloc");
	c->scount = c->slist + n;
	if (hipMalloc((void**)&c->rlist, (n + 1) * sizeof(int)) != hipSuccess) return fail("alloc");
	c->allocs.push_back(c->rlist);
	c->rcount = c->rlist + n;
	(void)hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device);
	c->shock_blocks = c->cus * shock_strong_blocks_per_cu();
	if (hipMalloc((void**)&c->F, 7 * NC * sizeof(float)) != hipSuccess) return fail("alloc"); c->allocs.push_back(c->F);
	if (hipDeviceSynchronize() != hipSuccess) return fail("sync");
	return c;
}

void spb6_destroy(spb6_ctx* c) {
	if (c) {
		for (auto& e : c->g_head) if (e) (void)hipGraphExecDestroy(e);
		for (auto& row : c->g_chunk) for (auto& e : row) if (e) (void)hipGraphExecDestroy(e);
	}
	if (!c) return;
	if (c->st) hipStreamSynchronize(c->st);
	for (void* p : c->allocs) hipFree(p);
	if (c->dX) { hipFree(c->dX); hipFree(c->dY); hipFree(c->dcount); }
	if (c->hs) hipHostFree(c->hs);
	if (c->st) hipStreamDestroy(c->st);
	delete c;
}

int spb6_set_state(spb6_ctx* c, const double* U_aos) {
	if (!c || !U_aos) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	const Geo& g = c->g;
	std::vector<double> h((size_t)g.NC * NV, 0.0);
	for (int j = 0; j < g.ny; j++) for (int i = 0; i < g.nx; i++) {
		const size_t cc = (size_t)(i + GH) + (size_t)(j + GH) * g.NXT;
		for (int k = 0; k < NV; k++) h[(size_t)k * g.NC + cc] = U_aos[((size_t)i + (size_t)j * g.nx) * NV + k];
	}
	HIPCHK(hipMemcpyAsync(c->U, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
	launch_bc_grid(g, c->U, c->st);
	// P_REF = min pbar, P = pbar - P_REF (PC:1222-1227)
	std::vector<double> P((size_t)g.n);
	double pref = INFINITY;
	for (int q = 0; q < g.n; q++) { const double pb = pbar_host(c->ph, &U_aos[(size_t)q * NV]); pref = (pb < pref) ? pb : pref; }
	if (!std::isfinite(pref)) pref = 0.0;
	for (int q = 0; q < g.n; q++) P[q] = pbar_host(c->ph, &U_aos[(size_t)q * NV]) - pref;
	c->P_REF = pref;
	HIPCHK(hipMemcpyAsync(c->P, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipMemcpyAsync(c->Pst, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	c->t = 0.0; c->dtprev = -1.0; c->startup = true; c->last_krylov = 0;
	HIPCHK(hipMemset(c->tmax, 0, 128 * sizeof(unsigned long long)));
	c->M0[0] = c->M0[1] = c->M0[2] = 0.0;   // conserved_totals (PC:1104-1107) in the reference's cell order
	for (int j = 0; j < g.ny; j++) for (int i = 0; i < g.nx; i++) {
		const double* Q = &U_aos[((size_t)i + (size_t)j * g.nx) * NV];
		c->M0[0] += Q[IM1]; c->M0[1] += Q[IM2]; c->M0[2] += Q[IEN];
	}
	c->guess_skip = c->guess_losses = c->guess_backoff = 0;
	std::memset(&c->stats, 0, sizeof c->stats);
	return SPB6_OK;
}

int spb6_get_state(spb6_ctx* c, double* U_aos) {
	if (!c || !U_aos) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	const Geo& g = c->g;
	std::vector<double> h((size_t)g.NC * NV);
	HIPCHK(hipMemcpyAsync(h.data(), c->U, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	for (int j = 0; j < g.ny; j++) for (int i = 0; i < g.nx; i++) {
		const size_t cc = (size_t)(i + GH) + (size_t)(j + GH) * g.NXT;
		for (int k = 0; k < NV; k++) U_aos[((size_t)i + (size_t)j * g.nx) * NV + k] = h[(size_t)k * g.NC + cc];
	}
	return SPB6_OK;
}

int spb6_get_pressure(spb6_ctx* c, double* P, double* p_ref) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	if (P) { HIPCHK(hipMemcpyAsync(P, c->P, (size_t)c->g.n * sizeof(double), hipMemcpyDeviceToHost, c->st)); HIPCHK(hipStreamSynchronize(c->st)); }
	if (p_ref) *p_ref = c->P_REF;
	return SPB6_OK;
}
int spb6_set_pressure(spb6_ctx* c, const double* P, double p_ref) {
	if (!c || !P) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	c->P_REF = p_ref;
	HIPCHK(hipMemcpyAsync(c->P, P, (size_t)c->g.n * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipMemcpyAsync(c->Pst, P, (size_t)c->g.n * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return SPB6_OK;
}

int spb6_caldt(spb6_ctx* c, int acoustic, double* inv_rate) {
	if (!c || !inv_rate) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	return caldt(c, acoustic != 0, *inv_rate);
}
int spb6_bc_fill(spb6_ctx* c) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	launch_bc_grid(c->g, c->U, c->st);
	HIPCHK(hipStreamSynchronize(c->st));
	return SPB6_OK;
}

int spb6_explicit_stage(spb6_ctx* c, double adt) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	HIPCHK(hipMemcpyAsync(c->Pst, c->P, (size_t)c->g.n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
	int rc = reset_stage(c); if (rc) return rc;
	rc = explicit_part(c, c->U, c->U, adt); if (rc) return rc;
	rc = fetch(c); if (rc) return rc;
	return c->hs->flags[FLAG_RHO] ? SPB6_RHO_NONPOSITIVE : SPB6_OK;
}
int spb6_stage_prep(spb6_ctx* c, double adt) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	int rc = reset_stage(c); if (rc) return rc;
	rc = prep_part(c, adt); if (rc) return rc;
	rc = fetch(c); if (rc) return rc;
	return c->hs->flags[FLAG_STAGE] ? SPB6_STAGE_INADMISSIBLE : SPB6_OK;
}
int spb6_implicit_stage(spb6_ctx* c, double adt, int* newton_iters, double* residual) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	int it = 0; double res = 0;
	if (reset_stage(c)) return SPB6_HIP_ERROR;   // the explicit part's flag was read by spb6_explicit_stage
	c->final_pending = false;
	const int rc = implicit_part(c, adt, c->U, c->H[0], it, res);
	if (newton_iters) *newton_iters = it;
	if (residual) *residual = res;
	return rc;
}
int spb6_step(spb6_ctx* c, double dt, int rk) {
	if (!c || !(dt > 0) || rk < 1 || rk > 3 || (rk == 3 && !c->H[2])) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	const int rc = attempt(c, dt, rk);
	if (rc == SPB6_OK) { std::swap(c->U, c->Ust); std::swap(c->P, c->Pst); c->t += dt; c->stats.steps++; }
	return rc;
}

int spb6_attempt_begin(spb6_ctx* c) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	if (c->ctype) launch_relax_activate(c->g, c->ph, c->rel, c->cfg.relax, c->U, c->eps_i, c->ctype, c->st);   // PC:1247
	HIPCHK(hipMemcpyAsync(c->Pst, c->P, (size_t)c->g.n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
	c->final_pending = false;
	c->qe_cur = c->qs_cur = c->U;
	return SPB6_OK;
}
int spb6_stage_combine(spb6_ctx* c, int rks, double dt, const double* ae, const double* ai) {
	if (!c || rks < 0 || rks > 2 || (rks > 0 && (!ae || !ai))) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	if (rks == 0) { c->qe_cur = c->qs_cur = c->U; return SPB6_OK; }
	launch_combine(c->g, c->U, c->H[0], c->H[1], rks, ae[0] * dt, rks > 1 ? ae[1] * dt : 0.0, ai[0] * dt, rks > 1 ? ai[1] * dt : 0.0,
	               c->Qe, c->Qs, c->st);
	launch_bc_grid(c->g, c->Qe, c->st); launch_bc_grid(c->g, c->Qs, c->st);
	c->qe_cur = c->Qe; c->qs_cur = c->Qs;
	return SPB6_OK;
}
int spb6_stage_update(spb6_ctx* c, double adt) {
	if (!c || !c->qe_cur || !(adt > 0)) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	int rc = reset_stage(c); if (rc) return rc;
	rc = explicit_part(c, c->qe_cur, c->qs_cur, adt); if (rc) return rc;
	rc = fetch(c); if (rc) return rc;
	return c->hs->flags[FLAG_RHO] ? SPB6_RHO_NONPOSITIVE : SPB6_OK;
}
int spb6_stage_pressure(spb6_ctx* c, double adt, int rks, int* newton_iters, double* residual) {
	if (!c || !c->qs_cur || !(adt > 0) || rks < 0 || rks > 2 || !c->H[rks]) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	int it = 0; double res = 0;
	if (reset_stage(c)) return SPB6_HIP_ERROR;   // the explicit part's flag was read by spb6_stage_update
	c->final_pending = false;
	const int rc = implicit_part(c, adt, c->qs_cur, c->H[rks], it, res);
	if (newton_iters) *newton_iters = it;
	if (residual) *residual = res;
	return rc;
}
int spb6_attempt_accept(spb6_ctx* c, double dt) {
	if (!c) return SPB6_BAD_ARG;
	std::swap(c->U, c->Ust); std::swap(c->P, c->Pst);
	c->t = c->t + dt; c->stats.steps++; c->stats.dt_last = dt; c->stats.t = c->t;
	c->qe_cur = c->qs_cur = nullptr;
	return SPB6_OK;
}

int spb6_run(spb6_ctx* c, double t_end, long max_steps) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	const double ramp_factor = 1.2;
	long done_steps = 0;
	while (c->t < t_end * (1 - 1e-12) && (max_steps < 0 || done_steps < max_steps)) {
		bool ramping = false; double dt, ir;
		int rc;
		if (c->cfg.dtfix > 0) dt = c->cfg.dtfix;
		else {
			rc = caldt(c, c->cfg.acdt != 0, ir); if (rc) return rc;
			dt = c->cfg.cfl * ir;
			if (c->cfg.ramp) {
				if (c->dtprev < 0) { rc = caldt(c, true, ir); if (rc) return rc; c->dtprev = c->cfg.cfl_ac * ir / ramp_factor; }
				if (dt > ramp_factor * c->dtprev) { dt = ramp_factor * c->dtprev; ramping = true; }
			}
			if (!std::isfinite(dt)) { rc = caldt(c, true, ir); if (rc) return rc; dt = c->cfg.cfl_ac * ir; }
		}
		dt = ((t_end - c->t) < dt) ? (t_end - c->t) : dt;
		if (!(dt > 1e-12 * t_end)) { c->stats.status = SPB6_DT_UNDERFLOW; return SPB6_DT_UNDERFLOW; }
		bool ok = false;
		for (int att = 0; att < 30 && !ok; att++) {
			const bool first_order = c->cfg.ramp && ramping && c->startup;
			const int rk = first_order ? 1 : c->cfg.rk; if (rk == 1) c->stats.rk1++;
			rc = attempt(c, dt, rk);
			if (rc == SPB6_HIP_ERROR || rc == SPB6_BAD_ARG) return rc;
			if (rc == SPB6_OK) { std::swap(c->U, c->Ust); std::swap(c->P, c->Pst); ok = true; }
			else { c->stats.rejects++; dt *= 0.5; }
		}
		if (!ok) { c->stats.status = SPB6_STEP_FAILED; return SPB6_STEP_FAILED; }
		c->t = c->t + dt; c->dtprev = dt; c->stats.steps++; c->stats.dt_last = dt; done_steps++;
		if (!ramping) c->startup = false;
	}
	c->stats.t = c->t;
	return SPB6_OK;
}

int spb6_hllc_step(spb6_ctx* c, double dt) {
	if (!c || !(dt > 0) || c->ctype || mg_any(c->ph)) return SPB6_BAD_ARG;   // relaxation level 0, NASG phases only
	hipSetDevice(c->device);
	int rc = hllc_alloc(c); if (rc) return rc;
	rc = hllc_attempt(c, dt);
	if (rc == SPB6_OK) { c->t += dt; c->stats.steps++; c->stats.dt_last = dt; c->stats.t = c->t; }
	return rc;
}
int spb6_hllc_run(spb6_ctx* c, double t_end, long max_steps, double cfl) {
	if (!c || !(cfl > 0) || c->ctype || mg_any(c->ph)) return SPB6_BAD_ARG;   // relaxation level 0, NASG phases only
	hipSetDevice(c->device);
	int rc = hllc_alloc(c); if (rc) return rc;
	long done_steps = 0;
	while (c->t < t_end * (1 - 1e-12) && (max_steps < 0 || done_steps < max_steps)) {
		double ir;
		rc = caldt(c, true, ir); if (rc) return rc;
		double dt = cfl * ir;
		dt = ((t_end - c->t) < dt) ? (t_end - c->t) : dt;
		if (!(dt > 1e-12 * t_end)) { c->stats.status = SPB6_DT_UNDERFLOW; return SPB6_DT_UNDERFLOW; }
		bool ok = false;
		for (int att = 0; att < 30 && !ok; att++) {
			rc = hllc_attempt(c, dt);
			if (rc == SPB6_HIP_ERROR || rc == SPB6_BAD_ARG) return rc;
			if (rc == SPB6_OK) ok = true;
			else { c->stats.rejects++; dt *= 0.5; }
		}
		if (!ok) { c->stats.status = SPB6_STEP_FAILED; return SPB6_STEP_FAILED; }
		c->t = c->t + dt; c->stats.steps++; c->stats.dt_last = dt; done_steps++;
	}
	c->stats.t = c->t;
	return SPB6_OK;
}

int spb6_relax_activate(spb6_ctx* c) {
	if (!c) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	if (!c->ctype) return SPB6_OK;
	launch_relax_activate(c->g, c->ph, c->rel, c->cfg.relax, c->U, c->eps_i, c->ctype, c->st);
	HIPCHK(hipStreamSynchronize(c->st));
	return SPB6_OK;
}

int spb6_get_stats(spb6_ctx* c, spb6_stats_t* s) {
	if (!c || !s) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	unsigned long long tm[128];
	HIPCHK(hipMemcpyAsync(tm, c->tmax, sizeof tm, hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	*s = c->stats; s->t = c->t;
	for (int k = 0; k < 2; k++) {
		unsigned long long m = 0;
		for (int q = 0; q < 64; q++) m = std::max(m, tm[64 * k + q]);
		std::memcpy(&s->tmax[k], &m, sizeof m);
	}
	return SPB6_OK;
}
double spb6_time(spb6_ctx* c) { return c ? c->t : 0.0; }

int spb6_get_profile(spb6_ctx* c, double* out) {
	if (!c || !out) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	const Geo& g = c->g;
	const size_t n = g.n;
	// an output routine: its 15 n doubles of device memory are allocated for the call (as spb6_get_field's face states)
	double* tmp = nullptr;
	HIPCHK(hipMalloc((void**)&tmp, n * 15 * sizeof(double)));
	launch_cell_diag(g, c->ph, c->U, tmp, c->st);
	std::vector<double> d(n * 15);
	const hipError_t e1 = hipMemcpyAsync(d.data(), tmp, d.size() * sizeof(double), hipMemcpyDeviceToHost, c->st);
	const hipError_t e2 = hipStreamSynchronize(c->st);
	(void)hipFree(tmp);
	HIPCHK(e1); HIPCHK(e2);
	for (int j = 0; j < g.ny; j++) for (int i = 0; i < g.nx; i++) {
		const size_t q = (size_t)i + (size_t)j * g.nx;
		double* o = out + q * 17;
		o[0] = c->cfg.x0 + (i + 0.5) * g.dx; o[1] = c->cfg.y0 + (j + 0.5) * g.dy;
		for (int k = 0; k < 15; k++) o[2 + k] = d[q * 15 + k];
	}
	return SPB6_OK;
}
int spb6_write_profile(spb6_ctx* c, const char* path, int prec) {
	if (!c || !path || prec < 1 || prec > 17) return SPB6_BAD_ARG;
	std::vector<double> o((size_t)c->g.n * 17);
	int rc = spb6_get_profile(c, o.data()); if (rc) return rc;
	FILE* f = std::fopen(path, "w");
	if (!f) { g_err = std::string("cannot open ") + path; return SPB6_BAD_ARG; }
	const int P = prec, Q = std::min(prec, 8), R = std::min(prec, 10);
	std::fprintf(f, "# x y p u v rho alpha2 Y2 rho1 rho2 c_frozen c_wood T1 T2 p1 p2 g1-g2\n");
	for (size_t q = 0; q < (size_t)c->g.n; q++) {
		const double* v = &o[q * 17];
		std::fprintf(f, "%.10e %.10e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e %.*e\n", v[0], v[1], P, v[2], P, v[3],
		             P, v[4], P, v[5], P, v[6], P, v[7], P, v[8], P, v[9], Q, v[10], Q, v[11], R, v[12], R, v[13], P, v[14], P, v[15], Q, v[16]);
	}
	return std::fclose(f) == 0 ? SPB6_OK : SPB6_BAD_ARG;
}
int spb6_conserved_totals(spb6_ctx* c, double M[3]) {
	if (!c || !M) return SPB6_BAD_ARG;
	std::vector<double> U((size_t)c->g.n * NV);
	int rc = spb6_get_state(c, U.data()); if (rc) return rc;
	M[0] = M[1] = M[2] = 0.0;
	for (size_t q = 0; q < (size_t)c->g.n; q++) { M[0] += U[q * NV + IM1]; M[1] += U[q * NV + IM2]; M[2] += U[q * NV + IEN]; }
	return SPB6_OK;
}
int spb6_stats_line(spb6_ctx* c, int case_id, double wall_s, char* buf, long len) {
	if (!c || !buf || len < 1) return SPB6_BAD_ARG;
	spb6_stats_t s;
	int rc = spb6_get_stats(c, &s); if (rc) return rc;
	double M1[3];
	rc = spb6_conserved_totals(c, M1); if (rc) return rc;
	std::snprintf(buf, (size_t)len,
	              "case %d mode 1 order %d rk %d N %d CFL %g | steps %ld (rk1 %ld) rejects %ld floor %ld | stages %ld newton it %ld (%.2f/stage) "
	              "maxres %.2e Tmax %.4g %.4g | closure fails %ld | dM1/M %.2e dM2/M %.2e dE/E %.2e | wall %.2f s",
	              case_id, c->cfg.order, c->cfg.rk, c->cfg.nx, c->cfg.cfl, s.steps, s.rk1, s.rejects, s.floor_acc, s.stages, s.newton_it,
	              s.stages ? (double)s.newton_it / s.stages : 0.0, s.maxres, s.tmax[0], s.tmax[1], s.closure_fails,
	              (M1[0] - c->M0[0]) / c->M0[0], (M1[1] - c->M0[1]) / c->M0[1], (M1[2] - c->M0[2]) / c->M0[2], wall_s);
	return SPB6_OK;
}

int spb6_residual(spb6_ctx* c, const double* P, double* G) {
	if (!c || !P || !G) return -1;
	hipSetDevice(c->device);
	const size_t n = c->g.n;
	if (hipMemcpyAsync(c->Pn, P, n * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess) return -1;
	Dev D = stage_dev(c, c->tau_last);
	if (reset_scalars(c)) return -1;
	launch_residual(D, c->Pn, nullptr, 0.0, nullptr, c->Gn, &c->nrm[0], FLAG_RES, c->st);
	if (hipMemcpyAsync(G, c->Gn, n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (hipStreamSynchronize(c->st) != hipSuccess) return -1;
	int bad = 0; for (size_t q = 0; q < n; q++) if (G[q] != G[q]) bad++;
	return bad;
}
int spb6_guess_residual(spb6_ctx* c, const double* P, double* Pg, double* G, double* G_full, double* norms, double* Pg_full) {
	if (!c || !P || !Pg || !G || !G_full || !norms || !Pg_full || !c->model) return -1;
	hipSetDevice(c->device);
	const size_t n = c->g.n;
	if (hipMemcpyAsync(c->Pplain, P, n * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess) return -1;
	Dev D = stage_dev(c, c->tau_last);
	if (reset_scalars(c)) return -1;
	// as implicit_part's learned branch: the plain residual, the guess and its residual (norm slots 128..)
	launch_residual(D, c->Pplain, nullptr, 0.0, nullptr, c->Gplain, &c->nrm[0], FLAG_RES, c->st);
	const spb6_mlp* M = c->model;
	MlpDev W{M->P, M->W1s, M->W1f, M->W2p, M->mub1, M->b3, M->mu_t, M->sd_t};
	// the guess from the features of every cell, then (features cleared to NaN) as a stage computes it
	mlp_feat_cell(D, c->Pplain, c->F, c->st);
	mlp_infer(&D, c->F, c->tau_last, (int)n, nullptr, W, c->Pplain, c->Gplain, 0.0, c->Pn, nullptr, c->st, c->slist, c->scount,
	          &c->nrm[0], &c->flags[FLAG_RES], &c->mail->gthr, c->Gn, c->rlist, c->rcount, &c->nrm[128], &c->flags[FLAG_RES2]);
	if (hipMemcpyAsync(Pg_full, c->Pn, n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (hipMemsetAsync(c->F, 0xFF, (size_t)7 * c->g.NC * sizeof(float), c->st) != hipSuccess) return -1;
	mlp_infer(&D, c->F, c->tau_last, (int)n, nullptr, W, c->Pplain, c->Gplain, 0.0, c->Pn, nullptr, c->st, c->slist, c->scount,
	          &c->nrm[0], &c->flags[FLAG_RES], &c->mail->gthr, c->Gn, c->rlist, c->rcount, &c->nrm[128], &c->flags[FLAG_RES2], true);
	launch_residual_list(D, c->Pn, c->rlist, c->rcount, c->Gn, &c->nrm[128], FLAG_RES2, 4 * c->cus, c->st);
	// the full evaluation at the guess (norm slots 192..)
	launch_residual(D, c->Pn, nullptr, 0.0, nullptr, c->kr, &c->nrm[192], FLAG_FLOOR, c->st);
	int cnt = 0;
	if (hipMemcpyAsync(Pg, c->Pn, n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (hipMemcpyAsync(G, c->Gn, n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (hipMemcpyAsync(G_full, c->kr, n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (hipMemcpyAsync(&cnt, c->rcount, sizeof(int), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (fetch(c)) return -1;
	norms[0] = nrm_val(c, 2); norms[1] = nrm_val(c, 3);
	return cnt;
}
int spb6_jacobian(spb6_ctx* c, const double* P, const double* G, double* val, double* sc) {
	if (!c || !P || !G || !val || !sc) return -1;
	hipSetDevice(c->device);
	const size_t n = c->g.n;
	if (hipMemcpyAsync(c->Pn, P, n * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess) return -1;
	if (hipMemcpyAsync(c->Gn, G, n * sizeof(double), hipMemcpyHostToDevice, c->st) != hipSuccess) return -1;
	Dev D = stage_dev(c, c->tau_last);
	if (reset_scalars(c)) return -1;
	launch_jacobian(D, c->Pn, c->Gn, c->sc, c->J, c->rhs, c->cfg.analytic_jacobian != 0, c->st);
	std::vector<double> J(5 * n);
	if (hipMemcpyAsync(J.data(), c->J, 5 * n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (hipMemcpyAsync(sc, c->sc, n * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (fetch(c)) return -1;
	for (size_t r = 0; r < n; r++) for (int w = 0; w < 5; w++) val[5 * r + w] = J[w * n + r];
	return c->hs->flags[FLAG_JAC] ? 1 : 0;
}
int spb6_linsolve(spb6_ctx* c, const double* val, const double* b, double* x, int use_thomas, int* iters) {
	if (!c || !val || !b || !x) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	const size_t n = c->g.n;
	std::vector<double> J(5 * n);
	for (size_t r = 0; r < n; r++) for (int w = 0; w < 5; w++) J[w * n + r] = val[5 * r + w];
	HIPCHK(hipMemcpyAsync(c->J, J.data(), 5 * n * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipMemcpyAsync(c->rhs, b, n * sizeof(double), hipMemcpyHostToDevice, c->st));
	int it = 0; bool ok = true;
	if (use_thomas) launch_thomas((int)n, c->J, c->rhs, c->dP, c->cpw, c->dpw, c->st);
	else { bool crit; double sn; int rc = krylov(c, it, ok, 1.0, true, crit, sn, false); if (rc) return rc; }
	HIPCHK(hipMemcpyAsync(x, c->dP, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	if (iters) *iters = it;
	return ok ? SPB6_OK : SPB6_KRYLOV_FAIL;
}

int spb6_linsolve_device(int nx, int ny, const int* bc, const double* J, const double* b, double* x, double tol, int maxit,
                         int use_thomas, int* iters, void* stream) {
	if (nx < 1 || ny < 1 || !bc || !J || !b || !x || maxit < 0) { g_err = "spb6_linsolve_device: bad argument"; return SPB6_BAD_ARG; }
	const long n = (long)nx * ny;
	if (n > NA_ROWS) {
		g_err = "spb6_linsolve_device: " + std::to_string(n) + " rows exceed the maximum of " + std::to_string(NA_ROWS);
		return SPB6_BAD_ARG;
	}
	if (use_thomas && ny != 1) { g_err = "spb6_linsolve_device: Thomas needs a 1D row (ny = 1)"; return SPB6_BAD_ARG; }
	int dev = 0;
	double* S = scratch_ptr(dev);
	if (!S) return SPB6_HIP_ERROR;
	hipStream_t st = static_cast<hipStream_t>(stream);
	Geo g{};
	g.nx = nx; g.ny = ny; g.NXT = nx + 2 * GH; g.NYT = ny + 2 * GH; g.NC = g.NXT * g.NYT; g.n = (int)n;
	g.twoD = ny > 1; g.xper = bc[0] == SPB6_BC_PERIODIC; g.yper = bc[2] == SPB6_BC_PERIODIC;
	for (int q = 0; q < 4; q++) g.bc[q] = bc[q];
	if (use_thomas) {
		launch_thomas((int)n, J, b, x, S, S + NA_ROWS, st);
		HIPCHK(hipStreamSynchronize(st));
		if (iters) *iters = 0;
		return SPB6_OK;
	}
	// every buffer is written before it is read: setup 1 writes the partials setup 2 reads, setup 2 sets the loop scalars
	// (done, iteration count, rho, alpha, omega of iteration 0) and x, p, v, setup 3 r and r0; nothing from an earlier call
	// is used
	KS* ks = reinterpret_cast<KS*>(S + NA_KS);
	KrylovBufs B{x, S, S + NA_ROWS, S + 2 * NA_ROWS, S + 3 * NA_ROWS, S + 4 * NA_ROWS, S + 5 * NA_ROWS, J, S + 6 * NA_ROWS,
	             (int)((n + 255) / 256) + 1, ks};
	launch_krylov_setup(g, B, b, st);
	KS hk{};
	int launched = 0;
	for (;;) {
		for (int k = 0; k < 4 && launched <= maxit; k++, launched++) launch_krylov_iter(g, B, st, launched, tol, maxit);
		HIPCHK(hipMemcpyAsync(&hk, ks, sizeof(KS), hipMemcpyDeviceToHost, st));
		HIPCHK(hipStreamSynchronize(st));
		if (hk.done || launched > maxit) break;
	}
	if (iters) *iters = hk.iters;
	return hk.done && hk.status == 0 ? SPB6_OK : SPB6_KRYLOV_FAIL;
}

long spb6_get_field(spb6_ctx* c, const char* name, double* out, long capacity) {
	if (!c || !name || !out) return -1;
	hipSetDevice(c->device);
	const std::string s(name);
	const Geo& g = c->g;
	const double* grid = nullptr; const double* arr = nullptr; long cnt = 0;
	const int face = s == "wxL" ? 0 : s == "wxR" ? 1 : s == "wyL" ? 2 : s == "wyR" ? 3 : -1;
	if (s == "U") grid = c->U; else if (s == "Ust") grid = c->Ust; else if (s == "Qe") grid = c->Qe; else if (s == "Qs") grid = c->Qs;
	else if (face >= 0) {}
	else if (s == "phx") arr = c->phx; else if (s == "phy") arr = c->phy; else if (s == "a2s") arr = c->a2s; else if (s == "ubx") arr = c->ubx; else if (s == "uby") arr = c->uby;
	else if (s == "ust") arr = c->ust; else if (s == "vst") arr = c->vst; else if (s == "cz") arr = c->cz; else if (s == "Kw") arr = c->Kw;
	else if (s == "ubsx") arr = c->ubsx; else if (s == "kpx") arr = c->kpx; else if (s == "wLx") arr = c->wLx; else if (s == "pvx") arr = c->pvx;
	else if (s == "ubsy") arr = c->ubsy; else if (s == "kpy") arr = c->kpy; else if (s == "wLy") arr = c->wLy; else if (s == "pvy") arr = c->pvy;
	else if (s == "Qph") { arr = c->Qph; cnt = 2L * g.NC; }
	else if (s == "P") { arr = c->P; cnt = g.n; } else if (s == "Pst") { arr = c->Pst; cnt = g.n; }
	else return -1;
	if (grid || face >= 0) {
		const long tot = (long)g.NC * NV; if (capacity < tot) return -1;
		std::vector<double> h(tot);
		double* tmp = nullptr;
		if (face >= 0) {
			// the explicit stage keeps only the face donors: the full face states of its Qe are rebuilt for inspection
			if (!c->qe_last) return -1;
			if (hipMalloc((void**)&tmp, 4 * (size_t)tot * sizeof(double)) != hipSuccess) return -1;
			double *xL = tmp, *xR = tmp + tot, *yL = g.twoD ? tmp + 2 * tot : xL, *yR = g.twoD ? tmp + 3 * tot : xR;
			launch_recon_full(make_dev(c), c->cfg.limiter, c->cfg.order, c->qe_last, xL, xR, yL, yR, c->st);
			grid = face == 0 ? xL : face == 1 ? xR : face == 2 ? yL : yR;
		}
		const bool ok = hipMemcpyAsync(h.data(), grid, tot * sizeof(double), hipMemcpyDeviceToHost, c->st) == hipSuccess &&
		                hipStreamSynchronize(c->st) == hipSuccess;
		if (tmp) (void)hipFree(tmp);
		if (!ok) return -1;
		for (long cc = 0; cc < g.NC; cc++) for (int k = 0; k < NV; k++) out[cc * NV + k] = h[(size_t)k * g.NC + cc];
		return tot;
	}
	if (!cnt) cnt = g.NC;
	if (capacity < cnt) return -1;
	if (hipMemcpyAsync(out, arr, cnt * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return -1;
	if (hipStreamSynchronize(c->st) != hipSuccess) return -1;
	return cnt;
}

int spb6_set_timing_mask(spb6_ctx* c, unsigned mask) {
	if (!c) return SPB6_BAD_ARG;
	c->timer.resolve();
	c->timer.mask = mask;
	return SPB6_OK;
}
int spb6_enable_timing(spb6_ctx* c, int on) {
	if (!c) return SPB6_BAD_ARG;
	c->timer.resolve();
	c->timer.on = on != 0;
	for (int k = 0; k < T_NCAT; k++) { c->timer.ms[k] = 0; c->timer.cnt[k] = 0; }
	c->krylov_launched = 0;
	HIPCHK(hipMemsetAsync(c->shock_steps, 0, sizeof(unsigned long long), c->st));
	return SPB6_OK;
}
int spb6_get_kernel_times(spb6_ctx* c, spb6_kernel_times_t* t) {
	if (!c || !t) return SPB6_BAD_ARG;
	c->timer.resolve();
	const Timer& T = c->timer;
	t->explicit_ms = T.ms[T_EXPLICIT]; t->prep_ms = T.ms[T_PREP]; t->shock_ms = T.ms[T_SHOCK]; t->residual_ms = T.ms[T_RESIDUAL];
	t->jacobian_ms = T.ms[T_JACOBIAN]; t->krylov_ms = T.ms[T_KRYLOV]; t->final_ms = T.ms[T_FINAL]; t->other_ms = T.ms[T_OTHER];
	t->explicit_n = T.cnt[T_EXPLICIT]; t->prep_n = T.cnt[T_PREP]; t->shock_n = T.cnt[T_SHOCK]; t->residual_n = T.cnt[T_RESIDUAL];
	t->jacobian_n = T.cnt[T_JACOBIAN]; t->krylov_n = T.cnt[T_KRYLOV]; t->final_n = T.cnt[T_FINAL];
	t->krylov_iters = c->stats.krylov_it;
	t->krylov_aux_ms = T.ms[T_KAUX]; t->guess_ms = T.ms[T_GUESS]; t->krylov_launched = c->krylov_launched;
	t->krylov_idle_ms = T.ms[T_KIDLE];
	unsigned long long rk = 0;
	HIPCHK(hipMemcpyAsync(&rk, c->shock_steps, sizeof rk, hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	t->shock_rk_steps = (long)rk;
	t->graph_launches = c->graph_launches; t->graph_rebuilds = c->graph_rebuilds;
	return SPB6_OK;
}


// ---- learned guess: dumps of (features, target) pairs from the running solver
int spb6_dump_enable(spb6_ctx* c, long capacity, int stride) {
	if (!c || capacity < 0 || stride < 1) return SPB6_BAD_ARG;
	hipSetDevice(c->device);
	if (c->dX) { hipFree(c->dX); hipFree(c->dY); hipFree(c->dcount); c->dX = nullptr; c->dY = nullptr; c->dcount = nullptr; }
	c->dcap = capacity; c->dstride = stride;
	if (capacity == 0) return SPB6_OK;
	HIPCHK(hipMalloc((void**)&c->dX, capacity * MLP_IN * sizeof(float)));
	HIPCHK(hipMalloc((void**)&c->dY, capacity * sizeof(float)));
	HIPCHK(hipMalloc((void**)&c->dcount, sizeof(unsigned long long)));
	HIPCHK(hipMemset(c->dcount, 0, sizeof(unsigned long long)));
	return SPB6_OK;
}
long spb6_dump_fetch(spb6_ctx* c, float* X, float* Y, long max_rows) {
	if (!c || !c->dX) return -1;
	hipSetDevice(c->device);
	unsigned long long cnt = 0;
	if (hipMemcpy(&cnt, c->dcount, sizeof cnt, hipMemcpyDeviceToHost) != hipSuccess) return -1;
	long rows = (long)std::min<unsigned long long>(cnt, (unsigned long long)c->dcap);
	rows = std::min(rows, max_rows);
	if (X && rows > 0 && hipMemcpy(X, c->dX, rows * MLP_IN * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return -1;
	if (Y && rows > 0 && hipMemcpy(Y, c->dY, rows * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return -1;
	hipMemset(c->dcount, 0, sizeof(unsigned long long));
	return rows;
}
int spb6_set_guess_model(spb6_ctx* c, spb6_mlp* m) {
	if (!c) return SPB6_BAD_ARG;
	c->model = m;
	return SPB6_OK;
}

// ---- the network: FP32 training on the device, BF16 MFMA inference
static int mlp_alloc(spb6_mlp* m, void** p, size_t bytes) {
	HIPCHK(hipMalloc(p, std::max<size_t>(bytes, 4)));
	m->allocs.push_back(*p);
	HIPCHK(hipMemset(*p, 0, std::max<size_t>(bytes, 4)));
	return SPB6_OK;
}
spb6_mlp* spb6_mlp_create(int device) {
	spb6_mlp* m = new spb6_mlp();
	m->device = device;
	auto fail = [&]() -> spb6_mlp* { spb6_mlp_destroy(m); return nullptr; };
	if (hipSetDevice(device) != hipSuccess) return fail();
	if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess) return fail();
	if (mlp_alloc(m, (void**)&m->P, MLP_NPARAM * 4) || mlp_alloc(m, (void**)&m->G, MLP_NPARAM * 4) || mlp_alloc(m, (void**)&m->m, MLP_NPARAM * 4) ||
	    mlp_alloc(m, (void**)&m->v, MLP_NPARAM * 4) || mlp_alloc(m, (void**)&m->mean, MLP_IN * 4) || mlp_alloc(m, (void**)&m->istd, MLP_IN * 4) ||
	    mlp_alloc(m, (void**)&m->loss, 4) || mlp_alloc(m, (void**)&m->W1s, MLP_H * MLP_KPAD * 2) || mlp_alloc(m, (void**)&m->W1f, MLP_H * MLP_KPAD * 2) ||
	    mlp_alloc(m, (void**)&m->W2p, MLP_H * MLP_H * 2) || mlp_alloc(m, (void**)&m->mub1, (7 + MLP_H) * 4))
		return fail();
	return m;
}
void spb6_mlp_destroy(spb6_mlp* m) {
	if (!m) return;
	if (m->st) hipStreamSynchronize(m->st);
	for (void* p : m->allocs) hipFree(p);
	if (m->st) hipStreamDestroy(m->st);
	delete m;
}
static int mlp_sync_infer(spb6_mlp* m) {   // refresh the BF16 copy and the host scalars after the FP32 parameters change
	mlp_pack(m->P, m->mean, m->istd, m->W1s, m->W1f, m->W2p, m->mub1, m->st);
	HIPCHK(hipMemcpyAsync(&m->b3, m->P + MLP_OFF_B3, sizeof(float), hipMemcpyDeviceToHost, m->st));
	HIPCHK(hipStreamSynchronize(m->st));
	return SPB6_OK;
}
int spb6_mlp_set_params(spb6_mlp* m, const float* params, const float* norm) {
	if (!m || !params) return SPB6_BAD_ARG;
	hipSetDevice(m->device);
	HIPCHK(hipMemcpy(m->P, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
	if (norm) {
		std::vector<float> istd(MLP_IN);
		for (int k = 0; k < MLP_IN; k++) istd[k] = 1.f / std::max(norm[MLP_IN + k], 1e-12f);
		HIPCHK(hipMemcpy(m->mean, norm, MLP_IN * sizeof(float), hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(m->istd, istd.data(), MLP_IN * sizeof(float), hipMemcpyHostToDevice));
		m->mu_t = norm[2 * MLP_IN]; m->sd_t = norm[2 * MLP_IN + 1];
	}
	HIPCHK(hipMemset(m->m, 0, MLP_NPARAM * 4)); HIPCHK(hipMemset(m->v, 0, MLP_NPARAM * 4)); m->step = 0;
	return mlp_sync_infer(m);
}
int spb6_mlp_get_params(spb6_mlp* m, float* params) {
	if (!m || !params) return SPB6_BAD_ARG;
	hipSetDevice(m->device);
	HIPCHK(hipMemcpy(params, m->P, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
	return SPB6_OK;
}
int spb6_mlp_nparams(void) { return MLP_NPARAM; }
static int mlp_batch_bufs(spb6_mlp* m, int Bn) {
	if (Bn <= m->bcap) return SPB6_OK;
	for (float** p : {&m->Xb, &m->Yb, &m->ws, &m->yb}) if (*p) { hipFree(*p); *p = nullptr; }
	if (m->idx) { hipFree(m->idx); m->idx = nullptr; }
	HIPCHK(hipMalloc((void**)&m->Xb, (size_t)Bn * MLP_IN * 4)); HIPCHK(hipMalloc((void**)&m->Yb, (size_t)Bn * 4));
	HIPCHK(hipMalloc((void**)&m->ws, (size_t)mlp_ws_floats(Bn) * 4)); HIPCHK(hipMalloc((void**)&m->yb, (size_t)Bn * 4));
	HIPCHK(hipMalloc((void**)&m->idx, (size_t)Bn * 4));
	m->bcap = Bn;
	return SPB6_OK;
}
int spb6_mlp_set_data(spb6_mlp* m, const float* X, const float* Y, long n) {
	if (!m || !X || !Y || n <= 0) return SPB6_BAD_ARG;
	hipSetDevice(m->device);
	if (m->X) { hipFree(m->X); hipFree(m->Y); m->X = m->Y = nullptr; }
	HIPCHK(hipMalloc((void**)&m->X, (size_t)n * MLP_IN * 4)); HIPCHK(hipMalloc((void**)&m->Y, (size_t)n * 4));
	HIPCHK(hipMemcpy(m->X, X, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(m->Y, Y, (size_t)n * 4, hipMemcpyHostToDevice));
	m->ndata = n;
	return SPB6_OK;
}
// epochs of Adam over the device data set (standardised features and targets), mini-batches in a fresh random order
// each epoch (SplitMix64-seeded Fisher-Yates on the host); loss_out[e] = mean squared error of epoch e
int spb6_mlp_train(spb6_mlp* m, int epochs, int batch, float lr, unsigned long long seed, float* loss_out) {
	if (!m || !m->X || epochs < 1 || batch < 1) return SPB6_BAD_ARG;
	hipSetDevice(m->device);
	int rc = mlp_batch_bufs(m, batch); if (rc) return rc;
	std::vector<int> perm(m->ndata);
	for (long i = 0; i < m->ndata; i++) perm[i] = (int)i;
	unsigned long long s = seed;
	auto rnd = [&]() { s += 0x9E3779B97F4A7C15ull; unsigned long long z = s; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); };
	int* didx = nullptr;
	HIPCHK(hipMalloc((void**)&didx, (size_t)m->ndata * 4));
	for (int e = 0; e < epochs; e++) {
		for (long i = m->ndata - 1; i > 0; i--) { const long j = (long)(rnd() % (unsigned long long)(i + 1)); std::swap(perm[i], perm[j]); }
		HIPCHK(hipMemcpyAsync(didx, perm.data(), (size_t)m->ndata * 4, hipMemcpyHostToDevice, m->st));
		HIPCHK(hipMemsetAsync(m->loss, 0, 4, m->st));
		long seen = 0;
		for (long b0 = 0; b0 + batch <= m->ndata; b0 += batch) {
			mlp_gather(batch, m->X, m->Y, didx + b0, m->Xb, m->Yb, m->st);
			mlp_fwd_bwd(m->P, m->Xb, m->Yb, batch, m->ws, m->G, m->yb, m->loss, m->st);
			mlp_adam(m->P, m->G, m->m, m->v, lr, ++m->step, m->st);
			seen += batch;
		}
		float l = 0.f;
		HIPCHK(hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, m->st));
		HIPCHK(hipStreamSynchronize(m->st));
		if (loss_out) loss_out[e] = seen ? l / (float)seen : 0.f;
	}
	hipFree(didx);
	return mlp_sync_infer(m);
}
// FP32 forward on host features (standardised), y[n]
int spb6_mlp_forward(spb6_mlp* m, const float* X, int n, float* y) {
	if (!m || !X || !y || n < 1) return SPB6_BAD_ARG;
	hipSetDevice(m->device);
	int rc = mlp_batch_bufs(m, n); if (rc) return rc;
	HIPCHK(hipMemcpyAsync(m->Xb, X, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, m->st));
	mlp_fwd_bwd(m->P, m->Xb, nullptr, n, m->ws, nullptr, m->yb, nullptr, m->st);
	HIPCHK(hipMemcpyAsync(y, m->yb, (size_t)n * 4, hipMemcpyDeviceToHost, m->st));
	HIPCHK(hipStreamSynchronize(m->st));
	return SPB6_OK;
}
// gradient of the batch loss mean((y - t)^2) with respect to every parameter (FP32 backward kernels)
int spb6_mlp_gradients(spb6_mlp* m, const float* X, const float* Y, int n, float* grad, float* loss) {
	if (!m || !X || !Y || !grad || n < 1) return SPB6_BAD_ARG;
	hipSetDevice(m->device);
	int rc = mlp_batch_bufs(m, n); if (rc) return rc;
	HIPCHK(hipMemcpyAsync(m->Xb, X, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, m->st));
	HIPCHK(hipMemcpyAsync(m->Yb, Y, (size_t)n * 4, hipMemcpyHostToDevice, m->st));
	HIPCHK(hipMemsetAsync(m->loss, 0, 4, m->st));
	mlp_fwd_bwd(m->P, m->Xb, m->Yb, n, m->ws, m->G, m->yb, m->loss, m->st);
	HIPCHK(hipMemcpyAsync(grad, m->G, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->st));
	float l = 0.f;
	HIPCHK(hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, m->st));
	HIPCHK(hipStreamSynchronize(m->st));
	if (loss) *loss = l / (float)n;
	return SPB6_OK;
}
// BF16 MFMA inference on host features (standardised), y[n]
int spb6_mlp_infer_bf16(spb6_mlp* m, const float* X, int n, float* y) {
	if (!m || !X || !y || n < 1) return SPB6_BAD_ARG;
	hipSetDevice(m->device);
	int rc = mlp_batch_bufs(m, n); if (rc) return rc;
	HIPCHK(hipMemcpyAsync(m->Xb, X, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice, m->st));
	MlpDev W{m->P, m->W1s, m->W1f, m->W2p, m->mub1, m->b3, m->mu_t, m->sd_t};
	mlp_infer(nullptr, nullptr, 0.0, n, m->Xb, W, nullptr, nullptr, 0.0, nullptr, m->yb, m->st);
	HIPCHK(hipMemcpyAsync(y, m->yb, (size_t)n * 4, hipMemcpyDeviceToHost, m->st));
	HIPCHK(hipStreamSynchronize(m->st));
	return SPB6_OK;
}

}   // extern "C"
