// =====================================================================================================================
// spb6_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement (C++17 + OpenMP, FP64) of the semi-implicit IMEX-RK time step of diffusionIMEX's SPB6_PC_IMEX /
// SPB6_PC_IMEX_omp solvers, used as the parity oracle of the HIP path and as the timed CPU baseline. Nothing in the
// product library links, includes or calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load it (through ctypes).
//
// Scope: stiffened-gas phases, relaxation level 0 (instantaneous p closure, CL_P), MUSCL-minmod + Rusanov (local speed),
// LSDIRK2 / backward Euler, transmissive / wall / periodic boundaries -- the hot path. Each function cites the reference
// function it restates ("PC" = src/SPB6_PC_IMEX.cpp, "PCO" = src/SPB6_PC_IMEX_omp.cpp, "PHY" = src/spb6_physics.h,
// "PHO" = src/spb6_physics_omp.h). Arithmetic is written operation-for-operation in the reference's evaluation order and
// must be compiled with -ffp-contract=off (no FMA contraction), like the reference's x86-64 -O2 build.
//
// Selectable variants (OraOpts):
//   closure_path : 0 = PHY closure_p (bracket first, PHY:462-494), 1 = PHO closure_p (Newton from the guess, PHO:472-548)
//   share_start  : 0 = PHY shock_share bracket [-40, 40] (PHY:363), 1 = PHO weak-shock-centred bracket (PHO:365-374)
//   dot_order    : 0 = block_sum B=4096 (PC:717-722), 1 = B=1024 (PCO:744-749), 2 = the HIP kernels' fixed tree order
//   portable_math: 0 = libm exp/log/pow/hypot (as the reference), 1 = the project's portable routines (bit-identical on
//                  host and device; used for GPU-vs-oracle bit parity)
//
// Parity pins (tests/test_oracle_*.py): the reference's archived full-precision profiles under data/ (cases 4, 5, 80-88),
// closed forms (EOS, weak-shock share, Hugoniot), invariants (closure root, energy consistency, conservation), brute-force
// bisection for closure_p, and a dense solve for the Thomas / BiCGStab linear algebra.
// =====================================================================================================================
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <vector>
#include <algorithm>
#include <limits>
#include <chrono>
#include <string>
#include <omp.h>

namespace ora {

static const double INF = std::numeric_limits<double>::infinity();
enum { IA1 = 0, IM1 = 1, IM2 = 2, IMU = 3, IMV = 4, IE1 = 5, IE2 = 6, IEN = 7 };   // PHY:25-26
static const int NV = 8, GH = 2;                                                   // PHY:25, PC:55
enum { BC_TRANS = 0, BC_WALL = 1, BC_PERIODIC = 2 };                               // PHY:906

// std::max / std::min semantics (first argument returned on ties and on NaN in the second comparison operand)
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

// ------------------------------------------------------------------------------------------------ portable math
// Used in place of libm where the HIP kernels must produce the same bits (shock_share, reconstruction scales).
// exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-13 Taylor polynomial in Horner form, exact 2^k scaling.
// log: x = 2^e m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(f/(2+f)) by its odd series, plus e ln2 in two parts.
// pow(x, y) = exp(y log x) for x > 0. All accurate to a few ulp (checked against libm in tests).
static const double LN2_HI = 6.93147180369123816490e-01;   // ln 2 rounded to 32 significant bits
static const double LN2_LO = 1.90821492927058770002e-10;   // ln 2 - LN2_HI
static const double INV_LN2 = 1.44269504088896338700e+00;
static double pexp(double x) {
	if (x != x) return x;
	if (x > 709.782712893384) return INF;
	if (x < -745.1332191019412) return 0.0;
	const double k = std::floor(x * INV_LN2 + 0.5);
	const double r = (x - k * LN2_HI) - k * LN2_LO;
	double p = 1.0 / 6227020800.0;                 // 1/13!
	p = 1.0 / 479001600.0 + r * p;                 // 1/12!
	p = 1.0 / 39916800.0 + r * p;
	p = 1.0 / 3628800.0 + r * p;
	p = 1.0 / 362880.0 + r * p;
	p = 1.0 / 40320.0 + r * p;
	p = 1.0 / 5040.0 + r * p;
	p = 1.0 / 720.0 + r * p;
	p = 1.0 / 120.0 + r * p;
	p = 1.0 / 24.0 + r * p;
	p = 1.0 / 6.0 + r * p;
	p = 0.5 + r * p;
	p = 1.0 + r * p;
	p = 1.0 + r * p;
	const int ki = (int)k;
	if (ki > 1000) return std::ldexp(std::ldexp(p, ki - 1000), 1000);
	if (ki < -1000) return std::ldexp(std::ldexp(p, ki + 1000), -1000);
	return std::ldexp(p, ki);
}
static double plog(double x) {
	if (x != x) return x;
	if (x < 0.0) return std::numeric_limits<double>::quiet_NaN();
	if (x == 0.0) return -INF;
	if (x == INF) return INF;
	int e; double m = std::frexp(x, &e);           // x = m 2^e, m in [0.5, 1)
	if (m < 0.70710678118654752440) { m = m + m; e = e - 1; }
	const double f = m - 1.0, s = f / (2.0 + f), z = s * s;
	double R = 1.0 / 23.0;
	R = 1.0 / 21.0 + z * R;
	R = 1.0 / 19.0 + z * R;
	R = 1.0 / 17.0 + z * R;
	R = 1.0 / 15.0 + z * R;
	R = 1.0 / 13.0 + z * R;
	R = 1.0 / 11.0 + z * R;
	R = 1.0 / 9.0 + z * R;
	R = 1.0 / 7.0 + z * R;
	R = 1.0 / 5.0 + z * R;
	R = 1.0 / 3.0 + z * R;
	R = z * R;                                     // log m = 2 s + 2 s R, and 2 s = f - s f since s (2 + f) = f
	const double logm = f - s * (f - 2.0 * R);
	const double de = (double)e;
	return de * LN2_HI + (logm + de * LN2_LO);
}
static double ppow(double x, double y) { return pexp(y * plog(x)); }
// e^x - 1: the Taylor series to x^17 / 17! by Horner on |x| < 0.5 (the next term is below 2^-70 relative), else
// pexp(x) - 1 (no cancellation beyond 2 bits there); shared bit for bit with the device
static double pexpm1(double x) {
	if (!(std::fabs(x) < 0.5)) return pexp(x) - 1.0;
	double f = 1.0; for (int k = 2; k <= 17; k++) f *= (double)k;   // 17! is exact in FP64
	double q = 1.0 / f;
	for (int k = 16; k >= 1; k--) { f /= (double)(k + 1); q = q * x + 1.0 / f; }
	return q * x;
}
static double phypot(double a, double b) { return std::sqrt(a * a + b * b); }

// ------------------------------------------------------------------------------------------------ options / model
// PHY:49-53: kind 0 is NASG (b = 0: SG) with g, pi, q, qp, cv, b; kind 1 (Cochran-Chan) and 2 (JWL) are Mie-Grueneisen
// phases p = G rho (e - er(rho)) + pr(rho) with G, c1, c2, e1, e2, rr (reference density) and e0 (JWL's energy shift)
struct Phase { double g, pi, q, qp, cv, b; int kind = 0; double G = 0, c1 = 0, c2 = 0, e1 = 0, e2 = 0, rr = 0, e0 = 0; };

// the saturation line T_sat(p) of the two phases (PHY:111-163): ln p on a uniform grid and the root T there
struct SatTable { std::vector<double> lp, T; double pcrit = std::numeric_limits<double>::infinity(), vcrit = 0.0, pmin = 0.0; };

struct Model {
	Phase PH[2];
	const SatTable* sat = nullptr;   // built for relaxation levels 2-3 (T_sat of the pTg closure)
	int closure_path = 1, share_start = 1, dot_order = 2, portable = 1;
	int shock_fused = 0;   // the strong-shock profile integrations in fused form (prof_integrate_fused)
	// the thermal and chemical relaxation rates (PHY:33; resolved by level as PC:1148-1149: inf = instantaneous)
	double theta_rate = 0.0, nu_rate = 0.0;
	double mu_rate = std::numeric_limits<double>::infinity();   // the mechanical relaxation rate (PHY:33: inf = instantaneous)
	int nu_law = 0; double nu_cr = 1.0, nu_pc = 22.064e6;           // the state-dependent transfer rate (NU_LAW, NU_CR, NU_PC; PHY:36)
	int relax_form = 0;   // RELAX_FORM (PHY:39-41): 0 backward Euler on the stage's step, 1 the exponential solution
	double exp_(double x) const { return portable ? pexp(x) : std::exp(x); }
	double expm1_(double x) const { return portable ? pexpm1(x) : std::expm1(x); }
	double log_(double x) const { return portable ? plog(x) : std::log(x); }
	double pow_(double x, double y) const { return portable ? ppow(x, y) : std::pow(x, y); }
	double hypot_(double a, double b) const { return portable ? phypot(a, b) : std::hypot(a, b); }
	bool is_mg(int k) const { return PH[k].kind != 0; }                                            // PHY:55-56
	bool mg_any() const { return is_mg(0) || is_mg(1); }
};
// the Mie-Grueneisen reference curve at density r: pr(r), pr'(r), er(r) (PHY:58-66). Cochran-Chan: pr = c1 x^e1 - c2 x^e2
// with x = r / rr, er its integral with er' = pr / r^2; JWL: pr = c1 exp(-e1 rr / r) + c2 exp(-e2 rr / r), er shifted by e0
static void mg_ref(const Model& M, int k, double r, double& pr, double& dpr, double& er) {
	const Phase& P = M.PH[k];
	if (P.kind == 1) {
		const double x = r / P.rr, t1 = P.c1 * M.pow_(x, P.e1), t2 = P.c2 * M.pow_(x, P.e2);
		pr = t1 - t2; dpr = (P.e1 * t1 - P.e2 * t2) / r; er = (t1 / (P.e1 - 1.0) - t2 / (P.e2 - 1.0)) / r;
	} else {
		const double t1 = P.c1 * M.exp_(-P.e1 * P.rr / r), t2 = P.c2 * M.exp_(-P.e2 * P.rr / r);
		pr = t1 + t2; dpr = (P.e1 * t1 + P.e2 * t2) * P.rr / (r * r); er = t1 / (P.e1 * P.rr) + t2 / (P.e2 * P.rr) - P.e0;
	}
}

// ------------------------------------------------------------------------------------------------ EOS (PHY:69-109)
static inline double eps_amp(const Model& M, int k, double a, double m, double p) { const Phase& P = M.PH[k];   // PHY:69-71
	if (P.kind) { double pr, dpr, er; mg_ref(M, k, m / a, pr, dpr, er); return m * er + a * (p - pr) / P.G; }
	return (p + P.g * P.pi) * (a - m * P.b) / (P.g - 1.0) + m * P.q; }
static inline double p_ame(const Model& M, int k, double a, double m, double e) { const Phase& P = M.PH[k];     // PHY:72-74
	if (P.kind) { double pr, dpr, er; mg_ref(M, k, m / a, pr, dpr, er); return P.G * (e - m * er) / a + pr; }
	return (e - m * P.q) * (P.g - 1.0) / (a - m * P.b) - P.g * P.pi; }
static inline double T_amp(const Model& M, int k, double a, double m, double p) { const Phase& P = M.PH[k];     // PHY:75-77
	if (P.kind) { if (!(P.cv > 0)) return NAN; double pr, dpr, er; mg_ref(M, k, m / a, pr, dpr, er); return (a * (p - pr) / (P.G * m)) / P.cv; }
	return (a / m - P.b) * (p + P.pi) / ((P.g - 1.0) * P.cv); }
static inline double c2_rp(const Model& M, int k, double r, double p) { const Phase& P = M.PH[k];               // PHY:83-85
	if (P.kind) { double pr, dpr, er; mg_ref(M, k, r, pr, dpr, er); return (P.G + 1.0) * (p - pr) / r + dpr; }
	return P.g * (p + P.pi) / (r * (1.0 - r * P.b)); }
static inline double dedp_am(const Model& M, int k, double a, double m) { const Phase& P = M.PH[k];             // PHY:86-88
	if (P.kind) return a / P.G;
	return (a - m * P.b) / (P.g - 1.0); }
static double p_mix(const Model& M, double a1, double m1, double m2, double rhoe) {                 // PHY:90-99
	const double a[2] = {a1, 1.0 - a1}, m[2] = {m1, m2};
	double num = rhoe, den = 0.0;
	for (int k = 0; k < 2; k++) {
		if (M.is_mg(k)) { double pr, dpr, er; mg_ref(M, k, m[k] / a[k], pr, dpr, er); num -= m[k] * er - a[k] * pr / M.PH[k].G; }
		else num -= m[k] * M.PH[k].q + M.PH[k].g * M.PH[k].pi * dedp_am(M, k, a[k], m[k]);
		den += dedp_am(M, k, a[k], m[k]);
	}
	return num / den;
}
static inline bool p_admissible(const Model& M, int k, double a, double m, double p) {              // PHY:101-104
	if (M.is_mg(k)) return std::isfinite(p) && c2_rp(M, k, m / a, p) > 0;
	return p + M.PH[k].pi > 0; }
static inline bool p_adm_sg(const Model& M, int k, double p) { return p + M.PH[k].pi > 0; }         // the explicit scheme's (SG / NASG)
static inline double eps_round(const Model& M, int k, double a, double m) {                         // PHY:106-109
	if (M.is_mg(k)) { double pr, dpr, er; mg_ref(M, k, m / a, pr, dpr, er); return std::fabs(m * er) + std::fabs(a * pr / M.PH[k].G); }
	return std::fabs(m * M.PH[k].q); }

// ------------------------------------------------------------------------------------------------ primitives (PHY:835-857)
struct Prim { double a1, r1, r2, u, v, p1, p2, p, rho, cf; };
static bool prim_of(const Model& M, const double* Q, Prim& w) {                                     // PHY:837-847
	const double a1 = Q[IA1], a2 = 1 - a1, m1 = Q[IM1], m2 = Q[IM2];
	if (!(a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0)) return false;
	w.a1 = a1; w.rho = m1 + m2; w.r1 = m1 / a1; w.r2 = m2 / a2; w.u = Q[IMU] / w.rho; w.v = Q[IMV] / w.rho;
	w.p = p_mix(M, a1, m1, m2, Q[IEN] - 0.5 * w.rho * (w.u * w.u + w.v * w.v));
	w.p1 = p_ame(M, 0, a1, m1, Q[IE1]); w.p2 = p_ame(M, 1, a2, m2, Q[IE2]);
	if (!(p_admissible(M, 0, a1, m1, w.p1) && p_admissible(M, 1, a2, m2, w.p2))) return false;
	w.cf = std::sqrt((m1 * c2_rp(M, 0, w.r1, w.p1) + m2 * c2_rp(M, 1, w.r2, w.p2)) / w.rho);
	return true;
}
static void prim2cell(const Model& M, const Prim& w, double* Q) {                                   // PHY:849-854
	const double a2 = 1 - w.a1, m1 = w.a1 * w.r1, m2 = a2 * w.r2, rho = m1 + m2;
	Q[IA1] = w.a1; Q[IM1] = m1; Q[IM2] = m2; Q[IMU] = rho * w.u; Q[IMV] = rho * w.v;
	Q[IE1] = eps_amp(M, 0, w.a1, m1, w.p1); Q[IE2] = eps_amp(M, 1, a2, m2, w.p2);
	Q[IEN] = Q[IE1] + Q[IE2] + 0.5 * rho * (w.u * w.u + w.v * w.v);
}
static inline double pbar_of(const Model& M, const double* Q) {                                     // PHY:855-857
	const double a1 = Q[IA1];
	return a1 * p_ame(M, 0, a1, Q[IM1], Q[IE1]) + (1 - a1) * p_ame(M, 1, 1 - a1, Q[IM2], Q[IE2]);
}

// ------------------------------------------------------------------------------------------------ shock budget (PHY:260-418)
// de/dp at fixed v and de/dv at fixed p plus p, of phase k at (p, v) (PHY:266-270)
static inline void eos_ep_ev(const Model& M, int k, double p, double v, double& ep, double& evp) {
	const Phase& P = M.PH[k];
	if (P.kind) { double pr, dpr, er; const double r = 1.0 / v; mg_ref(M, k, r, pr, dpr, er); ep = v / P.G; evp = (p - pr) * (1.0 + 1.0 / P.G) + r * dpr / P.G; return; }
	ep = (v - P.b) / (P.g - 1.0); evp = P.g * (p + P.pi) / (P.g - 1.0); }
static inline bool eos_ok(const Model& M, int k, double p, double v) {                             // PHY:271-275
	if (!(v > 0) || !std::isfinite(p)) return false;
	if (M.is_mg(k)) return c2_rp(M, k, 1.0 / v, p) > 0;
	return v > M.PH[k].b && p + M.PH[k].pi > 0; }
// phase k's Hugoniot from (p0, v0) to p1 (PHY:277-292): closed form for NASG; for a Mie-Grueneisen phase the root of
// e(p1, v) - e(p0, v0) + (p0 + p1) / 2 (v - v0) below v0, bracketed by steps of 0.9 v and bisected to 1e-14 v0
static double hug_v(const Model& M, int k, double p0, double p1, double v0) {
	const Phase& P = M.PH[k];
	if (!P.kind) {
		const double x0 = p0 + P.pi, x1 = p1 + P.pi;
		return P.b + (v0 - P.b) * ((P.g - 1.0) * x1 + (P.g + 1.0) * x0) / ((P.g + 1.0) * x1 + (P.g - 1.0) * x0);
	}
	auto e_pv = [&](double p, double v) { double pr, dpr, er; mg_ref(M, k, 1.0 / v, pr, dpr, er); return er + v * (p - pr) / P.G; };
	const double e0 = e_pv(p0, v0);
	auto hg = [&](double v) { return e_pv(p1, v) - e0 + 0.5 * (p0 + p1) * (v - v0); };
	double hi = v0, lo = v0, gl = hg(v0);
	if (!(gl > 0)) return v0;
	for (int it = 0; it < 80 && gl > 0; it++) { hi = lo; lo *= 0.9; gl = hg(lo); if (!std::isfinite(gl)) return NAN; }
	if (gl > 0) return NAN;
	for (int it = 0; it < 100; it++) {
		const double mid = 0.5 * (lo + hi), gm = hg(mid);
		if (gm > 0) hi = mid; else lo = mid;
		if (hi - lo <= 1e-14 * v0) break;
	}
	return 0.5 * (lo + hi);
}
// isentropic bulk modulus K and fundamental derivative G = (1 + dK/dp|_s) / 2 (PHY:294-299): closed form for NASG, a
// one-sided difference along the isentrope dv = -v dp / K (dp = 1e-6 K) for a Mie-Grueneisen phase
static void K_G(const Model& M, int k, double p, double v, double& K, double& G) {
	K = c2_rp(M, k, 1.0 / v, p) / v;
	if (!M.is_mg(k)) { G = 0.5 * (M.PH[k].g + 1.0); if (M.PH[k].b != 0.0) G *= v / (v - M.PH[k].b); return; }
	const double dp = 1e-6 * K, K2 = c2_rp(M, k, 1.0 / (v * (1.0 - dp / K)), p + dp) / (v * (1.0 - dp / K));
	G = 0.5 * (1.0 + (K2 - K) / dp);
}
struct ShockProf { double Y[2], vk0[2], v0, p0, j2, vend; };
static bool prof_rhs(const Model& M, const ShockProf& S, double c, double v, const double* y, double* dy) {   // PHY:302-310
	const double p = y[0], v1 = y[1], v2 = (v - S.Y[0] * v1) / S.Y[1], P = S.p0 + S.j2 * (S.v0 - v);
	if (!eos_ok(M, 0, p, v1) || !eos_ok(M, 1, p, v2)) return false;
	const double lam[2] = {c, 1.0 - c}, vk[2] = {v1, v2};
	double num = -1.0, den = 0.0, r[2], ep[2], evp[2];
	for (int k = 0; k < 2; k++) {
		eos_ep_ev(M, k, p, vk[k], ep[k], evp[k]);
		r[k] = lam[k] / S.Y[k] * (p - P); num += S.Y[k] * r[k] / evp[k]; den += S.Y[k] * ep[k] / evp[k];
	}
	dy[0] = num / den; dy[1] = (r[0] - ep[0] * dy[0]) / evp[0]; dy[2] = p - P;
	return std::isfinite(dy[0]) && std::isfinite(dy[1]);
}
// Cash-Karp embedded RK4(5) coefficients (Cash & Karp 1990)
static const double CK_A[6] = {0, 0.2, 0.3, 0.6, 1.0, 0.875};
static const double CK_B[6][5] = {{0}, {0.2}, {3.0 / 40, 9.0 / 40}, {0.3, -0.9, 1.2}, {-11.0 / 54, 2.5, -70.0 / 27, 35.0 / 27},
	{1631.0 / 55296, 175.0 / 512, 575.0 / 13824, 44275.0 / 110592, 253.0 / 4096}};
static const double CK_C5[6] = {37.0 / 378, 0, 250.0 / 621, 125.0 / 594, 0, 512.0 / 1771};
static const double CK_C4[6] = {2825.0 / 27648, 0, 18575.0 / 48384, 13525.0 / 55296, 277.0 / 14336, 0.25};
static bool prof_integrate(const Model& M, const ShockProf& S, double c, double& v1end, double& qend) {   // PHY:311-341
	double y[3] = {S.p0, S.vk0[0], 0.0}, v = S.v0;
	const double L = S.vend - S.v0, sc[3] = {std::fabs(S.p0) + S.j2 * std::fabs(L), S.vk0[0], S.j2 * L * L + 1e-300};
	double h = L / 32.0;
	for (int it = 0; it < 4000; it++) {
		const bool last = (L < 0 && v + h <= S.vend) || (L > 0 && v + h >= S.vend);
		if (last) h = S.vend - v;
		double k[6][3], yt[3];
		bool ok = true;
		for (int st = 0; st < 6 && ok; st++) {
			for (int n = 0; n < 3; n++) { yt[n] = y[n]; for (int q = 0; q < st; q++) yt[n] += h * CK_B[st][q] * k[q][n]; }
			ok = prof_rhs(M, S, c, v + CK_A[st] * h, yt, k[st]);
		}
		double err = 0.0, y5[3];
		if (ok) for (int n = 0; n < 3; n++) {
			double d5 = 0, d4 = 0; for (int st = 0; st < 6; st++) { d5 += CK_C5[st] * k[st][n]; d4 += CK_C4[st] * k[st][n]; }
			y5[n] = y[n] + h * d5; err = smax(err, std::fabs(h * (d5 - d4)) / sc[n]);
		}
		if (!ok || !(err <= 1e-8)) {
			h *= ok ? smax(0.1, 0.9 * M.pow_(1e-8 / err, 0.25)) : 0.25;
			if (std::fabs(h) < 1e-14 * std::fabs(L)) return false;
			continue;
		}
		for (int n = 0; n < 3; n++) y[n] = y5[n];
		if (last) { v1end = y[1]; qend = y[2]; return true; }
		v += h;
		h *= smin(5.0, 0.9 * M.pow_(1e-8 / smax(err, 1e-30), 0.2));
	}
	return false;
}
// The fused form of the profile integration (shock_fused = 1), as the strong-shock kernel's prof_step_fused computes
// it: the ODE of PHY:302-310 with its two phase terms over their common denominator, x_k = g_k (p + pi_k),
// n_k = v_k - b_k, lr_k = lam_k / Y_k, e_k = Y_k ((g_k - 1) lr_k), E0 = (g_0 - 1) lr_0:
//     dp/dv = num / den,  num = -x0 x1 + (p - P)(e0 x1 + e1 x0),  den = Y0 n0 x1 + Y1 n1 x0,
//     dv1/dv = (E0 (p - P) den - n0 num) / (den x0),
// both from the one correctly rounded division 1 / (den x0); v2 = (v - Y0 v1) (1 / Y1); the stage inputs, the fifth-
// and fourth-order sums (zero weights left out) and y5 as FMA chains in the Cash-Karp term order; the error norm
// multiplies by 1 / sc_k. The step control is prof_integrate's. Not the reference's rounding: this form is held to the
// HIP kernel bit for bit and to prof_rhs / prof_integrate within round-off (tests/test_oracle_shock_fused.py).
struct ProfFused { double rY1, rs[3], E0, e0, e1; };
static bool prof_rhs_fused(const Model& M, const ShockProf& S, const ProfFused& F, double v, const double* y, double* dy) {
	const Phase &A = M.PH[0], &B = M.PH[1];
	const double p = y[0], v1 = y[1], v2 = std::fma(-S.Y[0], v1, v) * F.rY1;
	const double P = std::fma(S.j2, S.v0 - v, S.p0), dp = p - P;
	const bool adm = eos_ok(M, 0, p, v1) && eos_ok(M, 1, p, v2);
	const double n0 = v1 - A.b, n1 = v2 - B.b, x0 = A.g * (p + A.pi), x1 = B.g * (p + B.pi);
	const double den = std::fma(S.Y[0] * n0, x1, (S.Y[1] * n1) * x0);
	const double num = std::fma(dp, std::fma(F.e0, x1, F.e1 * x0), -(x0 * x1));
	const double D = 1.0 / (den * x0);
	dy[0] = (num * x0) * D; dy[1] = std::fma(F.E0 * dp, den, -(n0 * num)) * D; dy[2] = dp;
	return adm && std::isfinite(dy[0]) && std::isfinite(dy[1]);
}
static void prof_fused_setup(const Model& M, const ShockProf& S, double c, ProfFused& F) {
	const double L = S.vend - S.v0;
	F.rY1 = 1.0 / S.Y[1];
	F.rs[0] = 1.0 / (std::fabs(S.p0) + S.j2 * std::fabs(L)); F.rs[1] = 1.0 / S.vk0[0]; F.rs[2] = 1.0 / (S.j2 * L * L + 1e-300);
	const double lr0 = c / S.Y[0], lr1 = (1.0 - c) / S.Y[1];
	F.E0 = (M.PH[0].g - 1.0) * lr0; F.e0 = S.Y[0] * F.E0; F.e1 = S.Y[1] * ((M.PH[1].g - 1.0) * lr1);
}
static bool prof_integrate_fused(const Model& M, const ShockProf& S, double c, double& v1end, double& qend) {
	ProfFused F; prof_fused_setup(M, S, c, F);
	double y[3] = {S.p0, S.vk0[0], 0.0}, v = S.v0;
	const double L = S.vend - S.v0;
	double h = L / 32.0;
	static const int W5[3] = {2, 3, 5}, W4[4] = {2, 3, 4, 5};   // the nonzero weights after the first
	for (int it = 0; it < 4000; it++) {
		const bool last = (L < 0 && v + h <= S.vend) || (L > 0 && v + h >= S.vend);
		if (last) h = S.vend - v;
		double k[6][3], yt[3];
		bool ok = true;
		for (int st = 0; st < 6; st++) {
			for (int n = 0; n < 3; n++) { yt[n] = y[n]; for (int q = 0; q < st; q++) yt[n] = std::fma(h * CK_B[st][q], k[q][n], yt[n]); }
			ok = prof_rhs_fused(M, S, F, v + CK_A[st] * h, yt, k[st]) && ok;
		}
		double err = 0.0, y5[3];
		for (int n = 0; n < 3; n++) {
			double d5 = CK_C5[0] * k[0][n], d4 = CK_C4[0] * k[0][n];
			for (int st : W5) d5 = std::fma(CK_C5[st], k[st][n], d5);
			for (int st : W4) d4 = std::fma(CK_C4[st], k[st][n], d4);
			y5[n] = std::fma(h, d5, y[n]); err = smax(err, std::fabs(h * (d5 - d4)) * F.rs[n]);
		}
		if (!ok || !(err <= 1e-8)) {
			h *= ok ? smax(0.1, 0.9 * M.pow_(1e-8 / err, 0.25)) : 0.25;
			if (std::fabs(h) < 1e-14 * std::fabs(L)) return false;
			continue;
		}
		for (int n = 0; n < 3; n++) y[n] = y5[n];
		if (last) { v1end = y[1]; qend = y[2]; return true; }
		v += h;
		h *= smin(5.0, 0.9 * M.pow_(1e-8 / smax(err, 1e-30), 0.2));
	}
	return false;
}
static bool shock_share(const Model& M, const double Y[2], const double vpre[2], double p0, double p1, double lam[2], double& q) {   // PHY:343-383, PHO:365-374
	double K[2], G[2], w[2], sw = 0.0, Kmin = INF;
	for (int k = 0; k < 2; k++) { K_G(M, k, p0, vpre[k], K[k], G[k]); if (!(K[k] > 0)) return false; Kmin = smin(Kmin, K[k]); }
	const double dp = p1 - p0;
	if (dp < 1e-2 * Kmin) {   // weak shock: closed form
		q = 0.0;
		for (int k = 0; k < 2; k++) { w[k] = Y[k] * vpre[k] * G[k] / (K[k] * K[k]); sw += w[k]; }
		for (int k = 0; k < 2; k++) lam[k] = w[k] / sw;
		q = sw * dp * dp * dp / 6.0;
		return true;
	}
	ShockProf S; S.p0 = p0;
	double vH[2];
	for (int k = 0; k < 2; k++) { S.Y[k] = Y[k]; S.vk0[k] = vpre[k]; vH[k] = hug_v(M, k, p0, p1, vpre[k]); if (!std::isfinite(vH[k])) return false; }
	S.v0 = Y[0] * vpre[0] + Y[1] * vpre[1]; S.vend = Y[0] * vH[0] + Y[1] * vH[1];
	if (!(S.vend < S.v0)) return false;
	S.j2 = dp / (S.v0 - S.vend);
	auto F = [&](double t, double& qq) -> double {
		const double c = 1.0 / (1.0 + M.exp_(-t)); double v1e;
		if (!(M.shock_fused ? prof_integrate_fused(M, S, c, v1e, qq) : prof_integrate(M, S, c, v1e, qq))) return NAN;
		return (v1e - vH[0]) / vpre[0]; };
	double ta, tb, qa, qb, fa, fb;
	if (M.share_start == 0) { ta = -40.0; tb = 40.0; fa = F(ta, qa); fb = F(tb, qb); if (!std::isfinite(fa) || !std::isfinite(fb)) return false; }
	else {
		double t0;
		{ double ww[2]; for (int k = 0; k < 2; k++) ww[k] = Y[k] * vpre[k] * G[k] / (K[k] * K[k]);
		  const double c0 = smin(1.0 - 1e-12, smax(1e-12, ww[0] / (ww[0] + ww[1])));
		  t0 = std::isfinite(c0) ? M.log_(c0 / (1.0 - c0)) : 0.0; }
		ta = smax(-40.0, t0 - 1.0); tb = smin(40.0, t0 + 1.0); fa = F(ta, qa); fb = F(tb, qb);
		if (!std::isfinite(fa) || !std::isfinite(fb)) return false;
		for (double st = 2.0; fa >= 0 && ta > -40.0; st *= 2.0) { tb = ta; fb = fa; qb = qa; ta = smax(-40.0, ta - st); fa = F(ta, qa); if (!std::isfinite(fa)) return false; }
		for (double st = 2.0; fb <= 0 && tb < 40.0; st *= 2.0) { ta = tb; fa = fb; qa = qb; tb = smin(40.0, tb + st); fb = F(tb, qb); if (!std::isfinite(fb)) return false; }
	}
	double t, qt;
	if (fa >= 0) { t = ta; qt = qa; }
	else if (fb <= 0) { t = tb; qt = qb; }
	else {   // Illinois on the logit of the share
		int side = 0; t = ta; qt = qa;
		for (int it = 0; it < 100; it++) {
			t = (ta * fb - tb * fa) / (fb - fa);
			if (!(t > ta && t < tb)) t = 0.5 * (ta + tb);
			const double ft = F(t, qt);
			if (!std::isfinite(ft)) return false;
			if (std::fabs(ft) < 1e-10 || tb - ta < 1e-10) break;
			if (ft < 0) { ta = t; fa = ft; if (side == -1) fb *= 0.5; side = -1; }
			else { tb = t; fb = ft; if (side == 1) fa *= 0.5; side = 1; }
		}
	}
	const double c = 1.0 / (1.0 + M.exp_(-t));
	lam[0] = c; lam[1] = 1.0 - c; q = qt;
	return q > 0;
}
static void stencil_state(const Model& M, const double* c, double* st) {                          // PHY:385-390
	const double a1 = c[IA1], m1 = c[IM1], m2 = c[IM2], r = m1 + m2;
	st[0] = a1; st[1] = 1.0 - a1; st[2] = m1; st[3] = m2; st[4] = NAN;
	st[5] = r > 0 ? c[IMU] / r : 0.0; st[6] = r > 0 ? c[IMV] / r : 0.0;
	if (a1 > 0 && a1 < 1 && m1 > 0 && m2 > 0) {
		const double p = p_mix(M, a1, m1, m2, c[IE1] + c[IE2]);
		if (p_admissible(M, 0, a1, m1, p) && p_admissible(M, 1, 1.0 - a1, m2, p)) st[4] = p;
	}
}
// cellptr(di, dj) returns a pointer to the 8 conserved values of the stencil cell
template <class F> static void shock_budget(const Model& M, F cellptr, bool twoD, double Q[2]) {  // PHY:395-418
	Q[0] = Q[1] = 0.0;
	double c0[7], st[7], lo[7] = {0, 0, 0, 0, 0, 0, 0}, plo = 1e300, phi = -1e300, div = 0.0;
	stencil_state(M, cellptr(0, 0), c0);
	for (int d = 0; d < (twoD ? 2 : 1); d++) for (int q = -1; q <= 1; q++) {
		if (d == 1 && q == 0) continue;
		stencil_state(M, cellptr(d == 0 ? q : 0, d == 1 ? q : 0), st);
		if (q == 1) div += st[5 + d]; else if (q == -1) div -= st[5 + d];
		if (!std::isfinite(st[4])) continue;
		if (st[4] < plo) { plo = st[4]; for (int n = 0; n < 7; n++) lo[n] = st[n]; }
		if (st[4] > phi) phi = st[4];
	}
	if (!(div < 0.0) || !(phi > plo)) return;
	double Y[2], vpre[2], lam[2], q;
	const double rho = c0[2] + c0[3];
	if (!(rho > 0)) return;
	for (int k = 0; k < 2; k++) {
		Y[k] = c0[2 + k] / rho;
		vpre[k] = (lo[k] > 0 && lo[2 + k] > 0) ? lo[k] / lo[2 + k] : (c0[2 + k] > 0 ? c0[k] / c0[2 + k] : 0.0);
	}
	if (!(Y[0] > 0) || !(Y[1] > 0) || !(vpre[0] > 0) || !(vpre[1] > 0)) return;
	if (!shock_share(M, Y, vpre, plo, phi, lam, q)) return;
	for (int k = 0; k < 2; k++) Q[k] = lam[k] * rho * q;
}

// ------------------------------------------------------------------------------------------------ p closure (PHY:420-494, PHO:472-548)
// closure types of PHY:226 used here: mechanical (p), thermal (pT; finite rate theta), chemical (pTg; finite rate nu)
enum { CL_P = 0, CL_MU = 1, CL_PT = 2, CL_THETA = 3, CL_NU = 4, CL_PTG = 5 };
// nodelta (PHY:234): the p closure without the energy defect, for the p-pTg activation's prediction only; tau: the
// stage's implicit step a_ii dt (the finite-rate closures)
struct ClIn { double atr[2], mtr[2], etr[2], s, rhoeC, pguess, Qph[2]; int type = CL_P; bool nodelta = false; double tau = 0.0;
               double Z[2] = {1.0, 1.0}; double nu = -1.0; };   // Z: the phases' frozen acoustic impedances (finite mu,
// PC:956-958); nu: the cell's transfer rate of the step (>= 0 replaces the model's, the nu law; PHY:241)
struct ClOut { double a[2], m[2], e[2], p[2], T[2], pbar; };
struct PCoef { double A[2], B[2], N[2], lam[2]; };
static double E_above(const Model& M, const ClIn& in, int k) {                                     // PHY:420-423
	if (M.is_mg(k)) {
		if (!(in.atr[k] > 0 && in.mtr[k] > 0)) return 0.0;
		double pr, dpr, er; mg_ref(M, k, in.mtr[k] / in.atr[k], pr, dpr, er); return smax(0.0, in.etr[k] - in.mtr[k] * er); }
	return smax(0.0, in.etr[k] - in.mtr[k] * M.PH[k].q - (in.atr[k] - in.mtr[k] * M.PH[k].b) * M.PH[k].pi); }
static void defect_split(const Model& M, const ClIn& in, double delta, double lam[2], double extra[2]) {   // PHY:428-439 (SHARE_MODE 0)
	double E[2], sE = 0.0; for (int k = 0; k < 2; k++) { E[k] = E_above(M, in, k); sE += E[k]; }
	double lE[2];
	if (sE > 0) for (int k = 0; k < 2; k++) lE[k] = E[k] / sE;
	else { const double sm = in.mtr[0] + in.mtr[1]; for (int k = 0; k < 2; k++) lE[k] = sm > 0 ? in.mtr[k] / sm : 0.5; }
	const double Q = in.Qph[0] + in.Qph[1];
	extra[0] = extra[1] = 0.0;
	if (!(delta > 0.0) || !(Q > 0.0)) { lam[0] = lE[0]; lam[1] = lE[1]; return; }
	if (delta < Q) { lam[0] = in.Qph[0] / Q; lam[1] = in.Qph[1] / Q; return; }
	for (int k = 0; k < 2; k++) { lam[k] = lE[k]; extra[k] = in.Qph[k] - Q * lE[k]; }
}
static bool p_coef(const Model& M, const ClIn& in, PCoef& c, double pseg) {                       // PHY:443-454
	if (!in.nodelta && !(in.mtr[0] + in.mtr[1] > 0)) return false;
	const double D0 = in.nodelta ? 0.0 : in.rhoeC - in.etr[0] - in.etr[1];
	double ex[2] = {0.0, 0.0};
	if (in.nodelta) c.lam[0] = c.lam[1] = 0.0; else defect_split(M, in, D0 + pseg * in.s, c.lam, ex);
	for (int k = 0; k < 2; k++) {
		c.A[k] = in.etr[k] - in.mtr[k] * M.PH[k].q + c.lam[k] * D0 + ex[k];
		c.B[k] = in.atr[k] + c.lam[k] * in.s;
		c.N[k] = c.A[k] - (c.B[k] - in.mtr[k] * M.PH[k].b) * M.PH[k].pi;
	}
	return true;
}
static inline double alpha_p(const Model& M, const PCoef& c, const ClIn& in, int k, double p) { const Phase& P = M.PH[k];   // PHY:455
	return ((P.g - 1.0) * (c.A[k] + c.B[k] * p) + (p + P.g * P.pi) * in.mtr[k] * P.b) / (P.g * (p + P.pi)); }
static inline double dalpha_p(const Model& M, const PCoef& c, int k, double p) { const Phase& P = M.PH[k];                   // PHY:456
	return -(P.g - 1.0) * c.N[k] / (P.g * (p + P.pi) * (p + P.pi)); }
static inline double plo_mix(const Model& M) { return smax(-M.PH[0].pi, -M.PH[1].pi); }            // PHY:458

static bool closure_out(const Model& M, const ClIn& in, double p, ClOut& o) {                      // PHY:487-493
	PCoef c; p_coef(M, in, c, p);
	for (int k = 0; k < 2; k++) {
		o.a[k] = alpha_p(M, c, in, k, p); o.m[k] = in.mtr[k];
		if (!(o.a[k] > o.m[k] * M.PH[k].b) || !(p + M.PH[k].pi > 0)) return false;
		o.e[k] = eps_amp(M, k, o.a[k], o.m[k], p); o.p[k] = p; o.T[k] = T_amp(M, k, o.a[k], o.m[k], p);
	}
	o.pbar = p; return std::isfinite(p);
}
static bool closure_p_phy(const Model& M, const ClIn& in, ClOut& o) {                              // PHY:462-494
	PCoef c; if (!p_coef(M, in, c, in.pguess)) return false;
	const double plo = plo_mix(M);
	auto Phi = [&](double p) { PCoef q; p_coef(M, in, q, p); return alpha_p(M, q, in, 0, p) + alpha_p(M, q, in, 1, p) - 1.0; };
	auto dPhi = [&](double p) { PCoef q; p_coef(M, in, q, p); return dalpha_p(M, q, 0, p) + dalpha_p(M, q, 1, p); };
	double d = smax(in.pguess - plo, 1e-3 * (std::fabs(plo) + 1.0));
	double lo = plo, hi = plo + d; double fh = Phi(hi);
	if (fh > 0) { lo = hi; for (int k = 0; k < 200 && fh > 0; k++) { d *= 4.0; hi = plo + d; fh = Phi(hi); } if (fh > 0) return false; }
	else { bool found = false; for (int k = 0; k < 200; k++) { const double dd = d * 0.25, pl = plo + dd, fl = Phi(pl); if (fl > 0) { lo = pl; found = true; break; } hi = pl; fh = fl; d = dd; } if (!found) return false; }
	double p = 0.5 * (lo + hi); if (in.pguess > lo && in.pguess < hi) p = in.pguess;
	for (int it = 0; it < 200; it++) {
		const double f = Phi(p);
		if (f == 0.0) break;
		if (f > 0) lo = p; else hi = p;
		const double df = dPhi(p);
		double pn = df < 0 ? p - f / df : 0.5 * (lo + hi);
		if (!(pn > lo && pn < hi)) pn = (hi - plo > 1e3 * (lo - plo + 1e-300) && lo > plo) ? plo + std::sqrt((lo - plo) * (hi - plo)) : 0.5 * (lo + hi);
		if (std::fabs(pn - p) <= 2e-16 * (std::fabs(p) + std::fabs(plo)) || hi - lo <= 4e-16 * (std::fabs(hi) + std::fabs(lo))) { p = pn; break; }
		p = pn;
	}
	return closure_out(M, in, p, o);
}
static bool closure_p_pho(const Model& M, const ClIn& in, ClOut& o) {                              // PHO:472-548
	if (!in.nodelta && !(in.mtr[0] + in.mtr[1] > 0)) return false;
	const double plo = plo_mix(M);
	const double D0 = in.nodelta ? 0.0 : in.rhoeC - in.etr[0] - in.etr[1];
	double lE[2], A0[2], sm = in.mtr[0] + in.mtr[1];
	{ double E[2], sE = 0.0; for (int k = 0; k < 2; k++) { E[k] = E_above(M, in, k); sE += E[k]; }
	  if (sE > 0) for (int k = 0; k < 2; k++) lE[k] = E[k] / sE;
	  else for (int k = 0; k < 2; k++) lE[k] = sm > 0 ? in.mtr[k] / sm : 0.5; }
	const double Qb = in.Qph[0] + in.Qph[1];
	for (int k = 0; k < 2; k++) A0[k] = in.etr[k] - in.mtr[k] * M.PH[k].q;
	auto PhiD = [&](double p, double& df) {
		double lam[2], ex[2] = {0.0, 0.0};
		const double delta = D0 + p * in.s;
		if (in.nodelta) lam[0] = lam[1] = 0.0;
		else if (!(delta > 0.0) || !(Qb > 0.0)) { lam[0] = lE[0]; lam[1] = lE[1]; }
		else if (delta < Qb) { lam[0] = in.Qph[0] / Qb; lam[1] = in.Qph[1] / Qb; }
		else for (int k = 0; k < 2; k++) { lam[k] = lE[k]; ex[k] = in.Qph[k] - Qb * lE[k]; }
		double f = 0.0; df = 0.0;
		for (int k = 0; k < 2; k++) {
			const Phase& P = M.PH[k];
			const double A = A0[k] + lam[k] * D0 + ex[k], B = in.atr[k] + lam[k] * in.s, N = A - (B - in.mtr[k] * P.b) * P.pi;
			f += ((P.g - 1.0) * (A + B * p) + (p + P.g * P.pi) * in.mtr[k] * P.b) / (P.g * (p + P.pi));
			df += -(P.g - 1.0) * N / (P.g * (p + P.pi) * (p + P.pi));
		}
		return f - 1.0; };
	double p = plo + smax(in.pguess - plo, 1e-3 * (std::fabs(plo) + 1.0));
	double lo = plo, hi = INF;
	bool conv = false;
	for (int it = 0; it < 400; it++) {
		double df; const double f = PhiD(p, df);
		if (f == 0.0) { conv = true; break; }
		if (f > 0) lo = p; else hi = p;
		double pn = df < 0 ? p - f / df : NAN;
		if (std::fabs(pn - p) <= 2e-16 * (std::fabs(p) + std::fabs(plo))) { p = pn; conv = true; break; }
		if (!(pn > lo && pn < hi)) {
			if (!std::isfinite(hi)) pn = plo + 4.0 * (p - plo);
			else if (lo == plo) pn = plo + 0.25 * (hi - plo);
			else pn = (hi - plo > 1e3 * (lo - plo + 1e-300) && lo > plo) ? plo + std::sqrt((lo - plo) * (hi - plo)) : 0.5 * (lo + hi);
		}
		if (std::fabs(pn - p) <= 2e-16 * (std::fabs(p) + std::fabs(plo)) || (hi < INF && hi - lo <= 4e-16 * (std::fabs(hi) + std::fabs(lo)))) { p = pn; conv = true; break; }
		p = pn;
	}
	if (!conv) return false;
	return closure_out(M, in, p, o);
}
// ---- scalar root finding (PHY:183-230): Brent's method on a sign-changing bracket, and the bracket search above a
// lower bound on y = ln(p - plo)
template <class F> static bool brent(F f, double a, double b, double fa, double fb, double tol, double& root) {   // PHY:183-204
	if (!(fa * fb <= 0)) return false;
	if (fa == 0) { root = a; return true; }
	if (fb == 0) { root = b; return true; }
	double c = a, fc = fa, d = b - a, e = d;
	for (int it = 0; it < 200; it++) {
		if (fb * fc > 0) { c = a; fc = fa; d = b - a; e = d; }
		if (std::fabs(fc) < std::fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
		const double tol1 = 2.0 * 1e-16 * std::fabs(b) + 0.5 * tol, xm = 0.5 * (c - b);
		if (std::fabs(xm) <= tol1 || fb == 0) { root = b; return true; }
		if (std::fabs(e) >= tol1 && std::fabs(fa) > std::fabs(fb)) {
			const double sv = fb / fa;
			double pv, qv;
			if (a == c) { pv = 2 * xm * sv; qv = 1 - sv; }
			else { const double qq = fa / fc, r = fb / fc; pv = sv * (2 * xm * qq * (qq - r) - (b - a) * (r - 1)); qv = (qq - 1) * (r - 1) * (sv - 1); }
			if (pv > 0) qv = -qv;
			pv = std::fabs(pv);
			if (2 * pv < smin(3 * xm * qv - std::fabs(tol1 * qv), std::fabs(e * qv))) { e = d; d = pv / qv; } else { d = xm; e = d; }
		} else { d = xm; e = d; }
		a = b; fa = fb;
		b += (std::fabs(d) > tol1) ? d : (xm > 0 ? tol1 : -tol1);
		fb = f(b); if (!std::isfinite(fb)) return false;
	}
	root = b; return true;
}
template <class F> static bool root_above(const Model& M, F f, double plo, double guess, double& root) {   // PHY:205-225
	auto Fy = [&](double y) { return f(plo + M.exp_(y)); };
	const double y0 = M.log_(smax(guess - plo, 1e-6 * (std::fabs(plo) + 1.0) + 1e-12));
	const double f0 = Fy(y0);
	if (f0 == 0.0) { root = plo + M.exp_(y0); return true; }
	double ya = y0, yb = y0, fa = f0, fb = f0, st = 0.1;
	if (!std::isfinite(f0)) {
		for (int k = 0; k < 200 && !std::isfinite(fb); k++) { yb += 1.0; fb = Fy(yb); }
		if (!std::isfinite(fb)) return false;
		ya = yb; fa = fb;
	}
	if (fa > 0) { for (int k = 0; k < 400 && !(fb < 0); k++) { ya = yb; fa = fb; yb += st; st *= 1.6; fb = Fy(yb); if (!std::isfinite(fb)) return false; } }
	else {
		for (int k = 0; k < 400 && !(fa > 0); k++) {
			yb = ya; fb = fa; ya -= st; st *= 1.6; fa = Fy(ya);
			if (!std::isfinite(fa)) {
				ya = yb; fa = fb;
				for (int q = 0; q < 60; q++) {
					const double ym = ya - st / std::ldexp(1.0, q + 1), fm = Fy(ym);   // 2^(q+1), as the reference's pow(2, q + 1)
					if (std::isfinite(fm)) { if (fm > 0) { ya = ym; fa = fm; break; } }
				}
				break;
			}
		}
	}
	if (fa == 0.0) { root = plo + M.exp_(ya); return true; }
	if (fb == 0.0) { root = plo + M.exp_(yb); return true; }
	if (!(fa > 0 && fb < 0)) return false;
	double y; if (!brent(Fy, ya, yb, fa, fb, 1e-15, y)) return false;
	root = plo + M.exp_(y); return true;
}
// pT closure (B6.3, PHY:660-672): mechanical and thermal equilibrium at the phases' masses; T(p) from the energy
static inline double v_pT(const Model& M, int k, double p, double T) { const Phase& P = M.PH[k]; return (P.g - 1.0) * P.cv * T / (p + P.pi) + P.b; }   // PHY:78
static inline double e_pT(const Model& M, int k, double p, double T) { const Phase& P = M.PH[k]; return P.cv * T * (p + P.g * P.pi) / (p + P.pi) + P.q; }   // PHY:79
static bool closure_pT(const Model& M, const ClIn& in, ClOut& o) {
	const double m[2] = {in.mtr[0], in.mtr[1]};
	const double Eth = in.rhoeC - m[0] * M.PH[0].q - m[1] * M.PH[1].q;
	if (!(Eth > 0)) return false;
	auto Tof = [&](double p) { double d = 0; for (int k = 0; k < 2; k++) d += m[k] * M.PH[k].cv * (p + M.PH[k].g * M.PH[k].pi) / (p + M.PH[k].pi); return Eth / d; };
	auto Phi = [&](double p) { const double T = Tof(p); return m[0] * v_pT(M, 0, p, T) + m[1] * v_pT(M, 1, p, T) - 1.0; };
	double p; if (!root_above(M, Phi, plo_mix(M), in.pguess, p)) return false;
	const double T = Tof(p);
	for (int k = 0; k < 2; k++) { o.m[k] = m[k]; o.a[k] = m[k] * v_pT(M, k, p, T); o.e[k] = m[k] * e_pT(M, k, p, T); o.p[k] = p; o.T[k] = T; }
	o.pbar = p; return T > 0;
}
// ---- the saturation line and the pTg closure (relaxation levels 2-3)
static inline double s_pT(const Model& M, int k, double p, double T) { const Phase& P = M.PH[k];            // PHY:80
	return P.cv * (P.g * M.log_(T) - (P.g - 1.0) * M.log_(p + P.pi)) + P.qp; }
static inline double g_pT(const Model& M, int k, double p, double T) { const Phase& P = M.PH[k];            // PHY:81
	return P.g * P.cv * T + P.b * p + P.q - T * s_pT(M, k, p, T); }
static double sat_f(const Model& M, double p, double T) { return g_pT(M, 0, p, T) - g_pT(M, 1, p, T); }   // PHY:113
// Newton on T for g_1 = g_2 at fixed p, the step kept within [T/2, 2T] (PHY:114-124)
static double sat_newton(const Model& M, double p, double Tg, bool& ok) {
	double T = Tg; ok = false;
	for (int it = 0; it < 60; it++) {
		const double f = sat_f(M, p, T), d = s_pT(M, 1, p, T) - s_pT(M, 0, p, T);
		if (!(d > 0) || !std::isfinite(f)) return T;
		double Tn = T - f / d; if (Tn < 0.5 * T) Tn = 0.5 * T; if (Tn > 2.0 * T) Tn = 2.0 * T;
		if (std::fabs(Tn - T) < 1e-14 * T) { ok = true; return Tn; }
		T = Tn;
	}
	ok = std::fabs(sat_f(M, p, T)) < 1e-8 * std::fabs(g_pT(M, 0, p, T)) + 1e-8; return T;
}
// the table (PHY:125-152): 1600 points in ln p over [ln 0.1, ln 5e9]; the first root by a scan of T in [20, 5000] K
// and bisection, each next one by Newton from the previous; it ends where no root is found or where the liquid volume
// reaches the vapour's (the critical point, refined by bisection in p)
static void build_sat(const Model& M, SatTable& S) {
	S = SatTable();
	const int NP = 1600; const double l0 = M.log_(1e-1), l1 = M.log_(5e9);
	double Tprev = -1.0;
	for (int n = 0; n < NP; n++) {
		const double lp = l0 + (l1 - l0) * n / (NP - 1), p = M.exp_(lp);
		double Troot = -1.0;
		if (Tprev < 0) {
			double Ta = 20.0, fa = sat_f(M, p, Ta);
			for (int m = 1; m <= 4000; m++) {
				const double Tb = 20.0 * M.pow_(250.0, m / 4000.0), fb = sat_f(M, p, Tb);
				if (fa < 0 && fb >= 0) {
					double a = Ta, b = Tb;
					for (int it = 0; it < 200; it++) { const double c = 0.5 * (a + b); if (sat_f(M, p, c) < 0) a = c; else b = c; }
					Troot = 0.5 * (a + b); break;
				}
				Ta = Tb; fa = fb;
			}
		} else { bool ok; const double T = sat_newton(M, p, Tprev, ok); if (ok) Troot = T; }
		if (Troot < 0) { if (n > 0) S.pcrit = M.exp_(S.lp.back()); break; }
		if (!(v_pT(M, 1, p, Troot) > v_pT(M, 0, p, Troot))) { S.pcrit = p; S.vcrit = v_pT(M, 0, p, Troot); break; }
		S.lp.push_back(lp); S.T.push_back(Troot); Tprev = Troot;
	}
	if (!S.lp.empty()) S.pmin = M.exp_(S.lp.front());
	if (S.vcrit == 0.0 && !S.T.empty()) S.vcrit = v_pT(M, 0, M.exp_(S.lp.back()), S.T.back());
	if (std::isfinite(S.pcrit) && !S.lp.empty()) {
		double a = M.exp_(S.lp.back()), b = S.pcrit, Ta = S.T.back();
		for (int it = 0; it < 100; it++) {
			const double c = std::sqrt(a * b); bool ok; const double T = sat_newton(M, c, Ta, ok);
			if (ok && v_pT(M, 1, c, T) > v_pT(M, 0, c, T)) { a = c; Ta = T; } else b = c;
		}
		S.pcrit = a; S.vcrit = v_pT(M, 0, a, Ta);
	}
}
// T_sat(p) (PHY:154-162): Newton from the table's linear interpolation in ln p; NaN outside (0, pcrit) or on failure
static double Tsat(const Model& M, double p) {
	const SatTable& S = *M.sat;
	if (!(p > 0) || !(p < S.pcrit) || S.lp.empty()) return NAN;
	const double lp = M.log_(p); double Tg;
	if (lp <= S.lp.front()) Tg = S.T.front();
	else if (lp >= S.lp.back()) Tg = S.T.back();
	else {
		const double h = S.lp[1] - S.lp[0];
		const int n = std::min((int)((lp - S.lp[0]) / h), (int)S.lp.size() - 2);
		const double w = (lp - S.lp[n]) / h; Tg = (1 - w) * S.T[n] + w * S.T[n + 1];
	}
	bool ok; const double T = sat_newton(M, p, Tg, ok); return ok ? T : NAN;
}
// the saturation pressure psat(T), the inverse of T_sat on the table's range (PHY:162-171): bisection in ln p between
// the table's lowest pressure and the critical one; 0 below the table's T, inf above the critical T
static double psat(const Model& M, double T) {
	if (!M.sat || M.sat->lp.empty()) return NAN;
	double a = M.sat->pmin, b = M.sat->pcrit * (1 - 1e-12);
	const double Ta = Tsat(M, a), Tb = Tsat(M, b);
	if (!(T > Ta)) return 0.0;
	if (!(T < Tb)) return INF;
	for (int it = 0; it < 80; it++) { const double c = std::sqrt(a * b); if (Tsat(M, c) < T) a = c; else b = c; if (b / a - 1 < 1e-14) break; }
	return std::sqrt(a * b);
}
// Pelanti 2022 eq. (84) (PHY:172-177): nu = c_r alpha_v^0.6 ((psat(T) - p) / (p_c - psat(T)))^1.76 in superheated liquid
// (p < psat(T)), else 0
static double nu_law(const Model& M, double p, double T, double av) {
	const double ps = psat(M, T);
	if (!std::isfinite(ps) || !(ps > p) || !(ps < M.nu_pc) || !(av > 0)) return 0.0;
	return M.nu_cr * M.pow_(av, 0.6) * M.pow_((ps - p) / (M.nu_pc - ps), 1.76);
}
static const double YMIN = 1e-10;                                                                    // PHY:242
// T at p and vapour mass fraction Y from the mixture's specific energy (PHY:701-704)
static inline double T_from_energy(const Model& M, double p, double Y, double eC) {
	const Phase &A = M.PH[0], &B = M.PH[1];
	const double d = (1 - Y) * A.cv * (p + A.g * A.pi) / (p + A.pi) + Y * B.cv * (p + B.g * B.pi) / (p + B.pi);
	return (eC - (1 - Y) * A.q - Y * B.q) / d;
}
// the pTg state at p (PHY:757-767): T = T_sat(p) and Y from the energy, clamped to [YMIN, 1 - YMIN] (then T from the
// energy); at p >= pcrit a single phase by v against vcrit
static bool pTg_state(const Model& M, double p, double rho, double eC, double& Y, double& T) {
	const double Ts = Tsat(M, p);
	if (std::isfinite(Ts)) {
		const double el = e_pT(M, 0, p, Ts), ev = e_pT(M, 1, p, Ts);
		Y = (eC - el) / (ev - el);
		if (Y > YMIN && Y < 1 - YMIN) { T = Ts; return true; }
		Y = smin(smax(Y, YMIN), 1 - YMIN);
	} else {
		if (!(p >= M.sat->pcrit)) return false;
		Y = (1.0 / rho < M.sat->vcrit) ? YMIN : 1 - YMIN;
	}
	T = T_from_energy(M, p, Y, eC); return T > 0;
}
// pTg closure (B6.6, PHY:768-778): mechanical, thermal and chemical equilibrium; the root of rho sum_k Y_k v_k = 1
static bool closure_pTg(const Model& M, const ClIn& in, ClOut& o) {
	const double rho = in.mtr[0] + in.mtr[1], eC = in.rhoeC / rho;
	auto Phi = [&](double p) {
		double Y, T; if (!pTg_state(M, p, rho, eC, Y, T)) return (double)NAN;
		return rho * ((1 - Y) * v_pT(M, 0, p, T) + Y * v_pT(M, 1, p, T)) - 1.0; };
	double p; if (!root_above(M, Phi, plo_mix(M), in.pguess, p)) return false;
	double Y, T; if (!pTg_state(M, p, rho, eC, Y, T)) return false;
	o.m[0] = rho * (1 - Y); o.m[1] = rho * Y;
	for (int k = 0; k < 2; k++) { o.a[k] = o.m[k] * v_pT(M, k, p, T); o.e[k] = o.m[k] * e_pT(M, k, p, T); o.p[k] = p; o.T[k] = T; }
	o.pbar = p; return true;
}
// ---- the p closure with a Mie-Grueneisen phase (alpha_mg PHY:498-528, closure_p_mg PHY:530-573; PHO:552-628 differ
// only in the start and the two early exits marked)
// alpha_k(p) of an MG phase is the root of f(a) = m er(m / a) + a (p - pr) / G + p a - A - B p, increasing where
// c^2 > 0 (f' = rho c^2 / G): the guess is first moved up past the c^2 <= 0 side, then the root is bracketed by
// doubling / halving (a step across c^2 = 0 bisects geometrically for a point of the admissible branch with f <= 0)
// and polished by safeguarded Newton. Returns alpha and d alpha / d p.
static bool alpha_mg(const Model& M, int k, double m, double A, double B, double p, double guess, double& al, double& dal) {
	const Phase& P = M.PH[k];
	const bool pho = M.closure_path != 0;
	auto F = [&](double a, double& df) {
		double pr, dpr, er; const double r = m / a; mg_ref(M, k, r, pr, dpr, er);
		df = ((P.G + 1.0) * (p - pr) + r * dpr) / P.G;
		return m * er + a * (p - pr) / P.G + p * a - A - B * p; };
	double lo = guess, hi = guess, d; double flo = F(lo, d), fhi = flo;
	if (!std::isfinite(flo)) return false;
	if (!(d > 0)) {
		for (int it = 0; it < 200 && !(d > 0); it++) { guess *= 2.0; flo = F(guess, d); if (!std::isfinite(flo)) return false; }
		if (!(d > 0)) return false;
		lo = hi = guess; fhi = flo;
	}
	if (flo < 0) {
		for (int it = 0; it < 200 && fhi < 0; it++) { lo = hi; flo = fhi; hi *= 2.0; fhi = F(hi, d); if (!std::isfinite(fhi)) return false; }
		if (fhi < 0) return false;
	} else {
		for (int it = 0; it < 200 && flo > 0; it++) {
			hi = lo; fhi = flo; lo *= 0.5; flo = F(lo, d); if (!std::isfinite(flo)) return false;
			if (d <= 0) {
				bool found = false;
				for (int b = 0; b < 200; b++) {
					const double mid = std::sqrt(lo * hi); double dm; const double fm = F(mid, dm);
					if (!std::isfinite(fm)) return false;
					if (!(dm > 0)) lo = mid;
					else if (fm > 0) { hi = mid; fhi = fm; }
					else { lo = mid; flo = fm; found = true; break; }
					if (hi - lo <= 1e-15 * hi) break;
				}
				if (!found) return false;
				break;
			}
		}
		if (flo > 0) return false;
	}
	const bool gin = pho ? (guess >= lo && guess <= hi) : (guess > lo && guess < hi);   // PHO:571 starts from an endpoint guess too
	double a = (flo == 0) ? lo : (fhi == 0 ? hi : (gin ? guess : std::sqrt(lo * hi)));
	for (int it = 0; it < 200; it++) {
		double df; const double fa = F(a, df);
		if (fa == 0) break;
		if (fa < 0) lo = a; else hi = a;
		double an = (df > 0) ? a - fa / df : -1.0;
		if (pho && an > 0 && std::fabs(an - a) <= 4e-16 * a) { a = an; break; }   // PHO:577: a converged Newton step first
		if (!(an > lo && an < hi)) an = std::sqrt(lo * hi);
		if (std::fabs(an - a) <= 4e-16 * a || hi - lo <= 4e-16 * hi) { a = an; break; }
		a = an;
	}
	double df; F(a, df); if (!(df > 0)) return false;
	al = a; dal = -(a / P.G + a - B) / df; return true;
}
// Phi(p) = sum_k alpha_k(p) - 1 on the guess's defect segment, each MG phase's alpha from alpha_mg (warm-started from the
// last accepted alphas), the NASG phases' in closed form; bracketed from the guess by doubling steps, then bisection-
// safeguarded Newton
static bool closure_p_mg(const Model& M, const ClIn& in, ClOut& o) {
	PCoef c; if (!p_coef(M, in, c, in.pguess)) return false;
	const bool pho = M.closure_path != 0;
	double ag[2] = {in.atr[0] / (in.atr[0] + in.atr[1]), in.atr[1] / (in.atr[0] + in.atr[1])};
	auto Phi = [&](double p, double& dPhi, bool upd) -> double {
		double sum = -1.0; dPhi = 0.0; double an[2];
		for (int k = 0; k < 2; k++) {
			double a, da;
			if (M.is_mg(k)) { if (!alpha_mg(M, k, in.mtr[k], c.A[k], c.B[k], p, ag[k], a, da)) return NAN; }
			else { if (!(p + M.PH[k].pi > 0) || !(c.N[k] > 0)) return NAN; a = alpha_p(M, c, in, k, p); da = dalpha_p(M, c, k, p); }
			an[k] = a; sum += a; dPhi += da;
		}
		if (upd) { ag[0] = an[0]; ag[1] = an[1]; }
		return sum; };
	double p = in.pguess, d;
	const double f = Phi(p, d, true);
	double scale = 1e-6 * (std::fabs(p) + 1e5);
	double lo, hi, flo, fhi;
	if (!std::isfinite(f) || f > 0) {
		lo = p; flo = f; hi = p;
		for (int it = 0; it < 300; it++) {
			hi = p + scale; fhi = Phi(hi, d, false);
			if (std::isfinite(fhi) && fhi <= 0) break;
			if (std::isfinite(fhi)) { lo = hi; flo = fhi; }
			scale *= 2.0;
		}
		if (!(std::isfinite(fhi) && fhi <= 0)) return false;
	} else {
		hi = p; fhi = f; lo = p;
		for (int it = 0; it < 300; it++) {
			lo = p - scale; flo = Phi(lo, d, false);
			if (std::isfinite(flo) && flo >= 0) break;
			if (!std::isfinite(flo)) { scale *= 0.5; if (scale < 1e-300) return false; continue; }
			hi = lo; fhi = flo; scale *= 2.0;
		}
		if (!(std::isfinite(flo) && flo >= 0)) return false;
	}
	(void)flo; (void)fhi;
	p = (std::isfinite(f) && in.pguess >= lo && in.pguess <= hi) ? in.pguess : 0.5 * (lo + hi);
	for (int it = 0; it < 300; it++) {
		double df; const double fp = Phi(p, df, true);
		if (fp == 0.0) break;
		if (!std::isfinite(fp) || fp > 0) lo = p; else hi = p;
		double pn = (std::isfinite(fp) && df < 0) ? p - fp / df : 0.5 * (lo + hi);
		if (pho && std::isfinite(fp) && df < 0 && std::fabs(pn - p) <= 2e-16 * (std::fabs(p) + std::fabs(pn))) { p = pn; break; }   // PHO:617
		if (!(pn > lo && pn < hi)) pn = 0.5 * (lo + hi);
		if (std::fabs(pn - p) <= 2e-16 * (std::fabs(p) + std::fabs(pn)) || hi - lo <= 4e-16 * (std::fabs(hi) + std::fabs(lo))) { p = pn; break; }
		p = pn;
	}
	double df; if (!std::isfinite(Phi(p, df, true))) return false;
	for (int k = 0; k < 2; k++) {
		o.a[k] = ag[k]; o.m[k] = in.mtr[k];
		if (!p_admissible(M, k, o.a[k], o.m[k], p)) return false;
		o.e[k] = eps_amp(M, k, o.a[k], o.m[k], p); o.p[k] = p; o.T[k] = T_amp(M, k, o.a[k], o.m[k], p);
	}
	o.pbar = p; return std::isfinite(p);
}
static bool closure_p(const Model& M, const ClIn& in, ClOut& o) {
	if (M.mg_any()) return closure_p_mg(M, in, o);                                                  // PHY:463, PHO:473
	return M.closure_path == 0 ? closure_p_phy(M, in, o) : closure_p_pho(M, in, o); }
// ---- finite mu (B6.2, PHY:575-656), backward-Euler form (RELAX_FORM 0): for a volume transfer R the phase pressures
// (p1, p2) solve a 2x2 linear system (each phase's energy with its own pressure work, the defect share on the guess's
// segment, the impedance weights w_k = Z_other / (Z0 + Z1) splitting the transfer's work); R is the root of
// F(R) = R - tau mu (p1 - p2)(R), which the relaxation keeps between 0 and tau mu dp(0): bracketed from F(0) by steps of
// the first estimate (an inadmissible trial halves back towards the last good point), then Brent's method.
static inline int mech_type(const Model& M) { return std::isinf(M.mu_rate) ? CL_P : CL_MU; }      // PHY:815
static bool closure_mu(const Model& M, const ClIn& in, ClOut& o) {
	const double sa = in.atr[0] + in.atr[1];
	const double om[2] = {in.atr[0] / sa, in.atr[1] / sa}, ak[2] = {in.atr[0] / sa, in.atr[1] / sa};
	const double D0 = in.rhoeC - in.etr[0] - in.etr[1];
	double lam[2], ex[2]; defect_split(M, in, D0 + in.pguess * in.s, lam, ex);
	const double w1 = in.Z[0] + in.Z[1] > 0 ? in.Z[1] / (in.Z[0] + in.Z[1]) : 0.5, w2 = 1.0 - w1;
	const double tm = std::isinf(M.mu_rate) ? INF : in.tau * M.mu_rate;
	const Phase &A = M.PH[0], &B = M.PH[1];
	double det = 0.0;   // the last solve's determinant
	auto solve = [&](double R, double& p1, double& p2) -> bool {
		const double a1 = ak[0] + R, a2 = ak[1] - R;
		const double C1 = dedp_am(M, 0, a1, in.mtr[0]) + om[0] * in.s, C2 = dedp_am(M, 1, a2, in.mtr[1]) + om[1] * in.s;
		const double M11 = C1 + R * w1 - lam[0] * in.s * om[0], M12 = R * w2 - lam[0] * in.s * om[1];
		const double M21 = -R * w1 - lam[1] * in.s * om[0], M22 = C2 - R * w2 - lam[1] * in.s * om[1];
		const double r1 = in.etr[0] - in.mtr[0] * A.q - A.g * A.pi * dedp_am(M, 0, a1, in.mtr[0]) + lam[0] * D0 + ex[0];
		const double r2 = in.etr[1] - in.mtr[1] * B.q - B.g * B.pi * dedp_am(M, 1, a2, in.mtr[1]) + lam[1] * D0 + ex[1];
		det = M11 * M22 - M12 * M21; if (!(std::fabs(det) > 0)) return false;
		p1 = (r1 * M22 - M12 * r2) / det; p2 = (M11 * r2 - M21 * r1) / det; return std::isfinite(p1) && std::isfinite(p2);
	};
	double R = 0.0, p1, p2;
	// the exponential form first finds the equilibrium transfer R_eq, dp(R_eq) = 0 (PHY:601-643), from a start r0 inside
	// the covolume bounds, its bracket not crossing the 2x2 system's singularity (det keeps its sign)
	const bool eq = M.relax_form == 1 || std::isinf(tm);
	if (tm > 0) {
		double det0 = 0.0, r0 = 0.0;
		if (eq) {
			const double Rlo = in.mtr[0] * A.b - ak[0], Rhi = ak[1] - in.mtr[1] * B.b;
			if (!(Rlo < Rhi)) return false;
			if (!(Rlo < 0.0 && 0.0 < Rhi)) r0 = Rlo >= 0.0 ? Rlo + 1e-9 * (Rhi - Rlo) : Rhi - 1e-9 * (Rhi - Rlo);
		}
		auto Fa = [&](double r, double& f) -> bool {
			double q1, q2; if (!solve(r, q1, q2)) return false;
			const double a1 = ak[0] + r, a2 = ak[1] - r;
			if (!(a1 > in.mtr[0] * A.b && a2 > in.mtr[1] * B.b)) return false;
			if (eq) { if (r != r0 && !(det * det0 > 0)) return false; }
			else if (!(q1 + A.pi > 0 && q2 + B.pi > 0)) return false;
			f = eq ? q2 - q1 : r - tm * (q1 - q2); return std::isfinite(f); };
		double f0; if (!Fa(r0, f0)) return false;
		det0 = det;
		if (eq && std::fabs(f0) <= 1e-13 * (std::fabs(in.pguess) + A.g * std::fabs(A.pi) + B.g * std::fabs(B.pi))) f0 = 0.0;
		R = r0;
		if (f0 != 0.0) {
			double hi = -f0;
			if (eq) {   // the acoustic linearisation dp(0) / sum_k rho_k c_k^2 / a_k (half the compressed phase's volume if inadmissible)
				double q[2], k = 0.0; solve(r0, q[0], q[1]);
				for (int jj = 0; jj < 2; jj++) { const double a = ak[jj] + (jj == 0 ? r0 : -r0), r = in.mtr[jj] / a; k += r * c2_rp(M, jj, r, q[jj]) / a; }
				hi = -f0 / k;
				if (!(std::isfinite(hi) && hi * -f0 > 0)) hi = f0 < 0 ? 0.5 * (ak[1] - r0 - in.mtr[1] * B.b) : -0.5 * (ak[0] + r0 - in.mtr[0] * A.b);
				hi += r0;
			}
			double lo = r0, flo = f0, fhi = NAN; bool ok = false, exact = false;
			for (int it = 0; it < 200; it++) {
				if (Fa(hi, fhi)) {
					if (eq && fhi == 0.0) { ok = exact = true; break; }
					if (fhi * flo < 0) { ok = true; break; }
					lo = hi; flo = fhi; hi = lo + (lo - r0 == 0 ? -f0 : (lo - r0));
				} else hi = 0.5 * (lo + hi);
			}
			if (!ok) return false;
			auto F = [&](double r) { double f; return Fa(r, f) ? f : (double)NAN; };
			if (exact) R = hi;
			else if (!brent(F, smin(lo, hi), smax(lo, hi), lo < hi ? flo : fhi, lo < hi ? fhi : flo, 1e-16 * (std::fabs(lo) + std::fabs(hi)) + 1e-300, R)) return false;
		}
		if (M.relax_form == 1 && !std::isinf(tm)) {   // PHY:644-647: R = R_eq (1 - exp(-tm D)) along the secant D = dp(0) / R_eq
			const double D = R != 0.0 && r0 == 0.0 ? -f0 / R : 0.0;
			if (D > 0 && std::isfinite(D)) R = -R * M.expm1_(-tm * D);
		}
	}
	if (!solve(R, p1, p2)) return false;
	const double pp[2] = {p1, p2};
	for (int k = 0; k < 2; k++) {
		o.a[k] = ak[k] + (k == 0 ? R : -R); o.m[k] = in.mtr[k]; o.p[k] = pp[k];
		if (!(o.a[k] > o.m[k] * M.PH[k].b) || !(pp[k] + M.PH[k].pi > 0)) return false;
		o.e[k] = eps_amp(M, k, o.a[k], o.m[k], pp[k]); o.T[k] = T_amp(M, k, o.a[k], o.m[k], pp[k]);
	}
	o.pbar = o.a[0] * p1 + o.a[1] * p2; return true;
}
// ---- finite relaxation rates, backward-Euler form (RELAX_FORM 0, the reference's default)
// finite theta (B6.4, PHY:679-700): at fixed p the phases' temperatures are linear in the exchanged heat, and
// dQ = tau theta (T2 - T1) / (1 + tau theta S), S = 1/Cp1 + 1/Cp2, shifts the p closure's volume fractions; the
// coefficient segment is the one at the guess (p_coef once), the root of sum alpha_k = 1 by the bracket search
static bool closure_theta(const Model& M, const ClIn& in, ClOut& o) {
	PCoef c; if (!p_coef(M, in, c, in.pguess)) return false;
	const double th = M.theta_rate, tt = std::isinf(th) ? INF : in.tau * th;
	if (tt > 0 && !std::isinf(tt)) for (int k = 0; k < 2; k++) if (!(c.N[k] > 0)) return false;
	const Phase &A = M.PH[0], &B = M.PH[1];
	const double Cp[2] = {A.g * A.cv * in.mtr[0], B.g * B.cv * in.mtr[1]};
	const double S = 1.0 / Cp[0] + 1.0 / Cp[1];
	auto eval = [&](double p, double a[2], double& dQ) {
		double a0[2], T0[2];
		for (int k = 0; k < 2; k++) { a0[k] = alpha_p(M, c, in, k, p); T0[k] = T_amp(M, k, a0[k], in.mtr[k], p); }
		if (M.relax_form == 1) dQ = (T0[1] - T0[0]) * (-M.expm1_(-tt * S)) / S;   // PHY:685: T2 - T1 decays by exp(-tt S)
		else if (std::isinf(tt)) dQ = (T0[1] - T0[0]) / S;
		else dQ = tt * (T0[1] - T0[0]) / (1.0 + tt * (1.0 / Cp[0] + 1.0 / Cp[1]));
		a[0] = a0[0] + (A.g - 1.0) * dQ / (A.g * (p + A.pi));
		a[1] = a0[1] - (B.g - 1.0) * dQ / (B.g * (p + B.pi));
	};
	auto Phi = [&](double p) { double a[2], dQ; eval(p, a, dQ); return a[0] + a[1] - 1.0; };
	double p; if (!root_above(M, Phi, plo_mix(M), in.pguess, p)) return false;
	double a[2], dQ; eval(p, a, dQ);
	for (int k = 0; k < 2; k++) {
		o.a[k] = a[k]; o.m[k] = in.mtr[k]; if (!(o.a[k] > o.m[k] * M.PH[k].b)) return false;
		o.e[k] = eps_amp(M, k, o.a[k], o.m[k], p); o.p[k] = p; o.T[k] = T_amp(M, k, o.a[k], o.m[k], p);
	}
	o.pbar = p; return true;
}
// finite nu (B6.5, PHY:716-753): at each p the vapour fraction Y solves rho Y - m2 + tau nu (g2 - g1)(p, T(p, Y)) = 0 on
// its admissible interval (clamped at the ends), T from the mixture energy; the root of rho sum_k Y_k v_k = 1
static bool closure_nu(const Model& M, const ClIn& in, ClOut& o) {
	const Phase &A = M.PH[0], &B = M.PH[1];
	const double rho = in.mtr[0] + in.mtr[1], eC = in.rhoeC / rho, nr = in.nu >= 0 ? in.nu : M.nu_rate, tn = std::isinf(nr) ? INF : in.tau * nr;
	auto Yof = [&](double p, double& Y, double& T) -> bool {
		auto G = [&](double y) { const double TT = T_from_energy(M, p, y, eC); if (!(TT > 0)) return (double)NAN; return g_pT(M, 1, p, TT) - g_pT(M, 0, p, TT); };
		auto R = [&](double y) {
			const double TT = T_from_energy(M, p, y, eC); if (!(TT > 0)) return (double)NAN;
			if (M.relax_form == 1 || std::isinf(tn)) return g_pT(M, 1, p, TT) - g_pT(M, 0, p, TT);   // the exponential form's Y_eq first
			return rho * y - in.mtr[1] + tn * (g_pT(M, 1, p, TT) - g_pT(M, 0, p, TT)); };
		double ya = YMIN, yb = 1.0 - YMIN;
		const double dq = B.q - A.q, yT = dq != 0.0 ? (eC - A.q) / dq : INF;
		if (dq > 0) yb = smin(yb, yT * (1.0 - 1e-12)); else if (dq < 0) ya = smax(ya, yT * (1.0 + 1e-12));
		if (!(ya < yb)) return false;
		const double Y0 = in.mtr[1] / rho;
		if (M.relax_form == 1 && tn == 0.0) { Y = smin(smax(Y0, ya), yb); T = T_from_energy(M, p, Y, eC); return T > 0; }   // no transfer
		const double Ra = R(ya), Rb = R(yb);
		if (!std::isfinite(Ra) || !std::isfinite(Rb)) return false;
		if (Ra >= 0) Y = ya; else if (Rb <= 0) Y = yb; else if (!brent(R, ya, yb, Ra, Rb, 1e-16, Y)) return false;
		if (M.relax_form == 1 && !std::isinf(tn)) {   // PHY:733-745: Y = Y_eq + (Y0 - Y_eq) exp(-tn D / rho) along the secant D of G
			const double Yeq = Y, Ys = smin(smax(Y0, ya), yb);
			const double D = Ys != Yeq ? G(Ys) / (Ys - Yeq) : 0.0;
			if (D > 0 && std::isfinite(D)) Y = smin(smax(Yeq + (Y0 - Yeq) * M.exp_(-(tn * D) / rho), ya), yb);
			else if (std::fabs(Ys - Yeq) <= 1e-14) Y = Yeq;
			else {   // G not monotone (outside B6.5's assumption): the backward-Euler root
				auto Rbe = [&](double y) { const double g = G(y); return std::isfinite(g) ? rho * y - in.mtr[1] + tn * g : (double)NAN; };
				const double ra = Rbe(ya), rb = Rbe(yb); if (!std::isfinite(ra) || !std::isfinite(rb)) return false;
				if (ra >= 0) Y = ya; else if (rb <= 0) Y = yb; else if (!brent(Rbe, ya, yb, ra, rb, 1e-16, Y)) return false;
			}
		}
		T = T_from_energy(M, p, Y, eC); return T > 0;
	};
	auto Phi = [&](double p) { double Y, T; if (!Yof(p, Y, T)) return (double)NAN; return rho * ((1 - Y) * v_pT(M, 0, p, T) + Y * v_pT(M, 1, p, T)) - 1.0; };
	double p; if (!root_above(M, Phi, plo_mix(M), in.pguess, p)) return false;
	double Y, T; if (!Yof(p, Y, T)) return false;
	o.m[0] = rho * (1 - Y); o.m[1] = rho * Y;
	for (int k = 0; k < 2; k++) { o.a[k] = o.m[k] * v_pT(M, k, p, T); o.e[k] = o.m[k] * e_pT(M, k, p, T); o.p[k] = p; o.T[k] = T; }
	o.pbar = p; return true;
}
static bool closure(const Model& M, const ClIn& in, ClOut& o) {                                    // PHY:786-810
	for (int k = 0; k < 2; k++) if (!(in.atr[k] > 0) || !(in.mtr[k] > 0) || !std::isfinite(in.etr[k])) return false;
	if (!std::isfinite(in.rhoeC) || !std::isfinite(in.s)) return false;
	bool ok;
	switch (in.type) {
	case CL_MU: ok = closure_mu(M, in, o); break;
	case CL_PT: ok = closure_pT(M, in, o); break;
	case CL_THETA: ok = closure_theta(M, in, o); break;
	case CL_NU: ok = closure_nu(M, in, o); break;
	case CL_PTG: ok = closure_pTg(M, in, o); break;
	default: ok = closure_p(M, in, o); break;
	}
	return ok && std::isfinite(o.pbar);
}
static double closure_roundoff(const Model& M, const ClIn& in, double p) {                         // PC:631-638
	if (in.type != CL_P || M.mg_any()) return 0.0;
	PCoef c; if (!p_coef(M, in, c, p)) return 0.0;
	double dphi = 0.0, sens = 0.0;
	for (int k = 0; k < 2; k++) { dphi += dalpha_p(M, c, k, p); sens += std::fabs(c.lam[k]) * (M.PH[k].g - 1.0) / (M.PH[k].g * (p + M.PH[k].pi)); }
	const double en = 2.220446049250313e-16 * (std::fabs(in.rhoeC) + std::fabs(in.etr[0]) + std::fabs(in.etr[1]) + std::fabs(in.mtr[0] * M.PH[0].q) + std::fabs(in.mtr[1] * M.PH[1].q));
	return dphi < 0 ? en * sens / -dphi : 0.0;
}

// ------------------------------------------------------------------------------------------------ fixed-order sums
// Device tree order: the terms are cut into blocks of 256*items consecutive terms, items = 4, 8, 16 or 32 growing with n
// (about a thousand blocks); in a block, thread t (0..255) adds terms t, t+256, t+512, ... in turn (starting from 0.0);
// each 64-lane wavefront combines with xor-butterflies (offsets 32..1); the block total is (w0 + w1) + (w2 + w3). The
// block totals are combined the same way by one 256-thread block (thread t adds totals t, t+256, ... in turn). The HIP
// reduction kernels use exactly this order.
static double tree256(const double* v) {
	double w[4];
	for (int q = 0; q < 4; q++) {
		double a[64]; for (int l = 0; l < 64; l++) a[l] = v[q * 64 + l];
		for (int off = 32; off >= 1; off >>= 1) { double b[64]; for (int l = 0; l < 64; l++) b[l] = a[l] + a[l ^ off]; std::memcpy(a, b, sizeof a); }
		w[q] = a[0];
	}
	return (w[0] + w[1]) + (w[2] + w[3]);
}
static int dev_items(int n) { int it = 4; while (it < 32 && (long)n > (long)256 * it * 2048) it *= 2; return it; }
template <class F> static double dev_sum(int n, F term) {
	const int items = dev_items(n), B = 256 * items, nb = (n + B - 1) / B; std::vector<double> part(nb > 0 ? nb : 1, 0.0);
	#pragma omp parallel for schedule(static)
	for (int b = 0; b < nb; b++) {
		double v[256];
		for (int t = 0; t < 256; t++) { double s = 0.0; for (int it = 0; it < items; it++) { const int i = b * B + it * 256 + t; if (i < n) s += term(i); } v[t] = s; }
		part[b] = tree256(v);
	}
	double v[256];
	for (int t = 0; t < 256; t++) { double s = 0.0; for (int c = t; c < nb; c += 256) s += part[c]; v[t] = s; }
	return tree256(v);
}
template <class F> static double blk_sum(int n, int B, F term) {                                   // PC:717-722 / PCO:744-749
	const int nb = (n + B - 1) / B; std::vector<double> part(nb, 0.0);
	#pragma omp parallel for schedule(static)
	for (int c = 0; c < nb; c++) { double s = 0; const int e = std::min(n, (c + 1) * B); for (int r = c * B; r < e; r++) s += term(r); part[c] = s; }
	double s = 0; for (int c = 0; c < nb; c++) s += part[c]; return s;
}
template <class F> static double osum(const Model& M, int n, F term) {
	if (M.dot_order == 2) return dev_sum(n, term);
	return blk_sum(n, M.dot_order == 0 ? 4096 : 1024, term);
}

// ================================================================================================ solver context
struct Cfg {
	int nx, ny; double x0, x1, y0, y1;
	int bc[4]; double bcu[4];
	double ph[2][6];
	int rk; double cfl, cfl_ac; int ramp; double dtfix; int acdt;
	double newton_tol; int newton_iter; double krylov_tol; int krylov_iter;
	int limiter, order, rus_local;
	int closure_path, share_start, dot_order, portable_math;
	int nthreads;
	int jreuse;   // JREUSE (PCO:83): modified Newton, the Jacobian rebuilt only as PCO:826-828 decides
	int variant;  // IMEX variant: 0 PC (SPB6_PC_IMEX.cpp), 1 TV, 2 TVpure, 3 TVinc, 4 TVsplit, 5 TVbb, 6 BB (SPB6_<name>_IMEX.cpp)
	int relax; double eps_i;   // RELAX_LEVEL (PHY:37; 0 p, 1 p-pT) and EPS_I (PHY:38)
	int shock_fused;           // 1: the strong-shock profile integrations in fused form (prof_integrate_fused)
	double theta_rate, nu_rate;   // theta_rate, nu_rate (PHY:33); < 0: the level's default (PC:1148-1149)
	int eos_kind[2];              // Phase::kind (PHY:50-53): 0 NASG / SG, 1 Cochran-Chan, 2 JWL
	double mg[2][7];              // a Mie-Grueneisen phase's G, c1, c2, e1, e2, rr, e0
	double mu_rate;               // mu_rate (PHY:33); <= 0: instantaneous (inf, the default)
	int nu_law; double nu_cr, nu_pc;   // NU_LAW, NU_CR, NU_PC (PHY:36); nu_cr, nu_pc <= 0: 1, 22.064e6
	int relax_form;               // RELAX_FORM (PHY:41): 1 the exponential form of the finite-rate closures
};
static Phase phase_of(const double* ph, int kind, const double* mg) {
	Phase P{ph[0], ph[1], ph[2], ph[3], ph[4], ph[5]};
	P.kind = kind;
	if (kind) { P.G = mg[0]; P.c1 = mg[1]; P.c2 = mg[2]; P.e1 = mg[3]; P.e2 = mg[4]; P.rr = mg[5]; P.e0 = mg[6]; }
	return P;
}
struct Stats { long steps, rk1, rejects, floor_acc, stages, newton_it, krylov_it, linsolves; double maxres, t, dt_last; int status;
               long jac_builds; };

struct SpRow { int n; int col[5]; double val[5]; };                                                  // PC:708

struct Ctx {
	Cfg C; Model M;
	int nx, ny, NXT, NYT; double dx, dy;
	bool twoD;
	// grids (AoS, PC:112-131), arrays (PC:133-152)
	std::vector<double> U, Un, Qe, Qs, Ust, wxL, wxR, wyL, wyR, H[3];
	std::vector<double> phx, phy, a2s, ubx, uby;
	std::vector<double> eTV1, eTV2, ETV, dA;   // the TV-type variants' explicit phase energies and total energy; BB's alpha_1 step
	std::vector<int> ctype;                    // STEP_TYPE (PC:537-560): each interior cell's closure type of the step
	std::vector<double> cnu;                   // STEP_NU (PC:162): the nu law's rate of each interior cell for the step
	SatTable SAT;                              // the saturation line (relaxation levels 2-3; PC:1151 builds it before the run)
	// stage data (PC:565-582)
	double tau, rx, ry;
	std::vector<double> ust, vst, ubsx, kpx, wLx, pvx, ubsy, kpy, wLy, pvy, Kw, Qph, cz, Pold;
	std::vector<double> Z;   // the phases' acoustic impedances of the stage's finite-mu cells (PC:956-958; 1 elsewhere)
	// pressure
	std::vector<double> P, Pst; double P_REF = 0.0;
	double t = 0.0, dtprev = -1.0; bool startup = true;
	Stats st{};
	double Ae[3][3], Ai[3][3];
	int last_newton_iters = 0, last_krylov = 0;
	inline size_t id(int i, int j) const { return (size_t)i + (size_t)j * NXT; }
	inline double& u(std::vector<double>& g, int i, int j, int k) { return g[NV * id(i, j) + k]; }
	inline double u(const std::vector<double>& g, int i, int j, int k) const { return g[NV * id(i, j) + k]; }
};

static void init_ctx(Ctx& X, const Cfg& C) {
	X.C = C;
	for (int k = 0; k < 2; k++) X.M.PH[k] = phase_of(C.ph[k], C.eos_kind[k], C.mg[k]);
	X.M.closure_path = C.closure_path; X.M.share_start = C.share_start; X.M.dot_order = C.dot_order; X.M.portable = C.portable_math;
	X.M.shock_fused = C.shock_fused;
	X.M.theta_rate = C.theta_rate < 0 ? (C.relax >= 1 ? INF : 0.0) : C.theta_rate;                 // PC:1148-1149
	X.M.nu_rate = C.nu_rate < 0 ? ((C.relax == 2 || C.relax == 3) ? INF : 0.0) : C.nu_rate;
	X.M.mu_rate = C.mu_rate > 0 ? C.mu_rate : INF;
	X.M.relax_form = C.relax_form;
	X.M.nu_law = C.nu_law; X.M.nu_cr = C.nu_cr > 0 ? C.nu_cr : 1.0; X.M.nu_pc = C.nu_pc > 0 ? C.nu_pc : 22.064e6;
	if (C.nu_law) X.M.nu_rate = 0.0;   // PC:1150: the chemical cells take finite-rate closures, each with its own rate
	if (C.relax >= 2) { build_sat(X.M, X.SAT); X.M.sat = &X.SAT; }
	X.nx = C.nx; X.ny = C.ny; X.NXT = C.nx + 2 * GH; X.NYT = C.ny + 2 * GH;
	X.dx = (C.x1 - C.x0) / C.nx; X.dy = (C.y1 - C.y0) / C.ny; X.twoD = C.ny > 1;
	const size_t nc = (size_t)X.NXT * X.NYT;
	for (auto* g : {&X.U, &X.Un, &X.Qe, &X.Qs, &X.Ust, &X.wxL, &X.wxR, &X.wyL, &X.wyR, &X.H[0], &X.H[1], &X.H[2]}) g->assign(nc * NV, 0.0);
	for (auto* a : {&X.phx, &X.phy, &X.a2s, &X.ubx, &X.uby, &X.ust, &X.vst, &X.ubsx, &X.kpx, &X.pvx, &X.ubsy, &X.kpy, &X.pvy, &X.Kw, &X.cz, &X.Pold,
	                &X.eTV1, &X.eTV2, &X.ETV, &X.dA}) a->assign(nc, 0.0);
	X.wLx.assign(nc, 0.5); X.wLy.assign(nc, 0.5); X.Qph.assign(nc * 2, 0.0); X.Z.assign(nc * 2, 1.0);
	X.P.assign((size_t)C.nx * C.ny, 0.0); X.Pst = X.P;
}

// ------------------------------------------------------------------------------------------------ boundaries
static void bc_apply(Ctx& X, std::vector<double>& g) {                                              // PHY:926-957 (TRANS/WALL/PERIODIC)
	const int nx = X.nx, ny = X.ny, G = GH; const int* BCS = X.C.bc; const double* BCU = X.C.bcu;
	auto ghost = [&](int i, int j, int si, int sj, int side) {
		for (int k = 0; k < 8; k++) X.u(g, i, j, k) = X.u(g, si, sj, k);
		if (BCS[side] != BC_WALL) return;
		const int kn = side < 2 ? IMU : IMV; const double rho = X.u(g, i, j, IM1) + X.u(g, i, j, IM2);
		if (BCU[side] == 0.0) { X.u(g, i, j, kn) = -X.u(g, i, j, kn); return; }
		const double mn = 2.0 * rho * BCU[side] - X.u(g, si, sj, kn);
		X.u(g, i, j, IEN) += 0.5 * (mn * mn - X.u(g, si, sj, kn) * X.u(g, si, sj, kn)) / rho; X.u(g, i, j, kn) = mn; };
	for (int j = 0; j < ny + 2 * G; j++)
		for (int q = 0; q < G; q++) {
			const int iL = G - 1 - q, iR = nx + G + q;
			const int sL = BCS[0] == BC_PERIODIC ? nx + G - 1 - q : (BCS[0] == BC_WALL ? G + q : G);
			const int sR = BCS[1] == BC_PERIODIC ? G + q : (BCS[1] == BC_WALL ? nx + G - 1 - q : nx + G - 1);
			ghost(iL, j, sL, j, 0); ghost(iR, j, sR, j, 1);
		}
	if (ny > 1 || BCS[2] == BC_PERIODIC) {
		for (int i = 0; i < nx + 2 * G; i++)
			for (int q = 0; q < G; q++) {
				const int jL = G - 1 - q, jR = ny + G + q;
				const int sL = BCS[2] == BC_PERIODIC ? ny + G - 1 - q : (BCS[2] == BC_WALL ? G + q : G);
				const int sR = BCS[3] == BC_PERIODIC ? G + q : (BCS[3] == BC_WALL ? ny + G - 1 - q : ny + G - 1);
				ghost(i, jL, i, sL, 2); ghost(i, jR, i, sR, 3);
			}
	} else {
		for (int i = 0; i < nx + 2 * G; i++)
			for (int q = 0; q < G; q++)
				for (int k = 0; k < 8; k++) { X.u(g, i, G - 1 - q, k) = X.u(g, i, G, k); X.u(g, i, G + 1 + q, k) = X.u(g, i, G, k); }
	}
}
static void bound_cond_array(Ctx& X, std::vector<double>& a) {                                      // PC:490-511 (signs +1)
	const int G = GH, nx = X.nx, ny = X.ny; const int* BCS = X.C.bc;
	auto src = [&](int side, int q, int n, int& sL) {
		if (BCS[side] == BC_PERIODIC) sL = (side % 2 == 0) ? n + G - 1 - q : G + q;
		else if (BCS[side] == BC_WALL) sL = (side % 2 == 0) ? G + q : n + G - 1 - q;
		else sL = (side % 2 == 0) ? G : n + G - 1; };
	auto A = [&](int i, int j) -> double& { return a[X.id(i, j)]; };
	for (int j = 0; j < ny + 2 * G; j++) for (int q = 0; q < G; q++) { int sL, sR; src(0, q, nx, sL); src(1, q, nx, sR); A(G - 1 - q, j) = A(sL, j); A(nx + G + q, j) = A(sR, j); }
	for (int i = 0; i < nx + 2 * G; i++) for (int q = 0; q < G; q++) {
		int sL, sR;
		if (ny == 1 && BCS[2] != BC_PERIODIC) { sL = G; sR = G; } else { src(2, q, ny, sL); src(3, q, ny, sR); }
		A(i, G - 1 - q) = A(i, sL); A(i, ny + G + q) = A(i, sR);
	}
}

// ------------------------------------------------------------------------------------------------ explicit stage
static void tvd_reconstruction(Ctx& X, const std::vector<double>& Q) {                             // PC:349-446
	const Model& M = X.M; const int NXT = X.NXT, NYT = X.NYT;
	X.wxL = Q; X.wxR = Q; X.wyL = Q; X.wyR = Q;
	std::fill(X.phx.begin(), X.phx.end(), 0.0); std::fill(X.phy.begin(), X.phy.end(), 0.0);
	if (X.C.order < 2) return;
	std::vector<double> vsc((size_t)NXT * NYT, 0.0), esc((size_t)NXT * NYT, 0.0);
	#pragma omp parallel for schedule(static)
	for (int c = 0; c < NXT * NYT; c++) {
		const double* q = &Q[NV * (size_t)c]; Prim w;
		const double rho = q[IM1] + q[IM2], mm = M.hypot_(q[IMU], q[IMV]);
		vsc[c] = (rho > 0 ? mm / rho : 0.0) + (prim_of(M, q, w) ? w.cf : 0.0);
		esc[c] = std::fabs(q[IE1]) + std::fabs(q[IE2]) + (rho > 0 ? 0.5 * mm * mm / rho : 0.0);
	}
	const int VOL[5] = {IA1, IM1, IM2, IE1, IE2};
	const bool lim = X.C.limiter != 0;
	for (int dir = 0; dir < (X.ny > 1 ? 2 : 1); dir++) {
		std::vector<double>& wL = dir == 0 ? X.wxL : X.wyL; std::vector<double>& wR = dir == 0 ? X.wxR : X.wyR;
		std::vector<double>& phis = dir == 0 ? X.phx : X.phy;
		#pragma omp parallel for schedule(static)
		for (int c = 0; c < NXT * NYT; c++) {
			const int i = c % NXT, j = c / NXT;
			int ia = i, ja = j, ib = i, jb = j;
			if (dir == 0) { ia = i - 1; ib = i + 1; if (ia < 0 || ib >= NXT) continue; }
			else { ja = j - 1; jb = j + 1; if (ja < 0 || jb >= NYT) continue; }
			const size_t ca = X.id(ia, ja), cc = X.id(i, j), cb = X.id(ib, jb);
			auto U = [&](size_t cell, int k) { return Q[NV * cell + k]; };
			double phi = 1.0;
			for (int q = 0; q < 5; q++) {
				const int k = VOL[q];
				const double xa = U(ca, k), xc = U(cc, k), xb = U(cb, k);
				double sc;
				if (k == IA1) sc = 3.0;
				else if (k == IM1 || k == IM2) sc = std::fabs(xa) + std::fabs(xc) + std::fabs(xb);
				else sc = esc[ca] + esc[cc] + esc[cb];
				const double a = xc - xa, b = xb - xc, tiny = 1e-12 * sc + 1e-300;
				if (!lim) break;
				if (std::fabs(a) <= tiny && std::fabs(b) <= tiny) continue;
				if (a * b <= 0) { phi = 0.0; break; }
				const double cm = 0.5 * (a + b), m = std::fabs(a) < std::fabs(b) ? a : b;
				phi = smin(phi, m / cm);
			}
			if (lim && phi > 0.0) {   // face admissibility cap (SG/NASG phases)
				const size_t cs[3] = {ca, cc, cb};
				for (int ph = 0; ph < 2; ph++) {
					if (M.is_mg(ph)) continue;
					double L1[3], L2[3];
					for (int n = 0; n < 3; n++) {
						const double a = ph == 0 ? U(cs[n], IA1) : 1.0 - U(cs[n], IA1), m = U(cs[n], ph == 0 ? IM1 : IM2), e = U(cs[n], ph == 0 ? IE1 : IE2);
						L1[n] = a - m * M.PH[ph].b; L2[n] = e - m * M.PH[ph].q - L1[n] * M.PH[ph].pi;
					}
					const double* Ls[2] = {L1, L2};
					for (int l = 0; l < 2; l++) {
						const double dL = std::fabs(Ls[l][2] - Ls[l][0]);
						if (Ls[l][1] > 0.0 && dL > 0.0) phi = smin(phi, 2.0 * Ls[l][1] / dL);
					}
				}
			}
			for (int q = 0; q < 5; q++) {
				const int k = VOL[q];
				const double sig = phi * 0.5 * (U(cb, k) - U(ca, k));
				wL[NV * cc + k] = U(cc, k) - 0.5 * sig;
				wR[NV * cc + k] = U(cc, k) + 0.5 * sig;
			}
			const double ra = U(ca, IM1) + U(ca, IM2), rc = U(cc, IM1) + U(cc, IM2), rb = U(cb, IM1) + U(cb, IM2);
			const double tiny = 1e-12 * (vsc[ca] + vsc[cc] + vsc[cb]) + 1e-300;
			double vL[2], vR[2], sm = 1.0;
			for (int d = 0; d < 2; d++) {
				const int k = d == 0 ? IMU : IMV;
				const double va = U(ca, k) / ra, vc = U(cc, k) / rc, vb = U(cb, k) / rb, a = vc - va, b = vb - vc;
				double sl = 0.0;
				if (!lim) sl = 0.5 * (a + b);
				else if (!(std::fabs(a) <= tiny && std::fabs(b) <= tiny)) {
					if (a * b > 0) { sl = std::fabs(a) < std::fabs(b) ? a : b; sm = smin(sm, sl / (0.5 * (a + b))); }
					else sm = 0.0;
				}
				vL[d] = vc - 0.5 * sl; vR[d] = vc + 0.5 * sl;
			}
			double* L = &wL[NV * cc]; double* R = &wR[NV * cc];
			const double rL = L[IM1] + L[IM2], rR = R[IM1] + R[IM2];
			L[IMU] = rL * vL[0]; L[IMV] = rL * vL[1]; L[IEN] = L[IE1] + L[IE2] + 0.5 * rL * (vL[0] * vL[0] + vL[1] * vL[1]);
			R[IMU] = rR * vR[0]; R[IMV] = rR * vR[1]; R[IEN] = R[IE1] + R[IE2] + 0.5 * rR * (vR[0] * vR[0] + vR[1] * vR[1]);
			phis[cc] = smin(phi, sm);
		}
	}
}
static void rusanov(const double* QL, const double* QR, int dirk, double* flux) {                 // PC:197-209, PC:224-246 (local speed)
	const double uL = QL[dirk] / (QL[IM1] + QL[IM2]), uR = QR[dirk] / (QR[IM1] + QR[IM2]);
	double fL[NV], fR[NV];
	for (int k = 0; k < NV; k++) { fL[k] = QL[k] * uL; fR[k] = QR[k] * uR; }
	const double a = smax(std::fabs(QL[dirk] / (QL[IM1] + QL[IM2])), std::fabs(QR[dirk] / (QR[IM1] + QR[IM2])));
	for (int k = 0; k < NV; k++) flux[k] = 0.5 * (fL[k] + fR[k]) - 0.5 * a * (QR[k] - QL[k]);
}
// IMEX variants (the reference's sibling programs): TV (SPB6_TV_IMEX.cpp), TVpure, TVinc, TVsplit, TVbb, BB
enum { VAR_PC = 0, VAR_TV, VAR_TVPURE, VAR_TVINC, VAR_TVSPLIT, VAR_TVBB, VAR_BB };
// explicit kinetic-energy flux of the TV-type variants: k = E - eps_1 - eps_2 transported with u, Rusanov dissipation on
// its jump (SPB6_TVpure_IMEX.cpp:252-258)
static double kin_flux(const double* QL, const double* QR, int IM) {
	const double rL = QL[IM1] + QL[IM2], rR = QR[IM1] + QR[IM2], uL = QL[IM] / rL, uR = QR[IM] / rR;
	const double kL = QL[IEN] - QL[IE1] - QL[IE2], kR = QR[IEN] - QR[IE1] - QR[IE2];
	const double a = smax(std::fabs(uL), std::fabs(uR));
	return 0.5 * (kL * uL + kR * uR) - 0.5 * a * (kR - kL);
}
// the explicit dissipation of phase energy k beyond the face velocity's own (SPB6_TVsplit_IMEX.cpp:259-263)
static double eps_diss(const double* QL, const double* QR, int IM, int k, double ube) {
	const double uL = QL[IM] / (QL[IM1] + QL[IM2]), uR = QR[IM] / (QR[IM1] + QR[IM2]);
	const double a = smax(std::fabs(uL), std::fabs(uR));
	return -0.5 * smax(a - std::fabs(ube), 0.0) * (QR[k] - QL[k]);
}
// BB: the face states of cell c moved to its explicit new alpha_1 (SPB6_BB_IMEX.cpp:310-318): phase masses and energies
// scaled by the volume fraction ratio, velocity kept
static void bb_rescale(double* w, double d) {
	const double a1 = w[IA1], a2 = 1.0 - a1;
	if (!(a1 > 0 && a2 > 0)) return;
	const double r1 = smax(a1 + d, 0.0) / a1, r2 = smax(a2 - d, 0.0) / a2;
	const double rho = w[IM1] + w[IM2], u = w[IMU] / rho, v = w[IMV] / rho;
	w[IM1] *= r1; w[IE1] *= r1; w[IM2] *= r2; w[IE2] *= r2;
	const double rn = w[IM1] + w[IM2];
	w[IMU] = rn * u; w[IMV] = rn * v; w[IEN] = w[IE1] + w[IE2] + 0.5 * rn * (u * u + v * v);
}
static bool explicit_update(Ctx& X, double adt) {                                                    // PC:251-326 (rus=1)
	const int g = GH, nx = X.nx, ny = X.ny, NXT = X.NXT, NYT = X.NYT;
	bool rok = true;
	#pragma omp parallel for schedule(static) reduction(&&:rok)
	for (int c = 0; c < NXT * NYT; c++) { const double rho = X.Qe[NV * (size_t)c + IM1] + X.Qe[NV * (size_t)c + IM2]; if (!(rho > 0)) rok = false; }
	if (!rok) return false;
	for (int j = g; j < ny + g; j++) for (int i = g - 1; i < nx + g; i++) {
		const double uL = X.u(X.wxR, i, j, IMU) / (X.u(X.wxR, i, j, IM1) + X.u(X.wxR, i, j, IM2)), uR = X.u(X.wxL, i + 1, j, IMU) / (X.u(X.wxL, i + 1, j, IM1) + X.u(X.wxL, i + 1, j, IM2));
		X.ubx[X.id(i, j)] = 0.5 * (uL + uR);
	}
	if (ny > 1) for (int j = g - 1; j < ny + g; j++) for (int i = g; i < nx + g; i++) {
		const double vL = X.u(X.wyR, i, j, IMV) / (X.u(X.wyR, i, j, IM1) + X.u(X.wyR, i, j, IM2)), vR = X.u(X.wyL, i, j + 1, IMV) / (X.u(X.wyL, i, j + 1, IM1) + X.u(X.wyL, i, j + 1, IM2));
		X.uby[X.id(i, j)] = 0.5 * (vL + vR);
	}
	const double rxd = adt / X.dx, ryd = adt / X.dy;
	const int V = X.C.variant;
	if (V == VAR_BB) {   // SPB6_BB_IMEX.cpp:289-324: alpha_1's explicit step from the unscaled faces, then every face state rescaled
		std::fill(X.dA.begin(), X.dA.end(), 0.0);
		#pragma omp parallel for schedule(static)
		for (int q = 0; q < nx * ny; q++) {
			const int i = q % nx + g, j = q / nx + g;
			double fp[NV], fm[NV];
			rusanov(&X.wxR[NV * X.id(i, j)], &X.wxL[NV * X.id(i + 1, j)], IMU, fp);
			rusanov(&X.wxR[NV * X.id(i - 1, j)], &X.wxL[NV * X.id(i, j)], IMU, fm);
			double a = X.u(X.Qs, i, j, IA1) - rxd * (fp[IA1] - fm[IA1]) + rxd * X.u(X.Qe, i, j, IA1) * (X.ubx[X.id(i, j)] - X.ubx[X.id(i - 1, j)]);
			if (ny > 1) {
				rusanov(&X.wyR[NV * X.id(i, j)], &X.wyL[NV * X.id(i, j + 1)], IMV, fp);
				rusanov(&X.wyR[NV * X.id(i, j - 1)], &X.wyL[NV * X.id(i, j)], IMV, fm);
				a += -ryd * (fp[IA1] - fm[IA1]) + ryd * X.u(X.Qe, i, j, IA1) * (X.uby[X.id(i, j)] - X.uby[X.id(i, j - 1)]);
			}
			X.dA[X.id(i, j)] = a - X.u(X.Qe, i, j, IA1);
		}
		#pragma omp parallel for schedule(static)
		for (int c = 0; c < NXT * NYT; c++) {
			const double d = X.dA[c];
			if (d == 0.0) continue;
			bb_rescale(&X.wxL[NV * (size_t)c], d); bb_rescale(&X.wxR[NV * (size_t)c], d);
			if (ny > 1) { bb_rescale(&X.wyL[NV * (size_t)c], d); bb_rescale(&X.wyR[NV * (size_t)c], d); }
		}
	}
	#pragma omp parallel for schedule(static)
	for (int q = 0; q < nx * ny; q++) {
		const int i = q % nx + g, j = q / nx + g;
		const size_t c = X.id(i, j);
		double fp[NV], fm[NV];
		rusanov(&X.wxR[NV * X.id(i, j)], &X.wxL[NV * X.id(i + 1, j)], IMU, fp);
		rusanov(&X.wxR[NV * X.id(i - 1, j)], &X.wxL[NV * X.id(i, j)], IMU, fm);
		double* o = &X.Ust[NV * X.id(i, j)];
		for (int k = 0; k < NV; k++) o[k] = X.u(X.Qs, i, j, k) - rxd * (fp[k] - fm[k]);
		double a2 = 0.0;
		const double *Le = &X.wxR[NV * c], *Re = &X.wxL[NV * X.id(i + 1, j)], *Lw = &X.wxR[NV * X.id(i - 1, j)], *Rw = &X.wxL[NV * c];
		if (V == VAR_TV)                                                                             // TV:298
			a2 = (1.0 - X.u(X.Qs, i, j, IA1)) - rxd * ((X.ubx[c] - fp[IA1]) - (X.ubx[X.id(i - 1, j)] - fm[IA1]));
		else if (V == VAR_TVSPLIT || V == VAR_TVBB) {                                                // TVsplit:313-316
			const double d1 = eps_diss(Le, Re, IMU, IE1, X.ubx[c]) - eps_diss(Lw, Rw, IMU, IE1, X.ubx[X.id(i - 1, j)]);
			const double d2 = eps_diss(Le, Re, IMU, IE2, X.ubx[c]) - eps_diss(Lw, Rw, IMU, IE2, X.ubx[X.id(i - 1, j)]);
			X.eTV1[c] = X.u(X.Qs, i, j, IE1) - rxd * d1; X.eTV2[c] = X.u(X.Qs, i, j, IE2) - rxd * d2;
			X.ETV[c] = X.u(X.Qs, i, j, IEN) - rxd * (kin_flux(Le, Re, IMU) - kin_flux(Lw, Rw, IMU) + d1 + d2);
		} else if (V != VAR_PC) {                                                                     // TVpure:306-307
			X.eTV1[c] = X.u(X.Qs, i, j, IE1); X.eTV2[c] = X.u(X.Qs, i, j, IE2);
			X.ETV[c] = X.u(X.Qs, i, j, IEN) - rxd * (kin_flux(Le, Re, IMU) - kin_flux(Lw, Rw, IMU));
		}
		if (V != VAR_TV) o[IA1] += rxd * X.u(X.Qe, i, j, IA1) * (X.ubx[X.id(i, j)] - X.ubx[X.id(i - 1, j)]);
		if (ny > 1) {
			double gp[NV], gm[NV];
			rusanov(&X.wyR[NV * X.id(i, j)], &X.wyL[NV * X.id(i, j + 1)], IMV, gp);
			rusanov(&X.wyR[NV * X.id(i, j - 1)], &X.wyL[NV * X.id(i, j)], IMV, gm);
			for (int k = 0; k < NV; k++) o[k] -= ryd * (gp[k] - gm[k]);
			const double *Ln = &X.wyR[NV * c], *Rn = &X.wyL[NV * X.id(i, j + 1)], *Ls = &X.wyR[NV * X.id(i, j - 1)], *Rs = &X.wyL[NV * c];
			if (V == VAR_TV) a2 -= ryd * ((X.uby[c] - gp[IA1]) - (X.uby[X.id(i, j - 1)] - gm[IA1]));
			else if (V == VAR_TVSPLIT || V == VAR_TVBB) {
				const double d1 = eps_diss(Ln, Rn, IMV, IE1, X.uby[c]) - eps_diss(Ls, Rs, IMV, IE1, X.uby[X.id(i, j - 1)]);
				const double d2 = eps_diss(Ln, Rn, IMV, IE2, X.uby[c]) - eps_diss(Ls, Rs, IMV, IE2, X.uby[X.id(i, j - 1)]);
				X.eTV1[c] -= ryd * d1; X.eTV2[c] -= ryd * d2;
				X.ETV[c] -= ryd * (kin_flux(Ln, Rn, IMV) - kin_flux(Ls, Rs, IMV) + d1 + d2);
			} else if (V != VAR_PC) X.ETV[c] -= ryd * (kin_flux(Ln, Rn, IMV) - kin_flux(Ls, Rs, IMV));
			if (V != VAR_TV) o[IA1] += ryd * X.u(X.Qe, i, j, IA1) * (X.uby[X.id(i, j)] - X.uby[X.id(i, j - 1)]);
		}
		X.a2s[X.id(i, j)] = V == VAR_TV ? a2 : 1.0 - o[IA1];
	}
	bc_apply(X, X.Ust);
	bound_cond_array(X, X.a2s);
	return true;
}

// ------------------------------------------------------------------------------------------------ implicit stage
static int M_P(const Ctx& X, int i, int j) {                                                        // PC:688-697
	int I, J; const bool xper = X.C.bc[0] == BC_PERIODIC, yper = X.C.bc[2] == BC_PERIODIC;
	if (i < GH) I = xper ? X.nx + i : GH;
	else if (i > X.nx + GH - 1) I = xper ? i - X.nx : X.nx + GH - 1;
	else I = i;
	if (j < GH) J = (yper && X.ny > 1) ? X.ny + j : GH;
	else if (j > X.ny + GH - 1) J = (yper && X.ny > 1) ? j - X.ny : X.ny + GH - 1;
	else J = j;
	return (I - GH) + X.nx * (J - GH);
}
static bool stage_prep(Ctx& X) {                                                                    // PC:851-969
	const Model& M = X.M; const int NXT = X.NXT, NYT = X.NYT, nx = X.nx, ny = X.ny;
	const std::vector<double>& Xs = X.Ust;
	auto XS = [&](int i, int j, int k) { return Xs[NV * X.id(i, j) + k]; };
	bool sok = true;
	#pragma omp parallel for schedule(static) reduction(&&:sok)
	for (int c = 0; c < NXT * NYT; c++) {
		const int i = c % NXT, j = c / NXT;
		const double sa = XS(i, j, IA1) + X.a2s[c], rho = XS(i, j, IM1) + XS(i, j, IM2);
		if (!(sa > 0 && rho > 0)) { sok = false; continue; }
		X.ust[c] = XS(i, j, IMU) / rho; X.vst[c] = XS(i, j, IMV) / rho;
	}
	if (!sok) return false;
	// frozen sound speed and residual scale K + P_E (PC:875-898)
	#pragma omp parallel for schedule(static)
	for (int c = 0; c < NXT * NYT; c++) {
		const int i = c % NXT, j = c / NXT;
		const double a1 = XS(i, j, IA1), a2 = X.a2s[c], m1 = XS(i, j, IM1), m2 = XS(i, j, IM2);
		double cz = INF;
		if (a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0) {
			const double p1 = p_ame(M, 0, a1, m1, XS(i, j, IE1)), p2 = p_ame(M, 1, a2, m2, XS(i, j, IE2));
			if (p_admissible(M, 0, a1, m1, p1) && p_admissible(M, 1, a2, m2, p2)) {
				const double c21 = c2_rp(M, 0, m1 / a1, p1), c22 = c2_rp(M, 1, m2 / a2, p2);
				const double c2 = (m1 * c21 + m2 * c22) / (m1 + m2);
				cz = c2 > 0 ? std::sqrt(c2) : INF;
			}
		}
		X.cz[c] = cz;
		const double sa = a1 + a2;
		double K = smax(M.PH[0].pi, M.PH[1].pi), PE = 0.0;
		if (a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0) {
			const double p1 = p_ame(M, 0, a1, m1, XS(i, j, IE1)), p2 = p_ame(M, 1, a2, m2, XS(i, j, IE2));
			if (p_admissible(M, 0, a1, m1, p1) && p_admissible(M, 1, a2, m2, p2)) K = 1.0 / ((a1 / sa) / (m1 / a1 * c2_rp(M, 0, m1 / a1, p1)) + (a2 / sa) / (m2 / a2 * c2_rp(M, 1, m2 / a2, p2)));
			const double dd = dedp_am(M, 0, a1, m1) + dedp_am(M, 1, a2, m2);
			if (dd > 0) PE = (std::fabs(XS(i, j, IE1)) + eps_round(M, 0, a1, m1) + std::fabs(XS(i, j, IE2)) + eps_round(M, 1, a2, m2)) / dd;
		}
		X.Kw[c] = K + PE;
		X.Pold[c] = X.Pst[M_P(X, i, j)];   // PC:899-900: the stage pressure (Newton guess)
		// the finite-mu cells' frozen impedances Z_k = rho_k c_k of the stage's content per unit volume (PC:956-958)
		X.Z[2 * c] = X.Z[2 * c + 1] = 1.0;
		if (i >= GH && i < X.nx + GH && j >= GH && j < X.ny + GH && sa > 0) {
			const int ty = X.ctype.empty() ? mech_type(M) : X.ctype[(size_t)(i - GH) + (size_t)(j - GH) * X.nx];
			if (ty == CL_MU) {
				const double cc[6] = {a1 / sa, a2 / sa, m1 / sa, m2 / sa, XS(i, j, IE1) / sa, XS(i, j, IE2) / sa};
				for (int k = 0; k < 2; k++) {
					const double a = cc[k], m = cc[2 + k], pk = p_ame(M, k, a, m, cc[4 + k]), r = m / a;
					X.Z[2 * c + k] = r * std::sqrt(smax(c2_rp(M, k, r, pk), 1e-30));
				}
			}
		}
	}
	const double tau = X.tau, dx = X.dx, dy = X.dy;
	auto rhoS = [&](int i, int j) { return XS(i, j, IM1) + XS(i, j, IM2); };
	auto minmod = [](double a, double b) { return a * b <= 0 ? 0.0 : (std::fabs(a) < std::fabs(b) ? a : b); };
	const bool lim = X.C.limiter != 0; const int order = X.C.order;
	auto psi_face = [&](double xm, double xL, double xR, double xp) {                               // PC:901-905
		if (order < 2) return 1.0;
		const double D = xR - xL; if (D == 0.0) return 1.0;
		if (!lim) return smin(1.0, smax(0.0, 1.0 - 0.25 * ((xR - xm) + (xp - xL)) / D));
		return 1.0 - 0.5 * (minmod(xL - xm, D) + minmod(D, xp - xR)) / D; };
	auto UST = [&](int i, int j) { return X.ust[X.id(i, j)]; };
	auto VST = [&](int i, int j) { return X.vst[X.id(i, j)]; };
	auto dil_jump = [&](int i, int j, bool ydir) -> double {                                         // PC:912-925
		if (!ydir) {
			const double d = UST(i + 1, j) - UST(i, j);
			if (ny == 1 || j < 1 || j + 1 >= NYT) return d;
			return 0.25 * (2.0 * d + ((UST(i + 1, j + 1) - UST(i, j + 1)) + (UST(i + 1, j - 1) - UST(i, j - 1))))
				+ 0.25 * (dx / dy) * ((VST(i, j + 1) - VST(i, j - 1)) + (VST(i + 1, j + 1) - VST(i + 1, j - 1)));
		}
		const double d = VST(i, j + 1) - VST(i, j);
		if (i < 1 || i + 1 >= NXT) return d;
		return 0.25 * (2.0 * d + ((VST(i + 1, j + 1) - VST(i + 1, j)) + (VST(i - 1, j + 1) - VST(i - 1, j))))
			+ 0.25 * (dy / dx) * ((UST(i + 1, j) - UST(i - 1, j)) + (UST(i + 1, j + 1) - UST(i - 1, j + 1)));
	};
	auto set_face = [&](int iL, int jL, int iR, int jR, double hh, bool ydir, double& ubs, double& kp, double& wL, double& pv) {   // PC:926-948
		const double rL = rhoS(iL, jL), rR = rhoS(iR, jR);
		const double uL = ydir ? VST(iL, jL) : UST(iL, jL), uR = ydir ? VST(iR, jR) : UST(iR, jR);
		ubs = 0.5 * (uL + uR); kp = tau / hh / (0.5 * (rL + rR)); wL = rR / (rL + rR); pv = 0.0;
		const double cL = X.cz[X.id(iL, jL)], cR = X.cz[X.id(iR, jR)];
		if (!(std::isfinite(cL) && std::isfinite(cR))) return;
		const double phi = smax(0.0, 1.0 - tau * smin(cL, cR) / hh);
		if (phi == 0.0) return;
		const int di = iR - iL, dj = jR - jL, im = iL - di, jm = jL - dj, ip = iR + di, jp = jR + dj;
		double psiP = 1.0, psiU = 1.0;
		if (im >= 0 && jm >= 0 && ip < NXT && jp < NYT) {
			const std::vector<double>& ph = ydir ? X.phy : X.phx;
			const double phis = smin(ph[X.id(iL, jL)], ph[X.id(iR, jR)]);
			auto PO = [&](int i, int j) { return X.Pold[X.id(i, j)]; };
			psiP = 1.0 - (1.0 - psi_face(PO(im, jm), PO(iL, jL), PO(iR, jR), PO(ip, jp))) * phis;
			const double vm = ydir ? VST(im, jm) : UST(im, jm), vp = ydir ? VST(ip, jp) : UST(ip, jp);
			psiU = psi_face(vm, uL, uR, vp);
		}
		const double ZL = rL * cL, ZR = rR * cR;
		kp += phi * psiP / (ZL + ZR); pv = phi * psiU * ZL * ZR / (ZL + ZR) * dil_jump(iL, jL, ydir);
	};
	#pragma omp parallel for schedule(static)
	for (int c = 0; c < NXT * NYT; c++) {
		const int i = c % NXT, j = c / NXT;
		if (i < NXT - 1) set_face(i, j, i + 1, j, dx, false, X.ubsx[c], X.kpx[c], X.wLx[c], X.pvx[c]);
		if (ny > 1 && j < NYT - 1) set_face(i, j, i, j + 1, dy, true, X.ubsy[c], X.kpy[c], X.wLy[c], X.pvy[c]);
	}
	// shock budget on the +-1 cross stencil of X* (PC:961-969)
	#pragma omp parallel for schedule(dynamic, 8)
	for (int q = 0; q < nx * ny; q++) {
		const int i = q % nx + GH, j = q / nx + GH;
		auto cell = [&](int di, int dj) { return &Xs[NV * X.id(i + di, j + dj)]; };
		shock_budget(M, cell, X.twoD, &X.Qph[X.id(i, j) * 2]);
	}
	return true;
}

struct FaceVals { double uf, du, pf, wL, wR, utL, utR; const double* cL; const double* cR; };     // PC:590
struct FaceFlux { double M, Mt, E; };
static inline void face_weights(FaceVals& f, double ube) { f.wL = smax(f.uf, 0.0) - smax(ube, 0.0); f.wR = smin(f.uf, 0.0) - smin(ube, 0.0); }   // PC:592
static inline int wall_x(const Ctx& X, int i) { if (i == GH - 1 && X.C.bc[0] == BC_WALL) return 0; if (i == X.nx + GH - 1 && X.C.bc[1] == BC_WALL) return 1; return -1; }   // PC:594
static inline int wall_y(const Ctx& X, int j) { if (j == GH - 1 && X.C.bc[2] == BC_WALL) return 2; if (j == X.ny + GH - 1 && X.C.bc[3] == BC_WALL) return 3; return -1; }   // PC:595
// donor face states: x face (i,j) = i+1/2 takes its left state from cell i's right face value (wxR) and its right
// state from cell i+1's left face value (wxL) (PC:866-873); the donors are read in place from the face grids
static inline void face_x(const Ctx& X, int i, int j, double PL, double PR, FaceVals& f) {           // PC:597-606
	const size_t c = X.id(i, j);
	f.pf = X.wLx[c] * PL + (1.0 - X.wLx[c]) * PR - X.pvx[c];
	f.cL = &X.wxR[NV * c]; f.cR = &X.wxL[NV * X.id(i + 1, j)]; f.utL = X.vst[c]; f.utR = X.vst[X.id(i + 1, j)];
	if (const int w = wall_x(X, i); w >= 0) {
		if (X.C.bcu[w] == 0.0) { f.uf = 0.0; f.du = 0.0; f.wL = f.wR = 0.0; return; }
		f.uf = X.C.bcu[w]; f.du = f.uf - X.ubx[c]; face_weights(f, X.ubx[c]); return; }
	f.uf = X.ubsx[c] - X.kpx[c] * (PR - PL);
	f.du = f.uf - X.ubx[c];
	face_weights(f, X.ubx[c]);
}
static inline void face_y(const Ctx& X, int i, int j, double PL, double PR, FaceVals& f) {           // PC:607-616
	const size_t c = X.id(i, j);
	f.pf = X.wLy[c] * PL + (1.0 - X.wLy[c]) * PR - X.pvy[c];
	f.cL = &X.wyR[NV * c]; f.cR = &X.wyL[NV * X.id(i, j + 1)]; f.utL = X.ust[c]; f.utR = X.ust[X.id(i, j + 1)];
	if (const int w = wall_y(X, j); w >= 0) {
		if (X.C.bcu[w] == 0.0) { f.uf = 0.0; f.du = 0.0; f.wL = f.wR = 0.0; return; }
		f.uf = X.C.bcu[w]; f.du = f.uf - X.uby[c]; face_weights(f, X.uby[c]); return; }
	f.uf = X.ubsy[c] - X.kpy[c] * (PR - PL);
	f.du = f.uf - X.uby[c];
	face_weights(f, X.uby[c]);
}
// donor component k of the reference's 6-vector (alpha1, 1 - alpha1, m1, m2, e1, e2) read from an 8-variable face state
static inline double dk(const double* w, int k) {
	return k == 0 ? w[IA1] : k == 1 ? 1.0 - w[IA1] : k == 2 ? w[IM1] : k == 3 ? w[IM2] : k == 4 ? w[IE1] : w[IE2]; }
static inline double inc(const FaceVals& f, int k) { return dk(f.cL, k) * f.wL + dk(f.cR, k) * f.wR; }   // PC:617
static inline FaceFlux face_flux(const Ctx& X, const FaceVals& f) {                                 // PC:620-625
	const double ML = (f.cL[IM1] + f.cL[IM2]) * f.wL, MR = (f.cR[IM1] + f.cR[IM2]) * f.wR, Mm = ML + MR;
	FaceFlux F; F.M = Mm; F.Mt = f.utL * ML + f.utR * MR;
	F.E = inc(f, 4) + inc(f, 5) + 0.5 * f.uf * f.uf * Mm + 0.5 * (f.utL * f.utL * ML + f.utR * f.utR * MR) + (X.P_REF + f.pf) * f.uf;
	return F;
}
struct CellOut { ClOut cl; double mx, my, E, s, sig; };
static bool cell_residual(const Ctx& X, int i, int j, double Pc, double PW, double PE, double PS, double PN, double& G, CellOut* out) {   // PC:640-679
	const Model& M = X.M;
	FaceVals fe, fw, fn, fs;
	face_x(X, i, j, Pc, PE, fe); face_x(X, i - 1, j, PW, Pc, fw);
	const bool twoD = X.twoD;
	if (twoD) { face_y(X, i, j, Pc, PN, fn); face_y(X, i, j - 1, PS, Pc, fs); }
	const size_t c = X.id(i, j); const double* xs = &X.Ust[NV * c];
	const int V = X.C.variant;
	// the TV-type variants: phase energies from the explicit step without transport, enthalpy flux (sum eps_k + p_f) u_f
	// with upwind donors (TVpure:655-675); TVinc / TVsplit add the mass increments and their kinetic energy (TVinc:662-687)
	const bool tvx = V >= VAR_TVPURE, tvinc = V == VAR_TVINC || V == VAR_TVSPLIT;
	auto up = [](const FaceVals& f, int k) { return f.uf > 0.0 ? dk(f.cL, k) : dk(f.cR, k); };
	auto hflux = [&](const FaceVals& f) { return (up(f, 4) + up(f, 5) + X.P_REF + f.pf) * f.uf; };
	auto kinc = [](const FaceVals& f, const FaceFlux& F) {
		return 0.5 * f.uf * f.uf * F.M + 0.5 * (f.utL * f.utL * (f.cL[IM1] + f.cL[IM2]) * f.wL + f.utR * f.utR * (f.cR[IM1] + f.cR[IM2]) * f.wR); };
	double Xv[6] = {xs[IA1], X.a2s[c], xs[IM1], xs[IM2], tvx ? X.eTV1[c] : xs[IE1], tvx ? X.eTV2[c] : xs[IE2]};
	if (tvx)
		for (int k = 4; k < 6; k++) {
			Xv[k] -= X.rx * (up(fe, k) * fe.uf - up(fw, k) * fw.uf);
			if (twoD) Xv[k] -= X.ry * (up(fn, k) * fn.uf - up(fs, k) * fs.uf);
		}
	const int k0 = V == VAR_TV ? 0 : 2, k1 = !tvx ? 6 : tvinc ? 4 : 2;
	for (int k = k0; k < k1; k++) {
		Xv[k] -= X.rx * (inc(fe, k) - inc(fw, k));
		if (twoD) Xv[k] -= X.ry * (inc(fn, k) - inc(fs, k));
	}
	double s;
	if (V == VAR_TV) s = 1.0 - Xv[0] - Xv[1];                                                        // TV:646
	else {
		Xv[1] = 1.0 - Xv[0];
		s = X.rx * (fe.uf - fw.uf) + (twoD ? X.ry * (fn.uf - fs.uf) : 0.0);
		Xv[0] *= 1.0 - s; Xv[1] *= 1.0 - s;
	}
	const FaceFlux FEe = face_flux(X, fe), FEw = face_flux(X, fw);
	const bool mom_inc = !tvx || tvinc;   // TVpure / TVbb / BB: mass and momentum transported in the explicit step only
	double mx = mom_inc ? xs[IMU] - X.rx * (fe.pf - fw.pf) - X.rx * (fe.uf * FEe.M - fw.uf * FEw.M) : xs[IMU] - X.rx * (fe.pf - fw.pf);
	double my = mom_inc ? xs[IMV] - X.rx * (FEe.Mt - FEw.Mt) : xs[IMV];
	double E = !tvx ? xs[IEN] - X.rx * (FEe.E - FEw.E)
	         : tvinc ? X.ETV[c] - X.rx * (hflux(fe) + kinc(fe, FEe) - hflux(fw) - kinc(fw, FEw)) : X.ETV[c] - X.rx * (hflux(fe) - hflux(fw));
	if (twoD) {
		const FaceFlux FEn = face_flux(X, fn), FEs = face_flux(X, fs);
		if (mom_inc) {
			my -= X.ry * (fn.pf - fs.pf) + X.ry * (fn.uf * FEn.M - fs.uf * FEs.M);
			mx -= X.ry * (FEn.Mt - FEs.Mt);
		} else my -= X.ry * (fn.pf - fs.pf);
		if (!tvx) E -= X.ry * (FEn.E - FEs.E);
		else if (tvinc) E -= X.ry * (hflux(fn) + kinc(fn, FEn) - hflux(fs) - kinc(fs, FEs));
		else E -= X.ry * (hflux(fn) - hflux(fs));
	}
	const double rho = Xv[2] + Xv[3];
	if (!(rho > 0)) return false;
	ClIn in; in.atr[0] = Xv[0]; in.atr[1] = Xv[1]; in.mtr[0] = Xv[2]; in.mtr[1] = Xv[3]; in.etr[0] = Xv[4]; in.etr[1] = Xv[5];
	in.s = s; in.rhoeC = E - 0.5 * (mx * mx + my * my) / rho; in.pguess = X.P_REF + Pc; in.Qph[0] = X.Qph[c * 2]; in.Qph[1] = X.Qph[c * 2 + 1];
	in.type = X.ctype.empty() ? mech_type(M) : X.ctype[(size_t)(i - GH) + (size_t)(j - GH) * X.nx];   // S.type (PC:951)
	in.tau = X.tau;
	if (in.type == CL_MU) { in.Z[0] = X.Z[2 * c]; in.Z[1] = X.Z[2 * c + 1]; }
	if (!X.cnu.empty()) in.nu = X.cnu[(size_t)(i - GH) + (size_t)(j - GH) * X.nx];   // PC:671
	ClOut o; if (!closure(M, in, o)) return false;
	G = (o.pbar - X.P_REF) - Pc;
	if (out) { out->cl = o; out->mx = mx; out->my = my; out->E = E; out->s = s; out->sig = closure_roundoff(M, in, o.pbar); }
	return std::isfinite(G);
}
static inline double Pnb(const Ctx& X, const std::vector<double>& P, int i, int j, int di, int dj) { return P[M_P(X, i + GH + di, j + GH + dj)]; }   // PC:698
static bool residual_at(const Ctx& X, const std::vector<double>& P, int i, int j, double dc, int which, double dn, double& G, CellOut* out = nullptr) {   // PC:699-705
	const double Pc = P[(size_t)i + (size_t)j * X.nx] + dc;
	double PW = Pnb(X, P, i, j, -1, 0), PE = Pnb(X, P, i, j, 1, 0), PS = Pnb(X, P, i, j, 0, -1), PN = Pnb(X, P, i, j, 0, 1);
	if (which == 0) PW += dn;
	if (which == 1) PE += dn;
	if (which == 2) PS += dn;
	if (which == 3) PN += dn;
	return cell_residual(X, i + GH, j + GH, Pc, PW, PE, PS, PN, G, out);
}

// ------------------------------------------------------------------------------------------------ linear solvers
static void spmv(const std::vector<SpRow>& A, const std::vector<double>& x, std::vector<double>& y) {   // PC:711-715
	const int n = (int)A.size();
	#pragma omp parallel for schedule(static)
	for (int r = 0; r < n; r++) { double s = 0; for (int q = 0; q < A[r].n; q++) s += A[r].val[q] * x[A[r].col[q]]; y[r] = s; }
}
static bool bicgstab(Ctx& X, std::vector<double>& x, const std::vector<double>& b, const std::vector<SpRow>& A, int& iters) {   // PC:725-758
	const Model& M = X.M; const int n = (int)b.size(); const double tol = X.C.krylov_tol; const int maxit = X.C.krylov_iter;
	std::vector<double> dinv(n, 1.0);
	for (int r = 0; r < n; r++) if (A[r].n > 0 && A[r].val[0] != 0) dinv[r] = 1.0 / A[r].val[0];
	auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) { return osum(M, n, [&](int r) { return a[r] * c[r]; }); };
	const double E = osum(M, n, [&](int r) { double e = 0; for (int q = 0; q < A[r].n; q++) e += A[r].val[q]; return e; });
	const double Einv = (E != 0.0 && std::isfinite(E)) ? 1.0 / E : 0.0;
	auto prec = [&](const std::vector<double>& in, std::vector<double>& out) {
		const double cc = osum(M, n, [&](int r) { return in[r]; }) * Einv;
		#pragma omp parallel for schedule(static)
		for (int r = 0; r < n; r++) out[r] = dinv[r] * in[r] + cc; };
	x.assign(n, 0.0); prec(b, x);
	std::vector<double> r(n), r0(n), p(n, 0.0), v(n, 0.0), s(n), t(n), ph(n), sh(n), Ax(n);
	spmv(A, x, Ax); for (int q = 0; q < n; q++) r[q] = b[q] - Ax[q]; r0 = r;
	const double bn = std::sqrt(dot(b, b)) + 1e-300; double rho = 1, alpha = 1, omega = 1;
	iters = 0;
	for (int it = 0; it < maxit; it++) {
		iters = it;
		if (std::sqrt(dot(r, r)) <= tol * bn) return true;
		const double rho1 = dot(r0, r); if (rho1 == 0) return false;
		const double beta = (rho1 / rho) * (alpha / omega); rho = rho1;
		#pragma omp parallel for schedule(static)
		for (int q = 0; q < n; q++) p[q] = r[q] + beta * (p[q] - omega * v[q]);
		prec(p, ph);
		spmv(A, ph, v); alpha = rho / dot(r0, v);
		#pragma omp parallel for schedule(static)
		for (int q = 0; q < n; q++) s[q] = r[q] - alpha * v[q];
		if (std::sqrt(dot(s, s)) <= tol * bn) { for (int q = 0; q < n; q++) x[q] += alpha * ph[q]; iters = it + 1; return true; }
		prec(s, sh);
		spmv(A, sh, t); omega = dot(t, s) / dot(t, t);
		#pragma omp parallel for schedule(static)
		for (int q = 0; q < n; q++) { x[q] += alpha * ph[q] + omega * sh[q]; r[q] = s[q] - omega * t[q]; }
	}
	iters = maxit;
	return std::sqrt(dot(r, r)) <= 1e3 * tol * bn;
}
static void thomas(const std::vector<SpRow>& A, const std::vector<double>& rhs, std::vector<double>& dP) {   // PC:807-812
	const int n = (int)A.size();
	std::vector<double> lo(n), di(n), up(n), cp(n), dp(n);
	for (int r = 0; r < n; r++) { lo[r] = up[r] = di[r] = 0; for (int q = 0; q < A[r].n; q++) { const int c = A[r].col[q]; if (c == r) di[r] += A[r].val[q]; else if (c == r - 1) lo[r] = A[r].val[q]; else if (c == r + 1) up[r] = A[r].val[q]; } }
	cp[0] = up[0] / di[0]; dp[0] = rhs[0] / di[0];
	for (int r = 1; r < n; r++) { const double m = di[r] - lo[r] * cp[r - 1]; cp[r] = up[r] / m; dp[r] = (rhs[r] - lo[r] * dp[r - 1]) / m; }
	dP[n - 1] = dp[n - 1]; for (int r = n - 2; r >= 0; r--) dP[r] = dp[r] - cp[r] * dP[r + 1];
}

// ------------------------------------------------------------------------------------------------ Newton (PC:761-847, PCO:799-904)
// jreuse (PCO:796-888): the Jacobian and its scale s are rebuilt at a stage's first iteration and whenever the previous
// step was damped or contracted the residual by less than 4x (fresh_ok, PCO:797); a line search that fails with a reused
// Jacobian rebuilds it and takes the step again from the same P (PCO:833).
struct NewtonTrace { std::vector<double>* J5; };   // optional dump of the first Jacobian (tests)
static bool newton(Ctx& X, std::vector<double>& P, int& iters, double& resfinal) {
	const int nx = X.nx, ny = X.ny, n = nx * ny; const double tol = X.C.newton_tol;
	std::vector<double> G(n), Gn(n), dP(n), rhs(n), sc(n), Pn = P;
	std::vector<SpRow> A(n);
	auto KW = [&](int q) { return X.Kw[X.id(q % nx + GH, q / nx + GH)]; };
	auto evalAll = [&](const std::vector<double>& Q, std::vector<double>& R) -> bool {
		bool ok = true;
		#pragma omp parallel for schedule(static) reduction(&&:ok)
		for (int q = 0; q < n; q++) { double g; if (!residual_at(X, Q, q % nx, q / nx, 0, -1, 0, g)) ok = false; else R[q] = g; }
		return ok; };
	auto norm = [&](const std::vector<double>& R, const std::vector<double>& Q) {
		double m = 0;
		#pragma omp parallel for schedule(static) reduction(max:m)
		for (int q = 0; q < n; q++) m = smax(m, std::fabs(R[q]) / (std::fabs(X.P_REF + Q[q]) + KW(q) + 1.0));
		return m; };
	if (!evalAll(P, G)) { iters = 0; resfinal = INF; return false; }
	double res = norm(G, P);
	const bool tri = (ny == 1 && X.C.bc[0] != BC_PERIODIC);
	bool fresh = false; double ratio = 1.0, lamprev = 1.0;
	for (int it = 0; it < X.C.newton_iter; it++) {
		iters = it; resfinal = res;
		if (res < tol) return true;
		const bool build = !X.C.jreuse || it == 0 || !(lamprev == 1.0 && ratio <= 0.25);
		if (build) {
		X.st.jac_builds++;
		#pragma omp parallel for schedule(static)
		for (int q = 0; q < n; q++) {
			const int I = q % nx + GH, J = q / nx + GH; const double K = X.Kw[X.id(I, J)];
			double D = X.rx * (X.kpx[X.id(I, J)] + X.kpx[X.id(I - 1, J)]);
			if (ny > 1) D += X.ry * (X.kpy[X.id(I, J)] + X.kpy[X.id(I, J - 1)]);
			sc[q] = (std::fabs(X.P_REF + P[q]) + K + 1.0) / (1.0 + K * D);
		}
		bool jok = true;
		#pragma omp parallel for schedule(static) reduction(&&:jok)
		for (int r = 0; r < n; r++) {
			const int i = r % nx, j = r / nx; const double h = 1e-7 * sc[r];
			double g1; if (!residual_at(X, P, i, j, h, -1, 0, g1)) { jok = false; continue; }
			SpRow& row = A[r]; row.n = 0; row.col[row.n] = r; row.val[row.n++] = (g1 - G[r]) / h;
			const int dd[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
			for (int w = 0; w < 4; w++) {
				if (ny == 1 && w >= 2) break;
				const int c = M_P(X, i + GH + dd[w][0], j + GH + dd[w][1]);
				const double hn = 1e-7 * sc[c];
				double g2; if (!residual_at(X, P, i, j, 0, w, hn, g2)) { jok = false; break; }
				row.col[row.n] = c; row.val[row.n++] = (g2 - G[r]) / hn;
			}
		}
		if (!jok) return false;
		}
		fresh = build;
		for (int q = 0; q < n; q++) rhs[q] = -G[q];
		X.st.linsolves++;
		if (tri) thomas(A, rhs, dP);
		else { int kit; const bool ok = bicgstab(X, dP, rhs, A, kit); X.st.krylov_it += kit; X.last_krylov = kit; if (!ok) return false; }
		double sn = 0.0;
		for (int q = 0; q < n; q++) sn = smax(sn, std::fabs(dP[q]) / (std::fabs(P[q]) + sc[q]));
		if (sn < tol) {
			for (int q = 0; q < n; q++) P[q] += dP[q];
			X.st.newton_it++; iters = it + 1;
			if (!evalAll(P, G)) return false;
			resfinal = norm(G, P); return true;
		}
		double lam = 1.0; bool acc = false; const double res0 = res;
		for (int ls = 0; ls < 40; ls++) {
			for (int q = 0; q < n; q++) Pn[q] = P[q] + lam * dP[q];
			if (evalAll(Pn, Gn)) { const double rn = norm(Gn, Pn); if (rn < res * (1.0 - 1e-4 * lam) || rn < tol) { acc = true; res = rn; break; } }
			lam *= 0.5;
		}
		if (!acc && !fresh) { ratio = 1.0; lamprev = 0.0; continue; }   // PCO:833: rebuild and retake the step from the same P
		ratio = res / res0; lamprev = lam;
		if (!acc) {   // round-off floor acceptance (PC:834-842)
			resfinal = res; bool floor_ok = true;
			#pragma omp parallel for schedule(static) reduction(&&:floor_ok)
			for (int q = 0; q < n; q++) {
				double g; CellOut o;
				if (!residual_at(X, P, q % nx, q / nx, 0, -1, 0, g, &o)) { floor_ok = false; continue; }
				if (std::fabs(g) > tol * (std::fabs(X.P_REF + P[q]) + KW(q) + 1.0) + o.sig) floor_ok = false;
			}
			if (floor_ok) X.st.floor_acc++;
			return floor_ok;
		}
		std::swap(P, Pn); std::swap(G, Gn);
		X.st.newton_it++;
	}
	iters = X.C.newton_iter; resfinal = res; return res < tol;
}

static bool pressure_subsystem(Ctx& X, double h) {                                                  // PC:851-993
	X.tau = h; X.rx = h / X.dx; X.ry = h / X.dy;
	if (!stage_prep(X)) return false;
	int iters; double res;
	const bool ok = newton(X, X.Pst, iters, res);
	X.last_newton_iters = iters;
	if (ok) X.st.maxres = smax(X.st.maxres, res);
	if (!ok) return false;
	bool fok = true; const int nx = X.nx;
	std::vector<double> out((size_t)nx * X.ny * NV);
	#pragma omp parallel for schedule(static) reduction(&&:fok)
	for (int q = 0; q < nx * X.ny; q++) {
		double Gv; CellOut o; if (!residual_at(X, X.Pst, q % nx, q / nx, 0, -1, 0, Gv, &o)) { fok = false; continue; }
		double* w = &out[(size_t)q * NV];
		w[IA1] = o.cl.a[0]; w[IM1] = o.cl.m[0]; w[IM2] = o.cl.m[1]; w[IMU] = o.mx; w[IMV] = o.my; w[IE1] = o.cl.e[0]; w[IE2] = o.cl.e[1]; w[IEN] = o.E;
	}
	if (!fok) return false;
	for (int q = 0; q < nx * X.ny; q++) for (int k = 0; k < NV; k++) X.u(X.Ust, q % nx + GH, q / nx + GH, k) = out[(size_t)q * NV + k];
	bc_apply(X, X.Ust);
	X.st.stages++;
	return true;
}

// ------------------------------------------------------------------------------------------------ driver
static double caldt(const Ctx& X, bool acoustic) {                                                  // PC:477-488
	double ux = 0, uy = 0;
	for (int j = GH; j < X.ny + GH; j++) for (int i = GH; i < X.nx + GH; i++) {
		Prim w; if (!prim_of(X.M, &X.U[NV * X.id(i, j)], w)) continue;
		const double c = acoustic ? w.cf : 0.0;
		ux = smax(ux, std::fabs(w.u) + c); uy = smax(uy, std::fabs(w.v) + c);
	}
	const double rate = ux / X.dx + (X.ny > 1 ? uy / X.dy : 0.0);
	return rate > 0 ? 1.0 / rate : INF;
}
static void set_tab(Ctx& X, int rk) {                                                               // PC:1195-1218
	for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) X.Ae[a][b] = X.Ai[a][b] = 0;
	if (rk == 1) X.Ai[0][0] = 1.0;
	else if (rk == 2) { const double bb = 1.0 - 1.0 / std::sqrt(2.0), cc = 1.0 / (2.0 * bb); X.Ae[1][0] = cc; X.Ai[0][0] = bb; X.Ai[1][0] = 1 - bb; X.Ai[1][1] = bb; }
	else if (rk == 3) { const double r12 = 0.5, r13 = 1.0 / 3.0, r14 = 1.0 / 4.0; X.Ae[1][0] = r12; X.Ae[2][0] = r12; X.Ae[2][1] = r12; X.Ai[0][0] = r14; X.Ai[1][1] = r14; X.Ai[2][0] = r13; X.Ai[2][1] = r13; X.Ai[2][2] = r13; }
}
static void init_pressure(Ctx& X) {                                                                 // PC:1222-1227
	X.P_REF = INF;
	for (int j = 0; j < X.ny; j++) for (int i = 0; i < X.nx; i++) X.P_REF = smin(X.P_REF, pbar_of(X.M, &X.U[NV * X.id(i + GH, j + GH)]));
	if (!std::isfinite(X.P_REF)) X.P_REF = 0.0;
	for (int j = 0; j < X.ny; j++) for (int i = 0; i < X.nx; i++) X.P[(size_t)i + (size_t)j * X.nx] = pbar_of(X.M, &X.U[NV * X.id(i + GH, j + GH)]) - X.P_REF;
}
// one attempt of a step of size dt with an rk-stage table; false = inadmissible (PC:1243-1285)
// ---- relaxation levels 1-3 (PC:513-560, PHY:812-830): each cell's closure type for the step, with U^n projected onto
// that closure's equilibrium (pT, pTg). Only cells with eps_I <= alpha_1 <= 1 - eps_I relax beyond p. Level 1 (p-pT):
// thermal equilibrium there. Level 2 (p-pT-pTg): chemical where the pT equilibrium's T_1 exceeds T_sat(p), else
// thermal. Level 3 (p-pTg): chemical where either phase's T at the mechanical equilibrium (without the energy defect)
// exceeds T_sat(p), else mechanical. The rates are instantaneous (mu = theta = nu = inf, the reference's defaults).
static int closure_type_of(const Model& M, int level, double eps_i, double a1, double m1, double m2, double e1, double e2, double rhoe) {   // PHY:814-830
	const int mech = mech_type(M);
	if (level == 0) return mech;
	if (!(a1 >= eps_i && a1 <= 1.0 - eps_i)) return mech;
	ClIn in{}; in.atr[0] = a1; in.atr[1] = 1 - a1; in.mtr[0] = m1; in.mtr[1] = m2; in.etr[0] = e1; in.etr[1] = e2; in.s = 0; in.rhoeC = rhoe;
	in.type = CL_PT; in.pguess = smax(p_mix(M, a1, m1, m2, rhoe), 1.0);
	const int chem = std::isinf(M.nu_rate) ? CL_PTG : CL_NU, therm = std::isinf(M.theta_rate) ? CL_PT : CL_THETA;
	if (level == 3) {
		in.type = CL_P; in.nodelta = true; ClOut o; if (!closure_p(M, in, o)) return mech;
		const double Ts = Tsat(M, o.p[0]); if (std::isfinite(Ts) && (o.T[0] > Ts || o.T[1] > Ts)) return chem; return mech;
	}
	if (level == 2) {
		ClOut o; if (closure_pT(M, in, o)) { const double Ts = Tsat(M, o.p[0]); if (std::isfinite(Ts) && o.T[0] > Ts) return chem; }
	}
	return therm;
}
static bool project_cell(const Model& M, double* Q, int type) {                                      // PC:524-535
	if (!(type == CL_PT || type == CL_PTG)) return true;
	const double rho = Q[IM1] + Q[IM2];
	ClIn in{}; in.atr[0] = Q[IA1]; in.atr[1] = 1.0 - Q[IA1]; in.mtr[0] = Q[IM1]; in.mtr[1] = Q[IM2];
	in.etr[0] = Q[IE1]; in.etr[1] = Q[IE2]; in.s = 0.0;
	in.rhoeC = Q[IEN] - 0.5 * (Q[IMU] * Q[IMU] + Q[IMV] * Q[IMV]) / rho;
	in.type = type; in.pguess = smax(p_mix(M, in.atr[0], in.mtr[0], in.mtr[1], in.rhoeC), 1.0);
	ClOut o; if (!closure(M, in, o)) return false;
	Q[IA1] = o.a[0]; Q[IM1] = o.m[0]; Q[IM2] = o.m[1]; Q[IE1] = o.e[0]; Q[IE2] = o.e[1];
	return true;
}
static void relaxation_activation(Ctx& X, std::vector<double>& U) {                                // PC:537-560
	const int nx = X.nx, n = nx * X.ny;
	X.ctype.assign((size_t)n, mech_type(X.M));
	if (X.M.nu_law) X.cnu.assign((size_t)n, 0.0);
	#pragma omp parallel for schedule(static)
	for (int q = 0; q < n; q++) {
		double* Q = &U[NV * X.id(q % nx + GH, q / nx + GH)];
		const double rho = Q[IM1] + Q[IM2], rhoe = Q[IEN] - 0.5 * (Q[IMU] * Q[IMU] + Q[IMV] * Q[IMV]) / rho;   // cell_type_U (PC:517-521)
		int ty = closure_type_of(X.M, X.C.relax, X.C.eps_i, Q[IA1], Q[IM1], Q[IM2], Q[IE1], Q[IE2], rhoe);
		if (!project_cell(X.M, Q, ty)) ty = mech_type(X.M);   // PC:544
		X.ctype[(size_t)q] = ty;
		if (X.M.nu_law && ty == CL_NU) {   // PC:546-556: the rate from the cell's pT equilibrium (p, T, vapour fraction)
			ClIn in{}; in.atr[0] = Q[IA1]; in.atr[1] = 1.0 - Q[IA1]; in.mtr[0] = Q[IM1]; in.mtr[1] = Q[IM2];
			in.etr[0] = Q[IE1]; in.etr[1] = Q[IE2]; in.s = 0.0; in.tau = 0.0; in.type = CL_PT; in.Qph[0] = in.Qph[1] = 0.0;
			in.rhoeC = Q[IEN] - 0.5 * (Q[IMU] * Q[IMU] + Q[IMV] * Q[IMV]) / rho;
			in.pguess = smax(p_mix(X.M, in.atr[0], in.mtr[0], in.mtr[1], in.rhoeC), 1.0);
			ClOut o; if (closure(X.M, in, o)) X.cnu[(size_t)q] = nu_law(X.M, o.p[0], o.T[0], o.a[1]);
		}
	}
	bc_apply(X, U);
}
static bool attempt(Ctx& X, double dt, int rk) {
	set_tab(X, rk);
	if (X.C.relax > 0) relaxation_activation(X, X.U);   // PC:1247, before u_con_n = u_con
	X.Un = X.U; X.Pst = X.P;
	const int nx = X.nx, ny = X.ny;
	bool ok = true;
	for (int rks = 0; rks < rk && ok; rks++) {
		X.Qe = X.Un; X.Qs = X.Un;
		for (int r = 0; r < rks; r++) {
			const double ae = X.Ae[rks][r], ai = X.Ai[rks][r];
			#pragma omp parallel for schedule(static)
			for (int q = 0; q < nx * ny; q++) {
				const size_t c = X.id(q % nx + GH, q / nx + GH);
				for (int k = 0; k < NV; k++) {
					X.Qe[NV * c + k] = X.Qe[NV * c + k] + ae * dt * X.H[r][NV * c + k];
					X.Qs[NV * c + k] = X.Qs[NV * c + k] + ai * dt * X.H[r][NV * c + k];
				}
			}
		}
		bc_apply(X, X.Qe); bc_apply(X, X.Qs);
		const double adt = X.Ai[rks][rks] * dt;
		tvd_reconstruction(X, X.Qe);
		ok = explicit_update(X, adt);
		ok = ok && pressure_subsystem(X, adt);
		if (ok) {
			#pragma omp parallel for schedule(static)
			for (int q = 0; q < nx * ny; q++) {
				const size_t c = X.id(q % nx + GH, q / nx + GH);
				for (int k = 0; k < NV; k++) X.H[rks][NV * c + k] = (X.Ust[NV * c + k] - X.Qs[NV * c + k]) / adt;
			}
		}
	}
	return ok;
}
// time loop (PC:1230-1296): runs until t_end or max_steps accepted steps; returns 0 ok, 2 failure
static int run(Ctx& X, double t_end, long max_steps) {
	const double ramp_factor = 1.2;
	long done_steps = 0;
	while (X.t < t_end * (1 - 1e-12) && (max_steps < 0 || done_steps < max_steps)) {
		bool ramping = false; double dt;
		if (X.C.dtfix > 0) dt = X.C.dtfix;
		else {
			dt = X.C.cfl * caldt(X, X.C.acdt != 0);
			if (X.C.ramp) { if (X.dtprev < 0) X.dtprev = X.C.cfl_ac * caldt(X, true) / ramp_factor; if (dt > ramp_factor * X.dtprev) { dt = ramp_factor * X.dtprev; ramping = true; } }
			if (!std::isfinite(dt)) dt = X.C.cfl_ac * caldt(X, true);
		}
		dt = smin(dt, t_end - X.t);
		if (!(dt > 1e-12 * t_end)) { X.st.status = 3; return 2; }
		bool ok = false;
		for (int att = 0; att < 30 && !ok; att++) {
			const bool first_order = X.C.ramp && ramping && X.startup;
			const int rk = first_order ? 1 : X.C.rk; if (rk == 1) X.st.rk1++;
			ok = attempt(X, dt, rk);
			if (ok) { X.U = X.Ust; X.P = X.Pst; }
			else { X.st.rejects++; dt *= 0.5; }
		}
		if (!ok) { X.st.status = 2; return 2; }
		X.t = X.t + dt; X.dtprev = dt; X.st.steps++; X.st.dt_last = dt; done_steps++; if (!ramping) X.startup = false;
	}
	X.st.t = X.t;
	return 0;
}

// ================================================================================================ explicit HLLC scheme
// The reference's fully explicit scheme, src/SPB6_explicit.cpp ("EX"), at relaxation level 0: minmod reconstruction of
// the primitive variables (EX:228-256), HLLC fluxes with the phase Hugoniot states in the star region (EX:145-192), the
// Godunov update with the non-conservative alpha_1 and phase-energy terms (EX:194-225), then per cell the mechanical
// relaxation by the p closure with the shock budget's defect split (EX:305-361), in Heun's two-stage scheme with the
// reject-and-halve time loop at the acoustic CFL (EX:537-566). The IMEX run's statistics (steps, rejects) are reused;
// hllc_cl_fail counts the relaxation's closure failures (st_cl_fail).
struct HFace { double F[5], ustar, ua1, ue[2]; };                                                       // EX:145
static double hugoniot_p(const Model& M, int k, double p, double v, double vs) {                     // EX:147-157 (SG / NASG)
	const Phase& P = M.PH[k];
	const double num = (p + P.g * P.pi) * (v - P.b) - P.g * P.pi * (vs - P.b) - 0.5 * (P.g - 1.0) * p * (vs - v);
	const double den = (vs - P.b) + 0.5 * (P.g - 1.0) * (vs - v);
	return num / den;
}
static HFace hllc(const Model& M, const Prim& L, const Prim& R, int dir) {                          // EX:159-192
	auto un = [&](const Prim& w) { return dir == 0 ? w.u : w.v; };
	auto ut = [&](const Prim& w) { return dir == 0 ? w.v : w.u; };
	const double uL = un(L), uR = un(R), pL = L.a1 * L.p1 + (1 - L.a1) * L.p2, pR = R.a1 * R.p1 + (1 - R.a1) * R.p2;
	const double SL = smin(uL - L.cf, uR - R.cf), SR = smax(uL + L.cf, uR + R.cf);
	const double SM = (pR - pL + L.rho * uL * (SL - uL) - R.rho * uR * (SR - uR)) / (L.rho * (SL - uL) - R.rho * (SR - uR));
	auto EK = [&](const Prim& w) { const double a2 = 1 - w.a1;
		return eps_amp(M, 0, w.a1, w.a1 * w.r1, w.p1) + eps_amp(M, 1, a2, a2 * w.r2, w.p2) + 0.5 * w.rho * (w.u * w.u + w.v * w.v); };
	auto flux = [&](const Prim& w, double* F) { const double u = un(w), p = w.a1 * w.p1 + (1 - w.a1) * w.p2;
		F[0] = w.a1 * w.r1 * u; F[1] = (1 - w.a1) * w.r2 * u; F[2] = w.rho * u * u + p; F[3] = w.rho * u * ut(w); F[4] = (EK(w) + p) * u; };
	HFace f;
	const bool left = SM >= 0; const Prim& K = left ? L : R;
	const double uK = left ? uL : uR, SK = left ? SL : SR, pK = left ? pL : pR;
	double FK[5]; flux(K, FK);
	if ((left && SL >= 0) || (!left && SR <= 0)) {
		for (int q = 0; q < 5; q++) f.F[q] = FK[q];
		f.ustar = uK; f.ua1 = uK * K.a1;
		f.ue[0] = eps_amp(M, 0, K.a1, K.a1 * K.r1, K.p1) * uK; f.ue[1] = eps_amp(M, 1, 1 - K.a1, (1 - K.a1) * K.r2, K.p2) * uK;
		return f;
	}
	const double fac = (SK - uK) / (SK - SM), EKv = EK(K);
	const double rs = K.rho * fac;
	const double Es = fac * (EKv + (SM - uK) * (K.rho * SM + pK / (SK - uK)));
	const double UK[5] = {K.a1 * K.r1, (1 - K.a1) * K.r2, K.rho * uK, K.rho * ut(K), EKv};
	const double Us[5] = {UK[0] * fac, UK[1] * fac, rs * SM, rs * ut(K), Es};
	for (int q = 0; q < 5; q++) f.F[q] = FK[q] + SK * (Us[q] - UK[q]);
	f.ustar = SM; f.ua1 = SM * K.a1;
	const double ak[2] = {K.a1, 1 - K.a1}, rk[2] = {K.r1, K.r2}, pk[2] = {K.p1, K.p2};
	for (int k = 0; k < 2; k++) {
		const double v0 = 1.0 / rk[k], vs = 1.0 / (rk[k] * fac);
		double ps = hugoniot_p(M, k, pk[k], v0, vs);
		if (!std::isfinite(ps) || !p_adm_sg(M, k, ps)) ps = pk[k];
		f.ue[k] = eps_amp(M, k, ak[k], ak[k] * rk[k] * fac, ps) * SM;
	}
	return f;
}
// cell primitives of every ghosted cell, then the minmod face states per direction (EX:228-256); false where prim_of fails
static bool hllc_recon(Ctx& X, const std::vector<double>& U, std::vector<Prim>& w, std::vector<Prim>* WL, std::vector<Prim>* WR) {
	const int NXT = X.NXT, NYT = X.NYT; const size_t NC = (size_t)NXT * NYT;
	w.assign(NC, Prim{}); bool ok = true;
	#pragma omp parallel for schedule(static) reduction(&&:ok)
	for (size_t c = 0; c < NC; c++) if (!prim_of(X.M, &U[NV * c], w[c])) ok = false;
	if (!ok) return false;
	for (int d = 0; d < 2; d++) { WL[d] = w; WR[d] = w; }
	if (X.C.order < 2) return true;
	auto mm = [&](double a, double b) { if (!X.C.limiter) return 0.5 * (a + b); return a * b > 0 ? (std::fabs(a) < std::fabs(b) ? a : b) : 0.0; };
	for (int dir = 0; dir < (X.ny > 1 ? 2 : 1); dir++)
		#pragma omp parallel for schedule(static)
		for (int j = 0; j < NYT; j++) for (int i = 0; i < NXT; i++) {
			int im = i, ip = i, jm = j, jp = j;
			if (dir == 0) { im = i - 1; ip = i + 1; if (im < 0 || ip >= NXT) continue; }
			else { jm = j - 1; jp = j + 1; if (jm < 0 || jp >= NYT) continue; }
			const Prim &c = w[X.id(i, j)], &a = w[X.id(im, jm)], &b = w[X.id(ip, jp)];
			const double cv[7] = {c.a1, c.r1, c.r2, c.u, c.v, c.p1, c.p2}, av[7] = {a.a1, a.r1, a.r2, a.u, a.v, a.p1, a.p2};
			const double bv[7] = {b.a1, b.r1, b.r2, b.u, b.v, b.p1, b.p2};
			double l[7], r[7];
			for (int q = 0; q < 7; q++) { const double sl = mm(cv[q] - av[q], bv[q] - cv[q]); l[q] = cv[q] - 0.5 * sl; r[q] = cv[q] + 0.5 * sl; }
			Prim PL = c, PR = c;
			PL.a1 = l[0]; PL.r1 = l[1]; PL.r2 = l[2]; PL.u = l[3]; PL.v = l[4]; PL.p1 = l[5]; PL.p2 = l[6];
			PR.a1 = r[0]; PR.r1 = r[1]; PR.r2 = r[2]; PR.u = r[3]; PR.v = r[4]; PR.p1 = r[5]; PR.p2 = r[6];
			for (Prim* P : {&PL, &PR}) {
				const double m1 = P->a1 * P->r1, m2 = (1 - P->a1) * P->r2; P->rho = m1 + m2;
				P->cf = std::sqrt((m1 * c2_rp(X.M, 0, P->r1, P->p1) + m2 * c2_rp(X.M, 1, P->r2, P->p2)) / P->rho);
			}
			WL[dir][X.id(i, j)] = PL; WR[dir][X.id(i, j)] = PR;
		}
	return true;
}
// Godunov update u_new = u + dU of the interior cells (EX:194-225); false where alpha_1, m_1 or m_2 leave their range
static bool hllc_update(Ctx& X, std::vector<double>& Unew, const std::vector<double>& U, const std::vector<Prim>& w,
                        const std::vector<Prim>* WL, const std::vector<Prim>* WR, double h) {
	const int nx = X.nx, ny = X.ny, NXT = X.NXT, NYT = X.NYT;
	const double rx = h / X.dx, ry = h / X.dy;
	std::vector<HFace> FX((size_t)NXT * NYT), FY;
	#pragma omp parallel for schedule(static)
	for (int j = GH; j < ny + GH; j++) for (int i = GH - 1; i < nx + GH; i++) FX[X.id(i, j)] = hllc(X.M, WR[0][X.id(i, j)], WL[0][X.id(i + 1, j)], 0);
	if (ny > 1) {
		FY.resize((size_t)NXT * NYT);
		#pragma omp parallel for schedule(static)
		for (int j = GH - 1; j < ny + GH; j++) for (int i = GH; i < nx + GH; i++) FY[X.id(i, j)] = hllc(X.M, WR[1][X.id(i, j)], WL[1][X.id(i, j + 1)], 1);
	}
	Unew = U;
	bool ok = true;
	#pragma omp parallel for schedule(static) reduction(&&:ok)
	for (int j = GH; j < ny + GH; j++) for (int i = GH; i < nx + GH; i++) {
		const Prim& wc = w[X.id(i, j)];
		const HFace &fp = FX[X.id(i, j)], &fm = FX[X.id(i - 1, j)];
		double dU[NV] = {0};
		const double ak[2] = {wc.a1, 1 - wc.a1}, pk[2] = {wc.p1, wc.p2};
		dU[IM1] = -rx * (fp.F[0] - fm.F[0]); dU[IM2] = -rx * (fp.F[1] - fm.F[1]); dU[IMU] = -rx * (fp.F[2] - fm.F[2]);
		dU[IMV] = -rx * (fp.F[3] - fm.F[3]); dU[IEN] = -rx * (fp.F[4] - fm.F[4]);
		dU[IA1] = -rx * ((fp.ua1 - fm.ua1) - wc.a1 * (fp.ustar - fm.ustar));
		for (int k = 0; k < 2; k++) dU[IE1 + k] = -rx * ((fp.ue[k] - fm.ue[k]) + ak[k] * pk[k] * (fp.ustar - fm.ustar));
		if (ny > 1) {
			const HFace &gp = FY[X.id(i, j)], &gm = FY[X.id(i, j - 1)];
			dU[IM1] -= ry * (gp.F[0] - gm.F[0]); dU[IM2] -= ry * (gp.F[1] - gm.F[1]); dU[IMV] -= ry * (gp.F[2] - gm.F[2]);
			dU[IMU] -= ry * (gp.F[3] - gm.F[3]); dU[IEN] -= ry * (gp.F[4] - gm.F[4]);
			dU[IA1] -= ry * ((gp.ua1 - gm.ua1) - wc.a1 * (gp.ustar - gm.ustar));
			for (int k = 0; k < 2; k++) dU[IE1 + k] -= ry * ((gp.ue[k] - gm.ue[k]) + ak[k] * pk[k] * (gp.ustar - gm.ustar));
		}
		double* Q = &Unew[NV * X.id(i, j)];
		for (int k = 0; k < NV; k++) Q[k] = U[NV * X.id(i, j) + k] + dU[k];
		if (!(Q[IA1] > 0 && Q[IA1] < 1 && Q[IM1] > 0 && Q[IM2] > 0)) ok = false;
	}
	return ok;
}
// mechanical relaxation of one cell (EX:305-343 at relaxation level 0, mu = inf): a phase whose transported state is
// inadmissible restarts from the mixture pressure, then the p closure (s = 0, the cell's shock budget) sets alpha and eps
static bool hllc_relax_cell(const Model& M, double* Q, const double Qph[2]) {
	const double rho = Q[IM1] + Q[IM2], rhoe = Q[IEN] - 0.5 * (Q[IMU] * Q[IMU] + Q[IMV] * Q[IMV]) / rho;
	double a1 = Q[IA1], m1 = Q[IM1], m2 = Q[IM2], e1 = Q[IE1], e2 = Q[IE2];
	{
		const double ak[2] = {a1, 1 - a1}, mk[2] = {m1, m2}; double* ek[2] = {&e1, &e2};
		for (int k = 0; k < 2; k++)
			if (!p_adm_sg(M, k, p_ame(M, k, ak[k], mk[k], *ek[k]))) {
				const double pm = p_mix(M, a1, m1, m2, rhoe);
				if (p_adm_sg(M, k, pm)) *ek[k] = eps_amp(M, k, ak[k], mk[k], pm);
			}
	}
	ClIn in{}; in.atr[0] = a1; in.atr[1] = 1 - a1; in.mtr[0] = m1; in.mtr[1] = m2; in.etr[0] = e1; in.etr[1] = e2;
	in.s = 0; in.rhoeC = rhoe; in.pguess = smax(pbar_of(M, Q), 1.0); in.Qph[0] = Qph[0]; in.Qph[1] = Qph[1];
	ClOut o; if (!closure(M, in, o)) return false;
	Q[IA1] = o.a[0]; Q[IM1] = o.m[0]; Q[IM2] = o.m[1]; Q[IE1] = o.e[0]; Q[IE2] = o.e[1];
	return true;
}
static long hllc_cl_fail = 0;
static bool hllc_relax(Ctx& X, std::vector<double>& U) {                                             // EX:344-361
	bc_apply(X, U);
	const int nx = X.nx, n = nx * X.ny;
	std::vector<double> Qph((size_t)n * 2);
	#pragma omp parallel for schedule(static)
	for (int q = 0; q < n; q++) {
		const int i = q % nx + GH, j = q / nx + GH;
		auto cell = [&](int di, int dj) { return &U[NV * X.id(i + di, j + dj)]; };
		shock_budget(X.M, cell, X.twoD, &Qph[(size_t)q * 2]);
	}
	bool ok = true; long fails = 0;
	#pragma omp parallel for schedule(static) reduction(&&:ok) reduction(+:fails)
	for (int q = 0; q < n; q++)
		if (!hllc_relax_cell(X.M, &U[NV * X.id(q % nx + GH, q / nx + GH)], &Qph[(size_t)q * 2])) { ok = false; fails++; }
	hllc_cl_fail += fails;
	if (!ok) return false;
	bc_apply(X, U);
	return true;
}
static bool hllc_stage(Ctx& X, std::vector<double>& U, double h) {                                   // EX:363-372
	bc_apply(X, U);
	std::vector<Prim> w, WL[2], WR[2];
	if (!hllc_recon(X, U, w, WL, WR)) return false;
	std::vector<double> Unew;
	if (!hllc_update(X, Unew, U, w, WL, WR, h)) return false;
	if (!hllc_relax(X, Unew)) return false;
	U.swap(Unew);
	return true;
}
// one step (EX:546-556): first order, one stage; second order, Heun U2 = R(U1 + dt L(U1)) after U1, then (U + U2) / 2 relaxed
static bool hllc_attempt(Ctx& X, double dt) {
	std::vector<double> V = X.U;
	bool done;
	if (X.C.order < 2) done = hllc_stage(X, V, dt);
	else {
		const std::vector<double> W = V;
		done = hllc_stage(X, V, dt) && hllc_stage(X, V, dt);
		if (done) {
			#pragma omp parallel for schedule(static)
			for (size_t q = 0; q < V.size(); q++) V[q] = 0.5 * (V[q] + W[q]);
			done = hllc_relax(X, V);
		}
	}
	if (done) X.U = V;
	return done;
}
// the time loop (EX:537-566): dt = min(cfl * caldt(acoustic), t_end - t), 30 halvings at most
static int hllc_run(Ctx& X, double t_end, long max_steps, double cfl) {
	long done_steps = 0;
	while (X.t < t_end * (1 - 1e-12) && (max_steps < 0 || done_steps < max_steps)) {
		double dt = smin(cfl * caldt(X, true), t_end - X.t);
		if (!(dt > 1e-12 * t_end)) { X.st.status = 3; return 2; }
		bool ok = false;
		for (int att = 0; att < 30 && !ok; att++) {
			ok = hllc_attempt(X, dt);
			if (!ok) { X.st.rejects++; dt *= 0.5; }
		}
		if (!ok) { X.st.status = 2; return 2; }
		X.t = X.t + dt; X.st.steps++; X.st.dt_last = dt; done_steps++;
	}
	X.st.t = X.t;
	return 0;
}

// ------------------------------------------------------------------------------------------------ reference cases (PHY:961-1100 subset)
static void set_SG(Cfg& C, int k, double g, double pi) { C.ph[k][0] = g; C.ph[k][1] = pi; C.ph[k][2] = 0; C.ph[k][3] = 0; C.ph[k][4] = 1000.0; C.ph[k][5] = 0; }   // PHY:958
static bool case_setup(int test, Cfg& C, double& tend) {                                           // PHY:961-1022 (SG cases)
	C.x0 = 0; C.x1 = 1; C.y0 = 0; C.y1 = 1; for (int q = 0; q < 4; q++) { C.bc[q] = BC_TRANS; C.bcu[q] = 0; }
	switch (test) {
	case 0: set_SG(C, 0, 4.4, 6e8); set_SG(C, 1, 1.4, 0); tend = 1.85e-3; break;
	case 1: case 2:   // PHY:965-967: Pelanti-Shyue water cavitation tubes, liquid and vapour with their thermal parameters
		set_SG(C, 0, 2.35, 1e9); C.ph[0][2] = -1167e3; C.ph[0][4] = 1816.0;
		set_SG(C, 1, 1.43, 0.0); C.ph[1][2] = 2030e3; C.ph[1][3] = -23.4e3; C.ph[1][4] = 1040.0;
		tend = test == 1 ? 3.2e-3 : 0.58e-3; break;
	case 3:   // PHY:968-969: Pelanti-Shyue dodecane liquid-vapour shock tube (SG), material CFL 0.05
		set_SG(C, 0, 2.35, 4e8); C.ph[0][2] = -775.269e3; C.ph[0][4] = 1077.7;
		set_SG(C, 1, 1.025, 0.0); C.ph[1][2] = -237.547e3; C.ph[1][3] = -24.4e3; C.ph[1][4] = 1956.45;
		C.cfl = 0.05; tend = 473e-6; break;
	case 31: {   // PHY:980: Pelanti's dodecane shock tube with the NASG liquid (covolume 7.51e-4)
		const double l[6] = {1.206, 1681e5, -996054.0, 0.0, 2532.0, 7.51e-4}, v[6] = {1.021, 0.0, -384592.0, -4301.0, 2274.0, 0.0};
		for (int q = 0; q < 6; q++) { C.ph[0][q] = l[q]; C.ph[1][q] = v[q]; }
		tend = 473e-6; break; }
	case 30: {   // PHY:979: Pelanti's water cavitation tube, NASG liquid (covolume 6.61e-4) and vapour
		const double l[6] = {1.187, 7028e5, -1177788.0, 0.0, 3610.0, 6.61e-4}, v[6] = {1.467, 0.0, 2077616.0, 14317.0, 955.0, 0.0};
		for (int q = 0; q < 6; q++) { C.ph[0][q] = l[q]; C.ph[1][q] = v[q]; }
		tend = 3e-3; break; }
	case 60: case 61: {   // PHY:991-997: Saurel-Petitpas-Berry's nitromethane tube (both phases one Cochran-Chan fluid);
		// Pelanti's detonation gases (JWL) | NASG water tube on [-10, 10] m
		auto mgp = [&](int k, int kind, double G, double c1, double c2, double e1, double e2, double rr, double e0, double g) {
			for (int q = 0; q < 6; q++) C.ph[k][q] = 0.0;
			C.ph[k][0] = g; C.eos_kind[k] = kind;
			const double v[7] = {G, c1, c2, e1, e2, rr, e0}; for (int q = 0; q < 7; q++) C.mg[k][q] = v[q]; };
		if (test == 60) { for (int k = 0; k < 2; k++) mgp(k, 1, 1.19, 0.819181e9, 1.50835e9, 4.52969, 1.42144, 1134.0, 0.0, 2.19); tend = 67e-6; }
		else {
			mgp(0, 2, 0.25, 854.5e9, 20.5e9, 4.6, 1.35, 1840.0, 0.0, 1.25);
			const double w[6] = {1.187, 7028e5, -1177788.0, 0.0, 3610.0, 6.61e-4};
			for (int q = 0; q < 6; q++) C.ph[1][q] = w[q];
			C.x0 = -10.0; C.x1 = 10.0; tend = 0.9e-3;
		}
		break; }
	case 4: set_SG(C, 0, 4.4, 6e8); set_SG(C, 1, 1.4, 0); tend = 2.79e-3; break;
	case 5: set_SG(C, 0, 4.4, 6e8); set_SG(C, 1, 1.4, 0); tend = 240e-6; break;
	case 80: case 82: case 83: case 84: set_SG(C, 0, 1.4, 0); set_SG(C, 1, 1.4, 0); tend = test == 80 ? 0.5 : test == 82 ? 0.14 : test == 83 ? 2.5e-6 : 0.15; break;
	case 81: case 85: set_SG(C, 0, 4.4, 6e8); set_SG(C, 1, 4.4, 6e8); tend = test == 81 ? 1e-4 : 0.15e-3; break;
	case 86: case 87: set_SG(C, 0, 4.4, 6e8); set_SG(C, 1, 1.4, 0); tend = test == 86 ? 0.25 : 241e-6; break;
	case 88: set_SG(C, 0, 2.43, 5.3e9); set_SG(C, 1, 1.62, 141e9); tend = 30e-6; break;
	case 70:   // PHY:998-999: 0.3 x 0.08 m, air / helium ideal gases, walls at the bottom and top, 600 x 160 cells
		set_SG(C, 0, 1.4, 0); set_SG(C, 1, 1.67, 0); C.x1 = 0.3; C.y1 = 0.08; C.bc[2] = C.bc[3] = BC_WALL; tend = 600e-6; break;
	default: return false;
	}
	return true;
}
// sub-cell averaging of the initial state (TestSetup::sub, PHY:898-905): 4 x 4 for case 70, none for the 1D cases
static int case_sub(int test) { return test == 70 ? 4 : 1; }
static void case_state(const Model& M, int test, double x, double y, double* Q) {                 // PHY:1028-1100 (SG cases)
	auto setcell = [&](double a1, double r1, double r2, double u, double v, double p1, double p2) {
		Prim w; w.a1 = a1; w.r1 = r1; w.r2 = r2; w.u = u; w.v = v; w.p1 = p1; w.p2 = p2; prim2cell(M, w, Q); };
	switch (test) {
	case 0: { const double u = x < 0.5 ? -100.0 : 100.0; setcell(1 - 1e-2, 1000.0, 1.0, u, 0, 1e5, 1e5); break; }
	case 1: case 2: {   // PHY:1034-1036: T = 354.728 K, p = 1e5 Pa, u = -U0 | +U0 about x = 0.5, vapour fraction 1e-2
		const double T0 = 354.728, U0 = test == 1 ? 2.0 : 500.0, u = x < 0.5 ? -U0 : U0;
		setcell(1 - 1e-2, 1.0 / v_pT(M, 0, 1e5, T0), 1.0 / v_pT(M, 1, 1e5, T0), u, 0, 1e5, 1e5); break; }
	case 3: case 31: {   // PHY:1039: liquid at 1e8 Pa left of x = 0.75, vapour at 1e5 Pa right of it, at rest
		const bool L = x < 0.75; setcell(L ? 1 - 1e-8 : 1e-8, 500.0, 2.0, 0, 0, L ? 1e8 : 1e5, L ? 1e8 : 1e5); break; }
	case 30: {   // PHY:1085: T = 353 K, p = 1e5 Pa, u = -2 | +2 m/s about x = 0.5, vapour fraction 1e-2
		const double T0 = 353.0, u = x < 0.5 ? -2.0 : 2.0;
		setcell(1 - 1e-2, 1.0 / v_pT(M, 0, 1e5, T0), 1.0 / v_pT(M, 1, 1e5, T0), u, 0, 1e5, 1e5); break; }
	case 4: { const bool L = x < 0.5; setcell(L ? 1 - 1e-8 : 1e-8, 1000.0, 10.0, 100.0, 0, 1e5, 1e5); break; }
	case 60: { const bool L = x < 0.5; setcell(L ? 1 - 1e-8 : 1e-8, 1134.0, 1200.0, 0, 0, L ? 20e9 : 0.2e6, L ? 20e9 : 0.2e6); break; }   // PHY:1040
	case 61: { const bool L = x < 0.0; setcell(L ? 1 - 1e-8 : 1e-8, 2000.0, 1044.0, 0, 0, L ? 4.6406e10 : 1e5, L ? 4.6406e10 : 1e5); break; }   // PHY:1041
	case 5: { const bool L = x < 0.75; setcell(L ? 1 - 1e-6 : 1e-6, 1000.0, 1.0, 0, 0, L ? 1e9 : 1e5, L ? 1e9 : 1e5); break; }
	case 80: case 81: case 82: case 83: case 84: case 85: {
		double xi = 0.5, rL = 0, uL = 0, pL = 0, rR = 0, uR = 0, pR = 0;
		switch (test) {
		case 80: xi = 0.25; rL = 1000; uL = 1; pL = 1e5; rR = 0.01; uR = 1; pR = 1e5; break;
		case 81: rL = 1000; uL = 0; pL = 1e8; rR = 1000; uR = 15; pR = 0.98e8; break;
		case 82: rL = 0.445; uL = 0.698; pL = 3.528; rR = 0.5; uR = 0; pR = 0.571; break;
		case 83: rL = 1; uL = 0; pL = 1e10; rR = 0.125; uR = 0; pR = 0.1; break;
		case 84: rL = 1; uL = -1; pL = 0.4; rR = 1; uR = 1; pR = 0.4; break;
		default: rL = 1000; uL = 0; pL = 1e6; rR = 1000; uR = 0; pR = 1e5; break; }
		const bool L = x < xi; const double r = L ? rL : rR, u = L ? uL : uR, p = L ? pL : pR;
		setcell(1 - 1e-8, r, r, u, 0, p, p); break; }
	case 86: case 87: {
		const bool adv = test == 86, L = x < (adv ? 0.25 : 0.7);
		setcell(L ? 1 - 1e-7 : 1e-7, 1000.0, adv ? 0.01 : 50.0, adv ? 1.0 : 0.0, 0, adv ? 1e5 : (L ? 1e9 : 1e5), adv ? 1e5 : (L ? 1e9 : 1e5)); break; }
	case 88: { const double p = x < 0.6 ? 1e10 : 1e5; setcell(0.5954, 1185.0, 3622.0, 0, 0, p, p); break; }
	case 70: {   // PHY:1042-1048: Mach 1.5 shock in air moving left, its post-shock state right of x = 0.221 m; helium bubble
		const double g = 1.4, Ms = 1.5, p0 = 101325.0, r0 = 1.29, c0 = std::sqrt(g * p0 / r0);
		const double p1 = p0 * (1 + 2 * g / (g + 1) * (Ms * Ms - 1)), r1 = r0 * (g + 1) * Ms * Ms / ((g - 1) * Ms * Ms + 2), u1 = -Ms * c0 * (1 - r0 / r1);
		const bool in = (x - 0.2) * (x - 0.2) + (y - 0.04) * (y - 0.04) < 0.02 * 0.02;
		if (x > 0.221) setcell(1 - 1e-6, r1, 0.167, u1, 0, p1, p1);
		else setcell(in ? 1e-6 : 1 - 1e-6, r0, 0.167, 0, 0, p0, p0);
		break; }
	}
}
static void cell_diag(const Model& M, const double* Q, double* o) {                                 // PHY:1105-1116 (columns p u v rho alpha2 Y2 rho1 rho2 cf cw T1 T2 p1 p2)
	Prim w; const bool ok = prim_of(M, Q, w);
	const double a1 = Q[IA1], m1 = Q[IM1], m2 = Q[IM2], rho = m1 + m2;
	double cw = NAN, T1 = NAN, T2 = NAN;
	if (ok) {
		cw = std::sqrt(1.0 / (rho * (a1 / (w.r1 * c2_rp(M, 0, w.r1, w.p1)) + (1 - a1) / (w.r2 * c2_rp(M, 1, w.r2, w.p2)))));
		T1 = T_amp(M, 0, a1, m1, w.p1); T2 = T_amp(M, 1, 1 - a1, m2, w.p2);
	}
	o[0] = ok ? a1 * w.p1 + (1 - a1) * w.p2 : NAN; o[1] = ok ? w.u : NAN; o[2] = ok ? w.v : NAN; o[3] = rho; o[4] = 1 - a1; o[5] = m2 / rho;
	o[6] = ok ? w.r1 : NAN; o[7] = ok ? w.r2 : NAN; o[8] = ok ? w.cf : NAN; o[9] = cw; o[10] = T1; o[11] = T2; o[12] = ok ? w.p1 : NAN; o[13] = ok ? w.p2 : NAN;
}

}   // namespace ora

// ================================================================================================ C ABI (tests only)
using namespace ora;
extern "C" {

typedef struct { double ph[2][6]; int closure_path, share_start, dot_order, portable_math, shock_fused; int eos_kind[2]; double mg[2][7]; } OraModel;
static Model mk_model(const OraModel* m) {
	Model M; for (int k = 0; k < 2; k++) M.PH[k] = phase_of(m->ph[k], m->eos_kind[k], m->mg[k]);
	M.closure_path = m->closure_path; M.share_start = m->share_start; M.dot_order = m->dot_order; M.portable = m->portable_math;
	M.shock_fused = m->shock_fused; return M;
}

int ora_version(void) { return 1; }
int ora_set_threads(int n) { if (n > 0) omp_set_num_threads(n); return omp_get_max_threads(); }

// NULL for a Mie-Grueneisen phase with a relaxation level above 0 (the reference fails every non-mechanical closure of
// such a phase, PHY:790), an IMEX variant or the fused profile form (both written for NASG phases only)
void* ora_create(const Cfg* C) {
	if ((C->eos_kind[0] || C->eos_kind[1]) && (C->relax || C->variant || C->shock_fused || C->mu_rate > 0)) return nullptr;
	Ctx* X = new Ctx(); init_ctx(*X, *C); return X; }
void ora_destroy(void* h) { delete (Ctx*)h; }
size_t ora_cfg_size(void) { return sizeof(Cfg); }
size_t ora_stats_size(void) { return sizeof(Stats); }

// interior state in AoS order U[(i + j nx) * 8 + k]; applies the boundary fill and the initial pressure (PC:1157-1227)
void ora_set_state(void* h, const double* Uin) {
	Ctx& X = *(Ctx*)h;
	for (int j = 0; j < X.ny; j++) for (int i = 0; i < X.nx; i++) for (int k = 0; k < NV; k++) X.u(X.U, i + GH, j + GH, k) = Uin[((size_t)i + (size_t)j * X.nx) * NV + k];
	bc_apply(X, X.U); init_pressure(X);
	X.t = 0.0; X.dtprev = -1.0; X.startup = true; X.st = Stats{};
}
void ora_get_state(void* h, double* Uout) {
	Ctx& X = *(Ctx*)h;
	for (int j = 0; j < X.ny; j++) for (int i = 0; i < X.nx; i++) for (int k = 0; k < NV; k++) Uout[((size_t)i + (size_t)j * X.nx) * NV + k] = X.u(X.U, i + GH, j + GH, k);
}
void ora_get_pressure(void* h, double* P, double* pref) { Ctx& X = *(Ctx*)h; std::memcpy(P, X.P.data(), X.P.size() * sizeof(double)); *pref = X.P_REF; }
int ora_run(void* h, double t_end, long max_steps) { return run(*(Ctx*)h, t_end, max_steps); }
// RelaxationActivation on the state (the reference's main applies it once more before its output, PC:1300)
void ora_relax_activate(void* h) { Ctx& X = *(Ctx*)h; if (X.C.relax > 0) relaxation_activation(X, X.U); }
// the context's saturation line (relaxation levels 2-3): T_sat(p), and (pmin, pcrit, vcrit, table points); NaN / -1 without it
double ora_tsat(void* h, double p) { Ctx& X = *(Ctx*)h; return X.M.sat ? Tsat(X.M, p) : NAN; }
int ora_sat_info(void* h, double* out) {
	Ctx& X = *(Ctx*)h; if (!X.M.sat) return -1;
	out[0] = X.SAT.pmin; out[1] = X.SAT.pcrit; out[2] = X.SAT.vcrit; return (int)X.SAT.lp.size();
}
// the explicit HLLC scheme's time loop on the context's state (cfl: the acoustic CFL, 0.4 in the reference); the
// relaxation's closure failures accumulate in ora_hllc_closure_fails
int ora_hllc_run(void* h, double t_end, long max_steps, double cfl) { return hllc_run(*(Ctx*)h, t_end, max_steps, cfl); }
long ora_hllc_closure_fails(void) { return hllc_cl_fail; }
// one step of size dt (0: accepted; 1: the attempt failed, the state is unchanged)
int ora_hllc_step(void* h, double dt) {
	Ctx& X = *(Ctx*)h;
	if (!hllc_attempt(X, dt)) return 1;
	X.t = X.t + dt; X.st.steps++; X.st.dt_last = dt; X.st.t = X.t;
	return 0;
}
// resume from another solver's state: pressure unknown and P_REF, and the driver clock (t, previous dt, start-up flag)
void ora_set_pressure(void* h, const double* P, double pref) { Ctx& X = *(Ctx*)h; std::memcpy(X.P.data(), P, X.P.size() * sizeof(double)); X.P_REF = pref; }
void ora_set_clock(void* h, double t, double dtprev, int startup) { Ctx& X = *(Ctx*)h; X.t = t; X.dtprev = dtprev; X.startup = startup != 0; }
void ora_get_stats(void* h, Stats* s) { *s = ((Ctx*)h)->st; }
double ora_time(void* h) { return ((Ctx*)h)->t; }

// one stage of explicit + implicit work on the current state with step adt (rk=1 stage): exposes the intermediate fields
int ora_stage_explicit(void* h, double adt) {
	Ctx& X = *(Ctx*)h; X.Un = X.U; X.Qe = X.U; X.Qs = X.U; X.Pst = X.P;
	tvd_reconstruction(X, X.Qe);
	return explicit_update(X, adt) ? 0 : 1;
}
int ora_stage_prep(void* h, double adt) { Ctx& X = *(Ctx*)h; X.tau = adt; X.rx = adt / X.dx; X.ry = adt / X.dy; return stage_prep(X) ? 0 : 1; }
int ora_stage_newton(void* h, int* iters, double* res) { Ctx& X = *(Ctx*)h; int it; double r; const bool ok = newton(X, X.Pst, it, r); *iters = it; *res = r; return ok ? 0 : 1; }
// residual of every cell at pressure P (interior, nx*ny); returns the number of failed cells
int ora_residual_all(void* h, const double* P, double* G) {
	Ctx& X = *(Ctx*)h; const int n = X.nx * X.ny; std::vector<double> Pv(P, P + n); int bad = 0;
	for (int q = 0; q < n; q++) { double g; if (!residual_at(X, Pv, q % X.nx, q / X.nx, 0, -1, 0, g)) { G[q] = NAN; bad++; } else G[q] = g; }
	return bad;
}
// finite-difference Jacobian rows at P with residual G: val[5*n] (self, W, E, S, N), sc[n]
int ora_jacobian(void* h, const double* P, const double* G, double* val, double* scout) {
	Ctx& X = *(Ctx*)h; const int nx = X.nx, ny = X.ny, n = nx * ny; std::vector<double> Pv(P, P + n); int bad = 0;
	for (int q = 0; q < n; q++) {
		const int I = q % nx + GH, J = q / nx + GH; const double K = X.Kw[X.id(I, J)];
		double D = X.rx * (X.kpx[X.id(I, J)] + X.kpx[X.id(I - 1, J)]);
		if (ny > 1) D += X.ry * (X.kpy[X.id(I, J)] + X.kpy[X.id(I, J - 1)]);
		scout[q] = (std::fabs(X.P_REF + Pv[q]) + K + 1.0) / (1.0 + K * D);
	}
	for (int r = 0; r < n; r++) {
		const int i = r % nx, j = r / nx; const double hh = 1e-7 * scout[r];
		for (int w = 0; w < 5; w++) val[5 * r + w] = 0.0;
		double g1; if (!residual_at(X, Pv, i, j, hh, -1, 0, g1)) { bad++; continue; }
		val[5 * r] = (g1 - G[r]) / hh;
		const int dd[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
		for (int w = 0; w < 4; w++) {
			if (ny == 1 && w >= 2) break;
			const int c = M_P(X, i + GH + dd[w][0], j + GH + dd[w][1]);
			const double hn = 1e-7 * scout[c];
			double g2; if (!residual_at(X, Pv, i, j, 0, w, hn, g2)) { bad++; break; }
			val[5 * r + 1 + w] = (g2 - G[r]) / hn;
		}
	}
	return bad;
}
// BiCGStab / Thomas on the 5-point (2D) or 3-point (1D) pattern of the current grid: val[5n] as from ora_jacobian
int ora_linsolve(void* h, const double* val, const double* b, double* x, int use_thomas, int* iters) {
	Ctx& X = *(Ctx*)h; const int nx = X.nx, ny = X.ny, n = nx * ny; std::vector<SpRow> A(n);
	const int dd[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
	for (int r = 0; r < n; r++) {
		const int i = r % nx, j = r / nx; SpRow& row = A[r]; row.n = 0; row.col[row.n] = r; row.val[row.n++] = val[5 * r];
		for (int w = 0; w < 4; w++) { if (ny == 1 && w >= 2) break; row.col[row.n] = M_P(X, i + GH + dd[w][0], j + GH + dd[w][1]); row.val[row.n++] = val[5 * r + 1 + w]; }
	}
	std::vector<double> bv(b, b + n), xv(n);
	int it = 0; bool ok = true;
	if (use_thomas) thomas(A, bv, xv); else ok = bicgstab(X, xv, bv, A, it);
	std::memcpy(x, xv.data(), n * sizeof(double)); *iters = it; return ok ? 0 : 1;
}
// named intermediate fields (with ghosts, NXT*NYT, or NXT*NYT*8 for grids)
const double* ora_field(void* h, const char* name) {
	Ctx& X = *(Ctx*)h; const std::string s(name);
	if (s == "U") return X.U.data(); if (s == "Ust") return X.Ust.data(); if (s == "Qe") return X.Qe.data(); if (s == "Qs") return X.Qs.data();
	if (s == "wxL") return X.wxL.data(); if (s == "wxR") return X.wxR.data(); if (s == "wyL") return X.wyL.data(); if (s == "wyR") return X.wyR.data();
	if (s == "phx") return X.phx.data(); if (s == "phy") return X.phy.data(); if (s == "a2s") return X.a2s.data(); if (s == "ubx") return X.ubx.data(); if (s == "uby") return X.uby.data();
	if (s == "ust") return X.ust.data(); if (s == "vst") return X.vst.data(); if (s == "cz") return X.cz.data(); if (s == "Kw") return X.Kw.data();
	if (s == "ubsx") return X.ubsx.data(); if (s == "kpx") return X.kpx.data(); if (s == "wLx") return X.wLx.data(); if (s == "pvx") return X.pvx.data();
	if (s == "ubsy") return X.ubsy.data(); if (s == "kpy") return X.kpy.data(); if (s == "wLy") return X.wLy.data(); if (s == "pvy") return X.pvy.data();
	if (s == "Qph") return X.Qph.data(); if (s == "Pst") return X.Pst.data(); if (s == "P") return X.P.data();
	return nullptr;
}
double ora_pref(void* h) { return ((Ctx*)h)->P_REF; }

// reference test cases (SG): fills cfg geometry, boundaries and EOS and the interior state U[(i + j nx) * 8] at the cfg's
// nx x ny cells (u_initialization, PC:450-465: cell centres, or the average of the conserved variables over SUB x SUB
// sub-points where the case asks for it); returns t_end (< 0: unknown case)
double ora_case_init(int test, Cfg* C, double* U) {
	double tend; if (!case_setup(test, *C, tend)) return -1.0;
	// the initial state as the reference's test_state builds it: libm in the Mie-Grueneisen reference curves
	Model M; for (int k = 0; k < 2; k++) M.PH[k] = phase_of(C->ph[k], C->eos_kind[k], C->mg[k]);
	M.portable = 0;
	const double dx = (C->x1 - C->x0) / C->nx, dy = (C->y1 - C->y0) / C->ny;
	const int sub = case_sub(test);
	for (int j = 0; j < C->ny; j++) for (int i = 0; i < C->nx; i++) {
		const double x = C->x0 + (i + 0.5) * dx, y = C->y0 + (j + 0.5) * dy;
		double* Q = &U[((size_t)i + (size_t)j * C->nx) * NV];
		if (sub <= 1) { case_state(M, test, x, y, Q); continue; }
		for (int k = 0; k < NV; k++) Q[k] = 0.0;
		for (int a = 0; a < sub; a++) for (int b = 0; b < sub; b++) {
			double Qs[NV];
			case_state(M, test, x + ((a + 0.5) / sub - 0.5) * dx, y + ((b + 0.5) / sub - 0.5) * dy, Qs);
			for (int k = 0; k < NV; k++) Q[k] += Qs[k] / (sub * sub);
		}
	}
	return tend;
}
// profile columns p u v rho alpha2 Y2 rho1 rho2 cf cw T1 T2 p1 p2 (14 per cell)
// phase k's EOS at (alpha, m) for the tests: out = {eps_amp(p), p_ame(e), c2_rp(m / alpha, p), dedp_am, T_amp(p),
// pr, pr', er of the Mie-Grueneisen reference curve at m / alpha (NaN for NASG)}
void ora_eos(const OraModel* mo, int k, double a, double m, double p, double e, double* out) {
	const Model M = mk_model(mo);
	out[0] = eps_amp(M, k, a, m, p); out[1] = p_ame(M, k, a, m, e); out[2] = c2_rp(M, k, m / a, p); out[3] = dedp_am(M, k, a, m);
	out[4] = T_amp(M, k, a, m, p);
	out[5] = out[6] = out[7] = NAN;
	if (M.is_mg(k)) mg_ref(M, k, m / a, out[5], out[6], out[7]);
}
void ora_cell_diag(const OraModel* m, const double* U, int ncell, double* out) { const Model M = mk_model(m); for (int c = 0; c < ncell; c++) cell_diag(M, &U[(size_t)c * NV], &out[(size_t)c * 14]); }

// ---- per-function entry points (unit tests)
double ora_p_mix(const OraModel* m, double a1, double m1, double m2, double rhoe) { return p_mix(mk_model(m), a1, m1, m2, rhoe); }
// in: atr0 atr1 mtr0 mtr1 etr0 etr1 s rhoeC pguess Qph0 Qph1; out: a0 a1 m0 m1 e0 e1 pbar sig
int ora_closure(const OraModel* m, const double* in, double* out) {
	const Model M = mk_model(m); ClIn c; c.atr[0] = in[0]; c.atr[1] = in[1]; c.mtr[0] = in[2]; c.mtr[1] = in[3]; c.etr[0] = in[4]; c.etr[1] = in[5];
	c.s = in[6]; c.rhoeC = in[7]; c.pguess = in[8]; c.Qph[0] = in[9]; c.Qph[1] = in[10];
	ClOut o; if (!closure(M, c, o)) return 0;
	out[0] = o.a[0]; out[1] = o.a[1]; out[2] = o.m[0]; out[3] = o.m[1]; out[4] = o.e[0]; out[5] = o.e[1]; out[6] = o.pbar; out[7] = closure_roundoff(M, c, o.pbar);
	return 1;
}
// Phi(p) = sum alpha_k(p) - 1 of the closure input (for brute-force root checks)
double ora_closure_phi(const OraModel* m, const double* in, double p) {
	const Model M = mk_model(m); ClIn c; c.atr[0] = in[0]; c.atr[1] = in[1]; c.mtr[0] = in[2]; c.mtr[1] = in[3]; c.etr[0] = in[4]; c.etr[1] = in[5];
	c.s = in[6]; c.rhoeC = in[7]; c.pguess = in[8]; c.Qph[0] = in[9]; c.Qph[1] = in[10];
	PCoef q; p_coef(M, c, q, p); return alpha_p(M, q, c, 0, p) + alpha_p(M, q, c, 1, p) - 1.0;
}
int ora_shock_share(const OraModel* m, const double* Y, const double* vpre, double p0, double p1, double* lam, double* q) { return shock_share(mk_model(m), Y, vpre, p0, p1, lam, *q) ? 1 : 0; }
// the closure of a given type at finite rates for the tests: in as ora_closure; type (CL_P 0, CL_MU 1, ...), tau, the
// impedances Z0, Z1, mu_rate (<= 0: inf) and the relaxation form; out: a0 a1 m0 m1 e0 e1 pbar p0 p1
int ora_closure_typed(const OraModel* m, const double* in, int type, double tau, double Z0, double Z1, double mu, int form, double* out) {
	Model M = mk_model(m); M.mu_rate = mu > 0 ? mu : INF; M.relax_form = form;
	ClIn c; c.atr[0] = in[0]; c.atr[1] = in[1]; c.mtr[0] = in[2]; c.mtr[1] = in[3]; c.etr[0] = in[4]; c.etr[1] = in[5];
	c.s = in[6]; c.rhoeC = in[7]; c.pguess = in[8]; c.Qph[0] = in[9]; c.Qph[1] = in[10];
	c.type = type; c.tau = tau; c.Z[0] = Z0; c.Z[1] = Z1;
	ClOut o; if (!closure(M, c, o)) return 0;
	out[0] = o.a[0]; out[1] = o.a[1]; out[2] = o.m[0]; out[3] = o.m[1]; out[4] = o.e[0]; out[5] = o.e[1]; out[6] = o.pbar;
	out[7] = o.p[0]; out[8] = o.p[1];
	return 1;
}
// the context's saturation pressure psat(T) and the nu law's rate at (p, T, alpha_v) (NaN / 0 without the table)
double ora_pexpm1(double x) { return pexpm1(x); }
double ora_psat(void* h, double T) { Ctx& X = *(Ctx*)h; return psat(X.M, T); }
double ora_nu_law(void* h, double p, double T, double av) { Ctx& X = *(Ctx*)h; return nu_law(X.M, p, T, av); }
// the nu law's rates of the last activation (n interior cells; -1 without the law)
int ora_cell_nu(void* h, double* out) { Ctx& X = *(Ctx*)h; if (X.cnu.empty()) return -1; std::memcpy(out, X.cnu.data(), X.cnu.size() * sizeof(double)); return (int)X.cnu.size(); }
double ora_hug_v(const OraModel* m, int k, double p0, double p1, double v0) { return hug_v(mk_model(m), k, p0, p1, v0); }
int ora_prof_integrate(const OraModel* m, const double* Y, const double* vpre, double p0, double p1, double c, double* v1end, double* qend) {
	const Model M = mk_model(m); ShockProf S; S.p0 = p0; double vH[2];
	for (int k = 0; k < 2; k++) { S.Y[k] = Y[k]; S.vk0[k] = vpre[k]; vH[k] = hug_v(M, k, p0, p1, vpre[k]); }
	S.v0 = Y[0] * vpre[0] + Y[1] * vpre[1]; S.vend = Y[0] * vH[0] + Y[1] * vH[1]; S.j2 = (p1 - p0) / (S.v0 - S.vend);
	return (M.shock_fused ? prof_integrate_fused(M, S, c, *v1end, *qend) : prof_integrate(M, S, c, *v1end, *qend)) ? 1 : 0;
}
// the profile right-hand side at (v, y) of the problem ora_prof_integrate sets up, in the model's form; returns its flag
int ora_prof_rhs(const OraModel* m, const double* Y, const double* vpre, double p0, double p1, double c, double v, const double* y, double* dy) {
	const Model M = mk_model(m); ShockProf S; S.p0 = p0; double vH[2];
	for (int k = 0; k < 2; k++) { S.Y[k] = Y[k]; S.vk0[k] = vpre[k]; vH[k] = hug_v(M, k, p0, p1, vpre[k]); }
	S.v0 = Y[0] * vpre[0] + Y[1] * vpre[1]; S.vend = Y[0] * vH[0] + Y[1] * vH[1]; S.j2 = (p1 - p0) / (S.v0 - S.vend);
	if (!M.shock_fused) return prof_rhs(M, S, c, v, y, dy) ? 1 : 0;
	ProfFused F; prof_fused_setup(M, S, c, F);
	return prof_rhs_fused(M, S, F, v, y, dy) ? 1 : 0;
}
// shock budget of one cell from its cross stencil: cells[5*8] = centre, W, E, S, N (S/N ignored when twoD = 0)
void ora_shock_budget(const OraModel* m, const double* cells, int twoD, double* Q) {
	const Model M = mk_model(m);
	auto cell = [&](int di, int dj) { const int k = (di == 0 && dj == 0) ? 0 : di == -1 ? 1 : di == 1 ? 2 : dj == -1 ? 3 : 4; return &cells[k * NV]; };
	shock_budget(M, cell, twoD != 0, Q);
}
double ora_pexp(double x) { return pexp(x); }
double ora_plog(double x) { return plog(x); }
double ora_ppow(double x, double y) { return ppow(x, y); }
// fixed-order sums (dot_order 0/1/2) of a vector
double ora_sum(int dot_order, const double* a, int n) { Model M; M.dot_order = dot_order; return osum(M, n, [&](int r) { return a[r]; }); }

}   // extern "C"
