// spb6_device.h -- per-cell FP64 physics of the SPB 6+1 IMEX step for gfx950 (stiffened-gas phases, p closure).
//
// Every function evaluates its operations in the reference's order (src/spb6_physics.h "PHY", src/spb6_physics_omp.h
// "PHO", src/SPB6_PC_IMEX.cpp "PC") and the library is compiled with -ffp-contract=off, so each result is the same IEEE
// double the reference's x86-64 build produces (division and square root are correctly rounded on both). The only
// transcendental functions of the path (shock_share's exp / log / pow) use the portable routines below, which run the
// same operations on host and device.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define SPB_HD __host__ __device__ __forceinline__

namespace spb {

enum { IA1 = 0, IM1 = 1, IM2 = 2, IMU = 3, IMV = 4, IE1 = 5, IE2 = 6, IEN = 7 };
constexpr int NV = 8, GH = 2;
enum { BC_TRANS = 0, BC_WALL = 1, BC_PERIODIC = 2 };

SPB_HD double dinf() { return __builtin_huge_val(); }
SPB_HD double dnan() { return __builtin_nan(""); }
// std::max / std::min semantics: max(a, b) = (a < b) ? b : a, min(a, b) = (b < a) ? b : a
SPB_HD double smax(double a, double b) { return (a < b) ? b : a; }
SPB_HD double smin(double a, double b) { return (b < a) ? b : a; }
SPB_HD bool isfin(double x) { return __builtin_isfinite(x); }
SPB_HD uint32_t umin(uint32_t a, uint32_t b) { return b < a ? b : a; }
SPB_HD uint32_t umax(uint32_t a, uint32_t b) { return a < b ? b : a; }

// forward-mode derivatives along N directions (value v, derivatives d), for the analytic Newton Jacobian. Division takes
// one reciprocal (values need not match the double path bit for bit here); smax / smin with a constant take the right
// derivative at the kink (the limit of a forward difference with a positive step).
template <int N> struct Dual {
	double v, d[N];
	SPB_HD Dual(double x = 0.0) : v(x) {
		#pragma unroll
		for (int i = 0; i < N; i++) d[i] = 0.0;
	}
	SPB_HD static Dual seed(double x, int i) { Dual r(x); r.d[i] = 1.0; return r; }
};
#define SPB_DUAL_OP(expr_v, expr_d)                                                       \
	{ Dual<N> o_; o_.v = (expr_v);                                                          \
	  _Pragma("unroll") for (int i = 0; i < N; i++) o_.d[i] = (expr_d);                     \
	  return o_; }
template <int N> SPB_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) SPB_DUAL_OP(a.v + b.v, a.d[i] + b.d[i])
template <int N> SPB_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) SPB_DUAL_OP(a.v - b.v, a.d[i] - b.d[i])
template <int N> SPB_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) SPB_DUAL_OP(a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
template <int N> SPB_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
	const double r = 1.0 / b.v, q = a.v * r;
	SPB_DUAL_OP(q, (a.d[i] - q * b.d[i]) * r)
}
template <int N> SPB_HD Dual<N> operator+(const Dual<N>& a, double b) SPB_DUAL_OP(a.v + b, a.d[i])
template <int N> SPB_HD Dual<N> operator+(double a, const Dual<N>& b) SPB_DUAL_OP(a + b.v, b.d[i])
template <int N> SPB_HD Dual<N> operator-(const Dual<N>& a, double b) SPB_DUAL_OP(a.v - b, a.d[i])
template <int N> SPB_HD Dual<N> operator-(double a, const Dual<N>& b) SPB_DUAL_OP(a - b.v, -b.d[i])
template <int N> SPB_HD Dual<N> operator*(const Dual<N>& a, double b) SPB_DUAL_OP(a.v * b, a.d[i] * b)
template <int N> SPB_HD Dual<N> operator*(double a, const Dual<N>& b) SPB_DUAL_OP(a * b.v, a * b.d[i])
template <int N> SPB_HD Dual<N> operator/(const Dual<N>& a, double b) { const double r = 1.0 / b; SPB_DUAL_OP(a.v * r, a.d[i] * r) }
#undef SPB_DUAL_OP
template <int N> SPB_HD Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { a = a - b; return a; }
template <int N> SPB_HD Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { a = a * b; return a; }
template <int N> SPB_HD Dual<N> smax(const Dual<N>& a, double b) {
	if (a.v < b) return Dual<N>(b);
	if (a.v > b) return a;
	Dual<N> r(a.v);
	#pragma unroll
	for (int i = 0; i < N; i++) r.d[i] = smax(a.d[i], 0.0);
	return r;
}
template <int N> SPB_HD Dual<N> smin(const Dual<N>& a, double b) {
	if (b < a.v) return Dual<N>(b);
	if (a.v < b) return a;
	Dual<N> r(a.v);
	#pragma unroll
	for (int i = 0; i < N; i++) r.d[i] = smin(a.d[i], 0.0);
	return r;
}

// ------------------------------------------------------------------ portable exp / log / pow / hypot
// exp: x = k ln2 + r (Cody-Waite, ln2 split in a 32-bit head and a tail), |r| <= ln2/2, Taylor degree 13, exact 2^k.
// log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f/(2+f), log m = 2 atanh s = f - s (f - 2 R),
//      R = sum_{j>=1} s^{2j}/(2j+1); plus e ln2 in two parts.
SPB_HD double pexp(double x) {
	if (x != x) return x;
	if (x > 709.782712893384) return dinf();
	if (x < -745.1332191019412) return 0.0;
	const double k = floor(x * 1.44269504088896338700e+00 + 0.5);
	const double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
	double p = 1.0 / 6227020800.0;
	p = 1.0 / 479001600.0 + r * p;
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
	if (ki > 1000) return ldexp(ldexp(p, ki - 1000), 1000);
	if (ki < -1000) return ldexp(ldexp(p, ki + 1000), -1000);
	return ldexp(p, ki);
}
SPB_HD double plog(double x) {
	if (x != x) return x;
	if (x < 0.0) return dnan();
	if (x == 0.0) return -dinf();
	if (x == dinf()) return dinf();
	int e; double m = frexp(x, &e);
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
	R = z * R;
	const double logm = f - s * (f - 2.0 * R);
	const double de = (double)e;
	return de * 6.93147180369123816490e-01 + (logm + de * 1.90821492927058770002e-10);
}
SPB_HD double ppow(double x, double y) { return pexp(y * plog(x)); }
// e^x - 1: the Taylor series to x^17 / 17! by Horner on |x| < 0.5 (the next term is below 2^-70 relative), else
// pexp(x) - 1 (no cancellation beyond 2 bits there)
SPB_HD double pexpm1(double x) {
	if (!(fabs(x) < 0.5)) return pexp(x) - 1.0;
	double f = 1.0; for (int k = 2; k <= 17; k++) f *= (double)k;   // 17! is exact in FP64
	double q = 1.0 / f;
	for (int k = 16; k >= 1; k--) { f /= (double)(k + 1); q = q * x + 1.0 / f; }
	return q * x;
}
SPB_HD double phypot(double a, double b) {
	if (b == 0.0) return fabs(a);
	if (a == 0.0) return fabs(b);
	return sqrt(a * a + b * b);
}

// ------------------------------------------------------------------ phases (stiffened gas / NASG covolume b; Mie-Grueneisen)
// kind[k] 0: NASG (b = 0: SG) with g, pi, q, qp, cv, b; 1 (Cochran-Chan) and 2 (JWL): a Mie-Grueneisen phase
// p = G rho (e - er(rho)) + pr(rho) with G, c1, c2, e1, e2, rr (reference density) and e0 (PHY:49-53); mg: either phase
// is one. Every MG branch below tests kind, a kernel argument, so a launch takes one side for all its lanes.
// The Mie-Grueneisen fields and branches exist only in the build with SPB6_MG = 1 (libspb6_hip_mg.so): with SPB6_MG = 0
// (libspb6_hip.so) the phase block is the NASG one, is_mg is the constant false, the branches fold away and the kernels
// keep their NASG-only registers and private memory.
#ifndef SPB6_MG
#define SPB6_MG 0
#endif
struct Phys {
	double g[2], pi[2], q[2], b[2], cv[2], qp[2];
#if SPB6_MG
	int kind[2], mg;
	double G[2], c1[2], c2[2], e1[2], e2[2], rr[2], e0[2];
#endif
};
#if SPB6_MG
SPB_HD bool is_mg(const Phys& P, int k) { return P.kind[k] != 0; }
SPB_HD bool mg_any(const Phys& P) { return P.mg != 0; }
SPB_HD double mg_G(const Phys& P, int k) { return P.G[k]; }
#else
SPB_HD constexpr bool is_mg(const Phys&, int) { return false; }
SPB_HD constexpr bool mg_any(const Phys&) { return false; }
SPB_HD constexpr double mg_G(const Phys&, int) { return 1.0; }
#endif
// the Mie-Grueneisen reference curve at density r: pr, pr' and er with er' = pr / r^2 (PHY:58-66), through the portable
// exp and pow (Cochran-Chan: pr = c1 x^e1 - c2 x^e2, x = r / rr; JWL: pr = c1 exp(-e1 rr / r) + c2 exp(-e2 rr / r))
SPB_HD void mg_ref(const Phys& P, int k, double r, double& pr, double& dpr, double& er) {
#if !SPB6_MG
	(void)P; (void)k; (void)r; pr = dpr = er = dnan();
#else
	if (P.kind[k] == 1) {
		const double x = r / P.rr[k], t1 = P.c1[k] * ppow(x, P.e1[k]), t2 = P.c2[k] * ppow(x, P.e2[k]);
		pr = t1 - t2; dpr = (P.e1[k] * t1 - P.e2[k] * t2) / r; er = (t1 / (P.e1[k] - 1.0) - t2 / (P.e2[k] - 1.0)) / r;
	} else {
		const double t1 = P.c1[k] * pexp(-P.e1[k] * P.rr[k] / r), t2 = P.c2[k] * pexp(-P.e2[k] * P.rr[k] / r);
		pr = t1 + t2; dpr = (P.e1[k] * t1 + P.e2[k] * t2) * P.rr[k] / (r * r);
		er = t1 / (P.e1[k] * P.rr[k]) + t2 / (P.e2[k] * P.rr[k]) - P.e0[k];
	}
#endif
}

SPB_HD double eps_amp(const Phys& P, int k, double a, double m, double p) {          // PHY:69-71
	if (is_mg(P, k)) { double pr, dpr, er; mg_ref(P, k, m / a, pr, dpr, er); return m * er + a * (p - pr) / mg_G(P, k); }
	return (p + P.g[k] * P.pi[k]) * (a - m * P.b[k]) / (P.g[k] - 1.0) + m * P.q[k]; }
SPB_HD double p_ame(const Phys& P, int k, double a, double m, double e) {            // PHY:72-74
	if (is_mg(P, k)) { double pr, dpr, er; mg_ref(P, k, m / a, pr, dpr, er); return mg_G(P, k) * (e - m * er) / a + pr; }
	return (e - m * P.q[k]) * (P.g[k] - 1.0) / (a - m * P.b[k]) - P.g[k] * P.pi[k]; }
SPB_HD double c2_rp(const Phys& P, int k, double r, double p) {                      // PHY:83-85
	if (is_mg(P, k)) { double pr, dpr, er; mg_ref(P, k, r, pr, dpr, er); return (mg_G(P, k) + 1.0) * (p - pr) / r + dpr; }
	return P.g[k] * (p + P.pi[k]) / (r * (1.0 - r * P.b[k])); }
SPB_HD double dedp_am(const Phys& P, int k, double a, double m) {                    // PHY:86-88
	if (is_mg(P, k)) return a / mg_G(P, k);
	return (a - m * P.b[k]) / (P.g[k] - 1.0); }
SPB_HD double p_mix(const Phys& P, double a1, double m1, double m2, double rhoe) {  // PHY:90-99
	const double a0 = a1, aa1 = 1.0 - a1;
	double num = rhoe, den = 0.0;
	const double d0 = dedp_am(P, 0, a0, m1);
	if (is_mg(P, 0)) { double pr, dpr, er; mg_ref(P, 0, m1 / a0, pr, dpr, er); num -= m1 * er - a0 * pr / mg_G(P, 0); }
	else num -= m1 * P.q[0] + P.g[0] * P.pi[0] * d0;
	den += d0;
	const double d1 = dedp_am(P, 1, aa1, m2);
	if (is_mg(P, 1)) { double pr, dpr, er; mg_ref(P, 1, m2 / aa1, pr, dpr, er); num -= m2 * er - aa1 * pr / mg_G(P, 1); }
	else num -= m2 * P.q[1] + P.g[1] * P.pi[1] * d1;
	den += d1;
	return num / den;
}
// phase admissibility (PHY:101-104): p + pi > 0 for NASG, c^2 > 0 for a Mie-Grueneisen phase
SPB_HD bool p_adm(const Phys& P, int k, double a, double m, double p) {
	if (is_mg(P, k)) return isfin(p) && c2_rp(P, k, m / a, p) > 0;
	return p + P.pi[k] > 0; }
SPB_HD bool p_adm_sg(const Phys& P, int k, double p) { return p + P.pi[k] > 0; }    // the explicit HLLC scheme's (NASG)
// the part of eps_k's round-off scale besides |eps_k| (PHY:106-109): |m q| for NASG, |m er| + |alpha pr / G| for MG
SPB_HD double eps_round(const Phys& P, int k, double a, double m) {
	if (is_mg(P, k)) { double pr, dpr, er; mg_ref(P, k, m / a, pr, dpr, er); return fabs(m * er) + fabs(a * pr / mg_G(P, k)); }
	return fabs(m * P.q[k]); }

// phase temperature T(alpha, m, p) (PHY:75-77); NaN for a Mie-Grueneisen phase without cv
SPB_HD double T_amp(const Phys& P, int k, double a, double m, double p) {
	if (is_mg(P, k)) {
		if (!(P.cv[k] > 0)) return dnan();
		double pr, dpr, er; mg_ref(P, k, m / a, pr, dpr, er); return (a * (p - pr) / (mg_G(P, k) * m)) / P.cv[k]; }
	return (a / m - P.b[k]) * (p + P.pi[k]) / ((P.g[k] - 1.0) * P.cv[k]); }
// the profile columns of cell_diag (PHY:1105-1116) from the 8 conserved values: p u v rho alpha2 Y2 rho1 rho2 c_frozen
// c_wood T1 T2 p1 p2 g1-g2, NaN where prim_of (PHY:837-847) fails. The Gibbs difference g1 - g2 (g_pT, PHY:78-81) takes
// the device log; every other column is the reference's arithmetic operation for operation.
SPB_HD void cell_diag(const Phys& P, const double* Q, double* o) {
	const double a1 = Q[IA1], a2 = 1 - a1, m1 = Q[IM1], m2 = Q[IM2], rho = m1 + m2;
	const double nan = dnan();
	bool ok = a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0;
	double r1 = nan, r2 = nan, u = nan, v = nan, p1 = nan, p2 = nan, cf = nan;
	if (ok) {
		r1 = m1 / a1; r2 = m2 / a2; u = Q[IMU] / rho; v = Q[IMV] / rho;
		p1 = p_ame(P, 0, a1, m1, Q[IE1]); p2 = p_ame(P, 1, a2, m2, Q[IE2]);
		ok = p_adm(P, 0, a1, m1, p1) && p_adm(P, 1, a2, m2, p2);
		if (ok) cf = sqrt((m1 * c2_rp(P, 0, r1, p1) + m2 * c2_rp(P, 1, r2, p2)) / rho);
	}
	double cw = nan, T1 = nan, T2 = nan, dg = nan;
	if (ok) {
		cw = sqrt(1.0 / (rho * (a1 / (r1 * c2_rp(P, 0, r1, p1)) + (1 - a1) / (r2 * c2_rp(P, 1, r2, p2)))));
		T1 = T_amp(P, 0, a1, m1, p1); T2 = T_amp(P, 1, 1 - a1, m2, p2);
		if (T1 > 0 && T2 > 0 && P.cv[0] > 0 && !mg_any(P)) {
			auto gpt = [&](int k, double pp, double T) {
				const double sk = P.cv[k] * (P.g[k] * log(T) - (P.g[k] - 1.0) * log(pp + P.pi[k])) + P.qp[k];
				return P.g[k] * P.cv[k] * T + P.b[k] * pp + P.q[k] - T * sk;
			};
			dg = gpt(0, p1, T1) - gpt(1, p2, T2);
		}
	}
	o[0] = ok ? a1 * p1 + (1 - a1) * p2 : nan; o[1] = ok ? u : nan; o[2] = ok ? v : nan; o[3] = rho; o[4] = 1 - a1; o[5] = m2 / rho;
	o[6] = ok ? r1 : nan; o[7] = ok ? r2 : nan; o[8] = ok ? cf : nan; o[9] = cw; o[10] = T1; o[11] = T2; o[12] = ok ? p1 : nan;
	o[13] = ok ? p2 : nan; o[14] = dg;
}

// frozen sound speed of prim_of (PHY:837-847); false where prim_of fails
SPB_HD bool prim_cf(const Phys& P, double a1, double m1, double m2, double e1, double e2, double& cf) {
	const double a2 = 1 - a1;
	if (!(a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0)) return false;
	const double rho = m1 + m2, r1 = m1 / a1, r2 = m2 / a2;
	const double p1 = p_ame(P, 0, a1, m1, e1), p2 = p_ame(P, 1, a2, m2, e2);
	if (!(p_adm(P, 0, a1, m1, p1) && p_adm(P, 1, a2, m2, p2))) return false;
	cf = sqrt((m1 * c2_rp(P, 0, r1, p1) + m2 * c2_rp(P, 1, r2, p2)) / rho);
	return true;
}

// ------------------------------------------------------------------ shock budget (PHY:260-418 with PHO:365-374)
struct Prof { double Y0, Y1, vk00, v0, p0, j2, vend; };
SPB_HD bool eos_ok(const Phys& P, int k, double p, double v) {                       // PHY:271-275
	if (!(v > 0) || !isfin(p)) return false;
	if (is_mg(P, k)) return c2_rp(P, k, 1.0 / v, p) > 0;
	return v > P.b[k] && p + P.pi[k] > 0; }
// phase k's Hugoniot volume from (p0, v0) to p1 (PHY:277-292): closed form for NASG; for a Mie-Grueneisen phase the
// root of e(p1, v) - e(p0, v0) + (p0 + p1) / 2 (v - v0) below v0, bracketed by steps of 0.9 v, bisected to 1e-14 v0
SPB_HD double hug_v(const Phys& P, int k, double p0, double p1, double v0) {
	if (!is_mg(P, k)) {
		const double x0 = p0 + P.pi[k], x1 = p1 + P.pi[k];
		return P.b[k] + (v0 - P.b[k]) * ((P.g[k] - 1.0) * x1 + (P.g[k] + 1.0) * x0) / ((P.g[k] + 1.0) * x1 + (P.g[k] - 1.0) * x0);
	}
	auto e_pv = [&](double p, double v) { double pr, dpr, er; mg_ref(P, k, 1.0 / v, pr, dpr, er); return er + v * (p - pr) / mg_G(P, k); };
	const double e0 = e_pv(p0, v0);
	auto hg = [&](double v) { return e_pv(p1, v) - e0 + 0.5 * (p0 + p1) * (v - v0); };
	double hi = v0, lo = v0, gl = hg(v0);
	if (!(gl > 0)) return v0;
	for (int it = 0; it < 80 && gl > 0; it++) { hi = lo; lo *= 0.9; gl = hg(lo); if (!isfin(gl)) return dnan(); }
	if (gl > 0) return dnan();
	for (int it = 0; it < 100; it++) {
		const double mid = 0.5 * (lo + hi), gm = hg(mid);
		if (gm > 0) hi = mid; else lo = mid;
		if (hi - lo <= 1e-14 * v0) break;
	}
	return 0.5 * (lo + hi);
}
// de/dp at fixed v and de/dv at fixed p plus p, of phase k at (p, v) (PHY:266-270)
SPB_HD void eos_ep_ev(const Phys& P, int k, double p, double v, double& ep, double& evp) {
	if (is_mg(P, k)) {
		double pr, dpr, er; const double r = 1.0 / v; mg_ref(P, k, r, pr, dpr, er);
		ep = v / mg_G(P, k); evp = (p - pr) * (1.0 + 1.0 / mg_G(P, k)) + r * dpr / mg_G(P, k); return; }
	ep = (v - P.b[k]) / (P.g[k] - 1.0); evp = P.g[k] * (p + P.pi[k]) / (P.g[k] - 1.0);
}
// right-hand side of the profile ODE in v (PHY:302-310); lr0 = lam_0 / Y_0 and lr1 = lam_1 / Y_1 are the integration's
// constant quotients (the reference evaluates lam_k / Y_k * (p - P) left to right, so hoisting the quotient keeps every
// result). Branch-free: the admissibility tests only decide the returned flag (lanes of one wavefront keep together).
SPB_HD bool prof_rhs(const Phys& P, const Prof& S, double lr0, double lr1, double v, const double* y, double* dy) {
	const double p = y[0], v1 = y[1], v2 = (v - S.Y0 * v1) / S.Y1, Pr = S.p0 + S.j2 * (S.v0 - v);
	const bool adm = eos_ok(P, 0, p, v1) && eos_ok(P, 1, p, v2);
	double num = -1.0, den = 0.0, ep0, evp0, ep1, evp1;
	eos_ep_ev(P, 0, p, v1, ep0, evp0);
	const double r0 = lr0 * (p - Pr);
	num += S.Y0 * r0 / evp0; den += S.Y0 * ep0 / evp0;
	eos_ep_ev(P, 1, p, v2, ep1, evp1);
	const double r1 = lr1 * (p - Pr);
	num += S.Y1 * r1 / evp1; den += S.Y1 * ep1 / evp1;
	dy[0] = num / den; dy[1] = (r0 - ep0 * dy[0]) / evp0; dy[2] = p - Pr;
	return adm && isfin(dy[0]) && isfin(dy[1]);
}
// Shared-reciprocal quotients (the strong-shock kernel's profile right-hand side). The compiler divides in FP64 by
// div_scale, rcp, two Newton-Raphson steps on the reciprocal r, q = a * r, e = fma(-b, q, a), div_fmas and div_fixup.
// When |a| and |b| both lie in [2^-384, 2^384) the two scale steps are identities (the exponent gap stays below 768,
// neither operand nor the quotient comes near the subnormal or the overflow range), div_fmas is the plain fma and
// div_fixup passes the value through, so quot(a, b, rcp_nr(b)) is the same correctly rounded a / b. rcp_nr depends on
// b alone: a divisor with several numerators pays for one reciprocal, and each quotient then costs a multiply and two
// fmas. A zero a gives a zero whose sign may differ from a / b's (-0 / b with b > 0 gives +0).
SPB_HD double rcp_nr(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
	double r = __builtin_amdgcn_rcp(b);
	double e = fma(-b, r, 1.0);
	r = fma(r, e, r);
	e = fma(-b, r, 1.0);
	return fma(r, e, r);
#else
	return 1.0 / b;   // the host never takes the shared-reciprocal path
#endif
}
SPB_HD double quot(double a, double b, double r) { const double q = a * r; return fma(fma(-b, q, a), r, q); }
// true in every lane of the wavefront when x is true in any (the value itself on the host)
SPB_HD bool wave_any(bool x) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __ballot(x) != 0ull;
#else
	return x;
#endif
}
// The exponent field of x where it lies in the high word (bits 20..30; the sign cleared). Bounds on |x| are taken as
// bounds on this field: 2^-k <= |x| < 2^(k+1) exactly when expo_lo(k) <= field <= expo_hi(k); zeros and subnormals
// have field 0, infinities and NaN 0x7ff00000, so they fall outside every band.
SPB_HD uint32_t expo_field(double x) { return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32) & 0x7ff00000u; }
constexpr uint32_t expo_lo(int k) { return (uint32_t)(1023 - k) << 20; }
constexpr uint32_t expo_hi(int k) { return (uint32_t)(1023 + k) << 20; }
SPB_HD bool in_band(double x, int k) { const uint32_t e = expo_field(x); return e >= expo_lo(k) && e <= expo_hi(k); }
constexpr int BAND_LEAF = 120, BAND_CONST = 8;
// what the strong-shock pass holds for all its cells: 1 / (g_k - 1) and whether g_k and g_k - 1 lie in band 8
struct ShockRcp { double rg0, rg1; bool ok; };
SPB_HD ShockRcp shock_rcp(const Phys& P) {
	ShockRcp c;
	c.rg0 = rcp_nr(P.g[0] - 1.0); c.rg1 = rcp_nr(P.g[1] - 1.0);
	c.ok = !mg_any(P) && in_band(P.g[0], BAND_CONST) && in_band(P.g[1], BAND_CONST) && in_band(P.g[0] - 1.0, BAND_CONST) &&
	       in_band(P.g[1] - 1.0, BAND_CONST);
	return c;
}
// prof_rhs with the ten quotients by Y1, g0 - 1, g1 - 1, evp0 and evp1 taken from shared reciprocals (rY1 = rcp_nr(Y1)
// is the step's, rc the pass's); num / den stays a division. Band k means 2^-k <= |x| < 2^(k+1). When rc.ok, Y0, Y1,
// lr0, lr1 lie in band 120 (both checked by the caller) and the leaves v - Y0 v1, v1 - b0, v2 - b1, p + pi0, p + pi1,
// r0 - ep0 dy0 and p - P (which may also be zero) lie in band 120, every numerator lies within [2^-360, 2^363) (the
// widest, Y_k lr_k (p - P), is a product of three such factors; ep_k and g_k (p + pi_k) lie within [2^-128, 2^130)) and
// every divisor (Y1, g_k - 1, evp_k) within [2^-136, 2^139): both operands of each quotient lie in [2^-384, 2^384), so
// it is a / b bit for bit. At p = P the two lam terms have zero numerators; their quotients, zeros of either sign, are
// added to num = -1 and leave it unchanged. Where a leaf is out of its band (NaN and infinities included), fast is
// cleared and the caller evaluates the step's stages again with prof_rhs. The band test reads the exponent fields
// with integer operations (half the issue cost of FP64 compares).
SPB_HD bool prof_rhs_rcp(const Phys& P, const Prof& S, const ShockRcp& rc, double rY1, double lr0, double lr1, double v,
                         const double* y, double* dy, bool& fast) {
	const double p = y[0], v1 = y[1], a1 = v - S.Y0 * v1;
	const double v2 = quot(a1, S.Y1, rY1), Pr = S.p0 + S.j2 * (S.v0 - v), dp = p - Pr;
	const bool adm = eos_ok(P, 0, p, v1) && eos_ok(P, 1, p, v2);
	const double g0m = P.g[0] - 1.0, g1m = P.g[1] - 1.0;
	const double n0 = v1 - P.b[0], pp0 = p + P.pi[0];
	const double ep0 = quot(n0, g0m, rc.rg0), evp0 = quot(P.g[0] * pp0, g0m, rc.rg0), re0 = rcp_nr(evp0);
	const double r0 = lr0 * dp;
	double num = -1.0, den = 0.0;
	num += quot(S.Y0 * r0, evp0, re0); den += quot(S.Y0 * ep0, evp0, re0);
	const double n1 = v2 - P.b[1], pp1 = p + P.pi[1];
	const double ep1 = quot(n1, g1m, rc.rg1), evp1 = quot(P.g[1] * pp1, g1m, rc.rg1), re1 = rcp_nr(evp1);
	const double r1 = lr1 * dp;
	num += quot(S.Y1 * r1, evp1, re1); den += quot(S.Y1 * ep1, evp1, re1);
	dy[0] = num / den;
	const double n2 = r0 - ep0 * dy[0];
	dy[1] = quot(n2, evp0, re0); dy[2] = dp;
	const uint32_t ea = expo_field(a1), eb = expo_field(n0), ec = expo_field(n1), ed = expo_field(pp0);
	const uint32_t ee = expo_field(pp1), ef = expo_field(n2), eg = expo_field(dp), egl = dp == 0.0 ? expo_lo(BAND_LEAF) : eg;
	const uint32_t lo = umin(umin(umin(ea, eb), umin(ec, ed)), umin(umin(ee, ef), egl));
	const uint32_t hi = umax(umax(umax(ea, eb), umax(ec, ed)), umax(umax(ee, ef), eg));
	fast = fast && lo >= expo_lo(BAND_LEAF) && hi <= expo_hi(BAND_LEAF);
	return adm && isfin(dy[0]) && isfin(dy[1]);
}
// adaptive Cash-Karp RK4(5) integration of the relaxed shock profile (PHY:311-341), written as a start and a single
// pass of the integration loop so that a caller can interleave the passes of many integrations (the strong-shock
// kernel advances one integration per lane, one pass at a time); prof_integrate runs the passes back to back
struct ProfRun { double y0, y1, y2, v, h, lr0, lr1; int it; };
// start the integration at share c (lam = (c, 1 - c))
SPB_HD void prof_begin(const Prof& S, double c, ProfRun& R) {
	R.y0 = S.p0; R.y1 = S.vk00; R.y2 = 0.0; R.v = S.v0; R.h = (S.vend - S.v0) / 32.0; R.it = 0;
	R.lr0 = c / S.Y0; R.lr1 = (1.0 - c) / S.Y1;
}
// one pass of the loop body: 0 = not done yet, 1 = reached v_end (v1end = R.y1, qend = R.y2), -1 = failed. RCP takes
// the right-hand side's quotients from shared reciprocals (prof_rhs_rcp with the pass's rc), with the same results, and
// may return 2 in every lane of the wavefront, R unchanged: then the step is to be taken with prof_step<false>.
template <bool RCP = false>
SPB_HD int prof_step(const Phys& P, const Prof& S, ProfRun& R, const ShockRcp* rc = nullptr) {
	constexpr double A1 = 0.2, A2 = 0.3, A3 = 0.6, A4 = 1.0, A5 = 0.875;
	constexpr double B10 = 0.2;
	constexpr double B20 = 3.0 / 40, B21 = 9.0 / 40;
	constexpr double B30 = 0.3, B31 = -0.9, B32 = 1.2;
	constexpr double B40 = -11.0 / 54, B41 = 2.5, B42 = -70.0 / 27, B43 = 35.0 / 27;
	constexpr double B50 = 1631.0 / 55296, B51 = 175.0 / 512, B52 = 575.0 / 13824, B53 = 44275.0 / 110592, B54 = 253.0 / 4096;
	constexpr double C50 = 37.0 / 378, C52 = 250.0 / 621, C53 = 125.0 / 594, C55 = 512.0 / 1771;
	constexpr double C40 = 2825.0 / 27648, C42 = 18575.0 / 48384, C43 = 13525.0 / 55296, C44 = 277.0 / 14336, C45 = 0.25;
	if (R.it >= 4000) return -1;
	R.it++;
	const double L = S.vend - S.v0;
	const double sc[3] = {fabs(S.p0) + S.j2 * fabs(L), S.vk00, S.j2 * L * L + 1e-300};
	const double y[3] = {R.y0, R.y1, R.y2}, v = R.v;
	double h = R.h;
	const bool last = (L < 0 && v + h <= S.vend) || (L > 0 && v + h >= S.vend);
	if (last) h = S.vend - v;
	// The six stages are evaluated unconditionally: where a stage is inadmissible the reference stops evaluating them
	// and rejects the step, and so does this code (the later stages' values are not used), without a branch per stage.
	// Stage inputs are accumulated term by term in the reference's order over q. v + 0.0 * h is v (h is finite).
	const double lr0 = R.lr0, lr1 = R.lr1;
	double k0[3], k1[3], k2[3], k3[3], k4[3], k5[3], yt[3];
	auto stages = [&](auto rhs) {
		bool ok = rhs(v, y, k0);
		for (int n = 0; n < 3; n++) { yt[n] = y[n]; yt[n] += h * B10 * k0[n]; }
		ok = rhs(v + A1 * h, yt, k1) && ok;
		for (int n = 0; n < 3; n++) { yt[n] = y[n]; yt[n] += h * B20 * k0[n]; yt[n] += h * B21 * k1[n]; }
		ok = rhs(v + A2 * h, yt, k2) && ok;
		for (int n = 0; n < 3; n++) { yt[n] = y[n]; yt[n] += h * B30 * k0[n]; yt[n] += h * B31 * k1[n]; yt[n] += h * B32 * k2[n]; }
		ok = rhs(v + A3 * h, yt, k3) && ok;
		for (int n = 0; n < 3; n++) { yt[n] = y[n]; yt[n] += h * B40 * k0[n]; yt[n] += h * B41 * k1[n]; yt[n] += h * B42 * k2[n]; yt[n] += h * B43 * k3[n]; }
		ok = rhs(v + A4 * h, yt, k4) && ok;
		for (int n = 0; n < 3; n++) { yt[n] = y[n]; yt[n] += h * B50 * k0[n]; yt[n] += h * B51 * k1[n]; yt[n] += h * B52 * k2[n]; yt[n] += h * B53 * k3[n]; yt[n] += h * B54 * k4[n]; }
		return rhs(v + A5 * h, yt, k5) && ok;
	};
	auto plain = [&](double vv, const double* yy, double* kk) { return prof_rhs(P, S, lr0, lr1, vv, yy, kk); };
	bool ok;
	if constexpr (RCP) {
		const double rY1 = rcp_nr(S.Y1);
		bool fast = rc->ok && in_band(S.Y0, BAND_LEAF) && in_band(S.Y1, BAND_LEAF) && in_band(lr0, BAND_LEAF) && in_band(lr1, BAND_LEAF);
		ok = stages([&](double vv, const double* yy, double* kk) { return prof_rhs_rcp(P, S, *rc, rY1, lr0, lr1, vv, yy, kk, fast); });
		// A bound failed somewhere in the step: every lane of the wavefront returns 2 with R as it was, and the caller
		// takes the step again with prof_step<false> (the lanes within bounds get the same results again)
		if (wave_any(!fast)) { R.it--; return 2; }
	} else {
		ok = stages(plain);
	}
	// fifth- and fourth-order sums in the reference's order; its zero-weight terms (0.0 * k1, 0.0 * k4 in the fifth-order
	// sum, 0.0 * k1 in the fourth) are left out: where the step is admissible every k is finite, so each such term is a
	// signed zero, and a sum that starts at +0.0 is never -0.0, so adding a signed zero leaves it unchanged
	double err = 0.0, y5[3];
	for (int n = 0; n < 3; n++) {
		double d5 = 0, d4 = 0;
		d5 += C50 * k0[n]; d4 += C40 * k0[n];
		d5 += C52 * k2[n]; d4 += C42 * k2[n];
		d5 += C53 * k3[n]; d4 += C43 * k3[n];
		d4 += C44 * k4[n];
		d5 += C55 * k5[n]; d4 += C45 * k5[n];
		y5[n] = y[n] + h * d5; err = smax(err, fabs(h * (d5 - d4)) / sc[n]);
	}
	const bool acc = ok && (err <= 1e-8);
	// the step-size factor of both outcomes from one power evaluation (reject: (1e-8 / err)^(1/4), accept:
	// (1e-8 / max(err, 1e-30))^(1/5)), so that lanes taking either branch share it
	const double pw = ppow(acc ? 1e-8 / smax(err, 1e-30) : (ok ? 1e-8 / err : 1.0), acc ? 0.2 : (ok ? 0.25 : 0.0));
	if (!acc) {
		h *= ok ? smax(0.1, 0.9 * pw) : 0.25;
		R.h = h;
		return fabs(h) < 1e-14 * fabs(L) ? -1 : 0;
	}
	R.y0 = y5[0]; R.y1 = y5[1]; R.y2 = y5[2];
	if (last) return 1;
	R.v = v + h;
	R.h = h * smin(5.0, 0.9 * pw);
	return 0;
}
SPB_HD bool prof_integrate(const Phys& P, const Prof& S, double c, double& v1end, double& qend) {
	ProfRun R; prof_begin(S, c, R);
	for (;;) {
		const int r = prof_step(P, S, R);
		if (r < 0) return false;
		if (r > 0) { v1end = R.y1; qend = R.y2; return true; }
	}
}
// The profile integration in fused form (the strong-shock kernel with shock_fused = 1): the same ODE and the same
// adaptive Cash-Karp step as prof_rhs / prof_step, rounded differently. With x_k = g_k (p + pi_k), n_k = v_k - b_k,
// the integration's quotients lr_k = lam_k / Y_k (prof_begin), e_k = Y_k ((g_k - 1) lr_k) and E_0 = (g_0 - 1) lr_0,
// the two phase terms of PHY:302-310 over their common denominator give
//     dp/dv  = (-x0 x1 + (p - P)(e0 x1 + e1 x0)) / (Y0 n0 x1 + Y1 n1 x0),
//     dv1/dv = (E0 (p - P) den - n0 num) / (den x0),
// so a right-hand side takes one correctly rounded division (1 / (den x0)) where PHY:302-310 takes ten quotients; v2
// multiplies by 1 / Y1, the error norm by 1 / sc_k (both once per cell), and the stage inputs, the fifth- and
// fourth-order sums and the products feeding them are FMA chains in the reference's term order. Every operation is a
// correctly rounded IEEE one (fma included), so the oracle's restatement gives the same bits on the CPU.
struct ProfF { double rY1, rs0, rs1, rs2, E0, e0, e1; };
// the per-cell part: 1 / Y1 and the reciprocal error scales 1 / sc_k
SPB_HD void proff_cell(const Prof& S, ProfF& F) {
	const double L = S.vend - S.v0;
	F.rY1 = 1.0 / S.Y1;
	F.rs0 = 1.0 / (fabs(S.p0) + S.j2 * fabs(L)); F.rs1 = 1.0 / S.vk00; F.rs2 = 1.0 / (S.j2 * L * L + 1e-300);
}
// the per-integration part, from the integration's lam_k / Y_k (prof_begin)
SPB_HD void proff_begin(const Phys& P, const Prof& S, const ProfRun& R, ProfF& F) {
	F.E0 = (P.g[0] - 1.0) * R.lr0;
	F.e0 = S.Y0 * F.E0;
	F.e1 = S.Y1 * ((P.g[1] - 1.0) * R.lr1);
}
SPB_HD bool prof_rhs_fused(const Phys& P, const Prof& S, const ProfF& F, double v, const double* y, double* dy) {
	const double p = y[0], v1 = y[1];
	const double v2 = fma(-S.Y0, v1, v) * F.rY1;
	const double Pr = fma(S.j2, S.v0 - v, S.p0), dp = p - Pr;
	const bool adm = eos_ok(P, 0, p, v1) && eos_ok(P, 1, p, v2);
	const double n0 = v1 - P.b[0], n1 = v2 - P.b[1];
	const double x0 = P.g[0] * (p + P.pi[0]), x1 = P.g[1] * (p + P.pi[1]);
	const double den = fma(S.Y0 * n0, x1, (S.Y1 * n1) * x0);
	const double num = fma(dp, fma(F.e0, x1, F.e1 * x0), -(x0 * x1));
	const double D = 1.0 / (den * x0);
	dy[0] = (num * x0) * D;
	dy[1] = fma(F.E0 * dp, den, -(n0 * num)) * D;
	dy[2] = dp;
	return adm && isfin(dy[0]) && isfin(dy[1]);
}
// one pass of the fused integration loop, with prof_step's return codes (0 not done, 1 reached v_end, -1 failed) and
// its step control (the same factors, from the same portable power)
SPB_HD int prof_step_fused(const Phys& P, const Prof& S, const ProfF& F, ProfRun& R) {
	constexpr double A1 = 0.2, A2 = 0.3, A3 = 0.6, A4 = 1.0, A5 = 0.875;
	constexpr double B10 = 0.2;
	constexpr double B20 = 3.0 / 40, B21 = 9.0 / 40;
	constexpr double B30 = 0.3, B31 = -0.9, B32 = 1.2;
	constexpr double B40 = -11.0 / 54, B41 = 2.5, B42 = -70.0 / 27, B43 = 35.0 / 27;
	constexpr double B50 = 1631.0 / 55296, B51 = 175.0 / 512, B52 = 575.0 / 13824, B53 = 44275.0 / 110592, B54 = 253.0 / 4096;
	constexpr double C50 = 37.0 / 378, C52 = 250.0 / 621, C53 = 125.0 / 594, C55 = 512.0 / 1771;
	constexpr double C40 = 2825.0 / 27648, C42 = 18575.0 / 48384, C43 = 13525.0 / 55296, C44 = 277.0 / 14336, C45 = 0.25;
	if (R.it >= 4000) return -1;
	R.it++;
	const double L = S.vend - S.v0;
	const double y[3] = {R.y0, R.y1, R.y2}, v = R.v;
	double h = R.h;
	const bool last = (L < 0 && v + h <= S.vend) || (L > 0 && v + h >= S.vend);
	if (last) h = S.vend - v;
	double k0[3], k1[3], k2[3], k3[3], k4[3], k5[3], yt[3];
	bool ok = prof_rhs_fused(P, S, F, v, y, k0);
	for (int n = 0; n < 3; n++) yt[n] = fma(h * B10, k0[n], y[n]);
	ok = prof_rhs_fused(P, S, F, v + A1 * h, yt, k1) && ok;
	for (int n = 0; n < 3; n++) yt[n] = fma(h * B21, k1[n], fma(h * B20, k0[n], y[n]));
	ok = prof_rhs_fused(P, S, F, v + A2 * h, yt, k2) && ok;
	for (int n = 0; n < 3; n++) yt[n] = fma(h * B32, k2[n], fma(h * B31, k1[n], fma(h * B30, k0[n], y[n])));
	ok = prof_rhs_fused(P, S, F, v + A3 * h, yt, k3) && ok;
	for (int n = 0; n < 3; n++) yt[n] = fma(h * B43, k3[n], fma(h * B42, k2[n], fma(h * B41, k1[n], fma(h * B40, k0[n], y[n]))));
	ok = prof_rhs_fused(P, S, F, v + A4 * h, yt, k4) && ok;
	for (int n = 0; n < 3; n++)
		yt[n] = fma(h * B54, k4[n], fma(h * B53, k3[n], fma(h * B52, k2[n], fma(h * B51, k1[n], fma(h * B50, k0[n], y[n])))));
	ok = prof_rhs_fused(P, S, F, v + A5 * h, yt, k5) && ok;
	const double rs[3] = {F.rs0, F.rs1, F.rs2};
	double err = 0.0, y5[3];
	for (int n = 0; n < 3; n++) {
		const double d5 = fma(C55, k5[n], fma(C53, k3[n], fma(C52, k2[n], C50 * k0[n])));
		const double d4 = fma(C45, k5[n], fma(C44, k4[n], fma(C43, k3[n], fma(C42, k2[n], C40 * k0[n]))));
		y5[n] = fma(h, d5, y[n]); err = smax(err, fabs(h * (d5 - d4)) * rs[n]);
	}
	const bool acc = ok && (err <= 1e-8);
	const double pw = ppow(acc ? 1e-8 / smax(err, 1e-30) : (ok ? 1e-8 / err : 1.0), acc ? 0.2 : (ok ? 0.25 : 0.0));
	if (!acc) {
		h *= ok ? smax(0.1, 0.9 * pw) : 0.25;
		R.h = h;
		return fabs(h) < 1e-14 * fabs(L) ? -1 : 0;
	}
	R.y0 = y5[0]; R.y1 = y5[1]; R.y2 = y5[2];
	if (last) return 1;
	R.v = v + h;
	R.h = h * smin(5.0, 0.9 * pw);
	return 0;
}
// the share's logistic parameter t -> c = 1 / (1 + exp(-t))
SPB_HD double share_c(double t) { return 1.0 / (1.0 + pexp(-t)); }
// phase shares lam and mixture dissipation q of a shock from (p0, vpre, Y) to p1 (PHY:343-383, PHO start bracket),
// in three parts: share_pre computes the phase stiffnesses K, G and settles the cheap cases -- 0: no share (shock_share
// false), 1: the weak-shock closed form (lam, q set; shock_share true) -- or returns 2 when the strong branch is needed;
// share_setup then builds the profile problem (false: no share), and the shooting over the adaptive RK45 profile runs
// as the state machine Shoot (shoot_begin, then shoot_next with each F(t) until the phase is SH_DONE or SH_FAIL)
SPB_HD int share_pre(const Phys& P, const double Y[2], const double vpre[2], double p0, double p1, double K[2], double G[2],
                     double lam[2], double& q) {
	double Kmin = dinf();
	for (int k = 0; k < 2; k++) {
		K[k] = c2_rp(P, k, 1.0 / vpre[k], p0) / vpre[k];   // K and G (PHY:294-299): MG by a difference along the isentrope
		if (is_mg(P, k)) {
			const double dp = 1e-6 * K[k], vv = vpre[k] * (1.0 - dp / K[k]), K2 = c2_rp(P, k, 1.0 / vv, p0 + dp) / vv;
			G[k] = 0.5 * (1.0 + (K2 - K[k]) / dp);
		} else { G[k] = 0.5 * (P.g[k] + 1.0); if (P.b[k] != 0.0) G[k] *= vpre[k] / (vpre[k] - P.b[k]); }
		if (!(K[k] > 0)) return 0; Kmin = smin(Kmin, K[k]);
	}
	const double dp = p1 - p0;
	if (dp < 1e-2 * Kmin) {
		double w[2], sw = 0.0;
		for (int k = 0; k < 2; k++) { w[k] = Y[k] * vpre[k] * G[k] / (K[k] * K[k]); sw += w[k]; }
		for (int k = 0; k < 2; k++) lam[k] = w[k] / sw;
		q = sw * dp * dp * dp / 6.0;
		return 1;
	}
	return 2;
}
// the strong branch's profile problem S, phase 0's Hugoniot volume vH0 and the start t0 of the shooting bracket (the
// weak-shock share's logit, PHO:365-374); false where the reference returns no share before shooting (PHY:359-364)
SPB_HD bool share_setup(const Phys& P, const double Y[2], const double vpre[2], double p0, double p1, const double K[2], const double G[2],
                        Prof& S, double& vH0, double& t0) {
	const double dp = p1 - p0;
	S.p0 = p0; S.Y0 = Y[0]; S.Y1 = Y[1]; S.vk00 = vpre[0];
	vH0 = hug_v(P, 0, p0, p1, vpre[0]);
	const double vH1 = hug_v(P, 1, p0, p1, vpre[1]);
	if (!isfin(vH0) || !isfin(vH1)) return false;
	S.v0 = Y[0] * vpre[0] + Y[1] * vpre[1]; S.vend = Y[0] * vH0 + Y[1] * vH1;
	if (!(S.vend < S.v0)) return false;
	S.j2 = dp / (S.v0 - S.vend);
	const double ww0 = Y[0] * vpre[0] * G[0] / (K[0] * K[0]), ww1 = Y[1] * vpre[1] * G[1] / (K[1] * K[1]);
	const double c0 = smin(1.0 - 1e-12, smax(1e-12, ww0 / (ww0 + ww1)));
	t0 = isfin(c0) ? plog(c0 / (1.0 - c0)) : 0.0;
	return true;
}
// F(t) of the shooting: the relative miss of phase 0's end volume, (v1(v_end; c(t)) - vH0) / vpre0, NaN where the
// integration fails
SPB_HD double share_miss(const ProfRun& R, int r, double vH0, double vpre0) { return r > 0 ? (R.y1 - vH0) / vpre0 : dnan(); }
// Illinois shooting on t = logit(c) (PHY:366-379) with the PHO start bracket t0 -+ 1 expanded by doubling steps
enum { SH_FA = 0, SH_FB, SH_EXPL, SH_EXPR, SH_ILL, SH_DONE, SH_FAIL };
constexpr int SHOCK_REC = 10;   // doubles per strong-shock record of the grid path (spb6_cell.h)
struct Shoot {
	double ta, tb, fa, fb, qa, qb, st, t, qt;
	int phase, side, it;
};
// returns the first t to evaluate F at
SPB_HD double shoot_begin(Shoot& H, double t0) {
	H.ta = smax(-40.0, t0 - 1.0); H.tb = smin(40.0, t0 + 1.0); H.phase = SH_FA; H.side = 0; H.it = 0; H.st = 2.0;
	H.t = H.ta; H.qt = 0.0; H.fa = H.fb = H.qa = H.qb = 0.0;
	return H.ta;
}
// F at the t asked for last is f with dissipation q: the next t to evaluate (phase < SH_DONE), or the end of the
// shooting (SH_DONE: the share's parameter H.t and dissipation H.qt; SH_FAIL: no share)
SPB_HD double shoot_next(Shoot& H, double f, double q) {
	bool right = false;
	switch (H.phase) {
	case SH_FA: H.fa = f; H.qa = q; H.phase = SH_FB; return H.tb;
	case SH_FB:
		H.fb = f; H.qb = q;
		if (!isfin(H.fa) || !isfin(H.fb)) { H.phase = SH_FAIL; return 0.0; }
		H.st = 2.0;
		break;
	case SH_EXPL:
		H.fa = f; H.qa = q;
		if (!isfin(H.fa)) { H.phase = SH_FAIL; return 0.0; }
		H.st *= 2.0;
		break;
	case SH_EXPR:
		H.fb = f; H.qb = q;
		if (!isfin(H.fb)) { H.phase = SH_FAIL; return 0.0; }
		H.st *= 2.0;
		right = true;
		break;
	case SH_ILL: {
		const double ft = f;
		if (!isfin(ft)) { H.phase = SH_FAIL; return 0.0; }
		H.qt = q;
		if (fabs(ft) < 1e-10 || H.tb - H.ta < 1e-10) { H.phase = SH_DONE; return 0.0; }
		// ft < 0: ta = t, fa = ft, and fb halved if the previous point fell on the same side; else tb = t, fb = ft, fa
		// halved likewise (written as selects, which keeps the bracket in registers)
		const bool neg = ft < 0;
		const double fa = neg ? ft : (H.side == 1 ? H.fa * 0.5 : H.fa);
		const double fb = neg ? (H.side == -1 ? H.fb * 0.5 : H.fb) : ft;
		H.ta = neg ? H.t : H.ta; H.tb = neg ? H.tb : H.t; H.fa = fa; H.fb = fb; H.side = neg ? -1 : 1;
		if (++H.it >= 100) { H.phase = SH_DONE; return 0.0; }
		H.t = (H.ta * H.fb - H.tb * H.fa) / (H.fb - H.fa);
		if (!(H.t > H.ta && H.t < H.tb)) H.t = 0.5 * (H.ta + H.tb);
		return H.t;
	}
	default: return 0.0;
	}
	if (!right) {   // expand the bracket to the left while F(ta) >= 0
		if (H.fa >= 0 && H.ta > -40.0) {
			H.tb = H.ta; H.fb = H.fa; H.qb = H.qa; H.ta = smax(-40.0, H.ta - H.st); H.phase = SH_EXPL;
			return H.ta;
		}
		H.st = 2.0;
	}
	if (H.fb <= 0 && H.tb < 40.0) {   // then to the right while F(tb) <= 0
		H.ta = H.tb; H.fa = H.fb; H.qa = H.qb; H.tb = smin(40.0, H.tb + H.st); H.phase = SH_EXPR;
		return H.tb;
	}
	if (H.fa >= 0) { H.t = H.ta; H.qt = H.qa; H.phase = SH_DONE; return 0.0; }
	if (H.fb <= 0) { H.t = H.tb; H.qt = H.qb; H.phase = SH_DONE; return 0.0; }
	H.side = 0; H.it = 0;
	H.t = (H.ta * H.fb - H.tb * H.fa) / (H.fb - H.fa);
	if (!(H.t > H.ta && H.t < H.tb)) H.t = 0.5 * (H.ta + H.tb);
	H.phase = SH_ILL;
	return H.t;
}
// shoot_next as straight-line code with selects, for the strong-shock kernel, where the lanes of a wavefront are in
// different phases: each lane evaluates every case's arithmetic once and keeps the fields of its own case, so a
// wavefront pays for one pass and one regula falsi quotient instead of one per phase present and the state is not
// copied between branches. Field by field the same values as shoot_next for every state and input (the regula falsi
// point is taken from the bracket that the lane's case continues from: the one just expanded, or the Illinois update).
SPB_HD double shoot_next_sel(Shoot& H, double f, double q) {
	const int ph = H.phase;
	if (ph < SH_FA || ph >= SH_DONE) return 0.0;
	const bool isFA = ph == SH_FA, isFB = ph == SH_FB, isEL = ph == SH_EXPL, isER = ph == SH_EXPR, isIL = ph == SH_ILL;
	// F and q go to the bracket end the phase evaluated
	const double fa = (isFA || isEL) ? f : H.fa, qa = (isFA || isEL) ? q : H.qa;
	const double fb = (isFB || isER) ? f : H.fb, qb = (isFB || isER) ? q : H.qb;
	const bool fail = (isFB && !(isfin(fa) && isfin(fb))) || (isEL && !isfin(fa)) || (isER && !isfin(fb)) || (isIL && !isfin(f));
	// bracket expansion (FB, EXPL, EXPR): L1 to the left, else R1 to the right, else an end with F of the right sign
	// (D1, D2), else the Illinois start
	const double st1 = isFB ? 2.0 : H.st * 2.0;
	const bool L1 = !isER && fa >= 0 && H.ta > -40.0;
	const double st2 = (!isER && !L1) ? 2.0 : st1;
	const bool R1 = !L1 && fb <= 0 && H.tb < 40.0;
	const bool D1 = !L1 && !R1 && fa >= 0, D2 = !L1 && !R1 && !D1 && fb <= 0;
	// Illinois update (ILL)
	const bool neg = f < 0, conv = fabs(f) < 1e-10 || H.tb - H.ta < 1e-10;
	const double fai = neg ? f : (H.side == 1 ? H.fa * 0.5 : H.fa), fbi = neg ? (H.side == -1 ? H.fb * 0.5 : H.fb) : f;
	const double tai = neg ? H.t : H.ta, tbi = neg ? H.tb : H.t;
	const int iti = H.it + 1;
	const double A = isIL ? tai : H.ta, B = isIL ? tbi : H.tb, FA = isIL ? fai : fa, FB = isIL ? fbi : fb;
	double tn = (A * FB - B * FA) / (FB - FA);
	if (!(tn > A && tn < B)) tn = 0.5 * (A + B);
	const bool ex = (isFB || isEL || isER) && !fail, il = isIL && !fail, ilc = il && !conv;
	const bool I1 = ex && !L1 && !R1 && !D1 && !D2, ilgo = ilc && iti < 100;
	double ret = isFA ? H.tb : 0.0;
	int nph = isFA ? SH_FB : (fail ? SH_FAIL : ph);
	if (ex) nph = L1 ? SH_EXPL : R1 ? SH_EXPR : (D1 || D2) ? SH_DONE : SH_ILL;
	if (il) nph = ilgo ? SH_ILL : SH_DONE;
	const double nta = (ex && L1) ? smax(-40.0, H.ta - st1) : (ex && R1) ? H.tb : ilc ? tai : H.ta;
	const double ntb = (ex && L1) ? H.ta : (ex && R1) ? smin(40.0, H.tb + st2) : ilc ? tbi : H.tb;
	ret = (ex && L1) ? nta : (ex && R1) ? ntb : (I1 || ilgo) ? tn : ret;
	const double nt = (ex && D1) ? H.ta : (ex && D2) ? H.tb : (I1 || ilgo) ? tn : H.t;
	const double nqt = (ex && D1) ? qa : (ex && D2) ? qb : il ? q : H.qt;
	H.fa = (ex && R1) ? fb : ilc ? fai : fa;
	H.qa = (ex && R1) ? qb : qa;
	H.fb = (ex && L1) ? fa : ilc ? fbi : fb;
	H.qb = (ex && L1) ? qa : qb;
	H.st = ex ? (L1 ? st1 : st2) : H.st;
	H.side = I1 ? 0 : ilc ? (neg ? -1 : 1) : H.side;
	H.it = I1 ? 0 : ilc ? iti : H.it;
	H.ta = nta; H.tb = ntb; H.t = nt; H.qt = nqt; H.phase = nph;
	return ret;
}
// the share of a finished shooting (PHY:380-382): lam = (c, 1 - c), q; false without a share
SPB_HD bool shoot_result(const Shoot& H, double lam[2], double& q) {
	if (H.phase != SH_DONE) return false;
	const double c = share_c(H.t);
	lam[0] = c; lam[1] = 1.0 - c; q = H.qt;
	return q > 0;
}
// the strong branch in one call: shooting with each F(t) integrated at once
SPB_HD bool share_strong(const Phys& P, const double Y[2], const double vpre[2], double p0, double p1, const double K[2],
                         const double G[2], double lam[2], double& q) {
	Prof S; double vH0, t0;
	if (!share_setup(P, Y, vpre, p0, p1, K, G, S, vH0, t0)) return false;
	Shoot H;
	double t = shoot_begin(H, t0);
	while (H.phase < SH_DONE) {
		ProfRun R; prof_begin(S, share_c(t), R);
		int r;
		do r = prof_step(P, S, R); while (r == 0);
		t = shoot_next(H, share_miss(R, r, vH0, vpre[0]), R.y2);
	}
	return shoot_result(H, lam, q);
}
SPB_HD bool shock_share(const Phys& P, const double Y[2], const double vpre[2], double p0, double p1, double lam[2], double& q) {
	double K[2], G[2];
	const int m = share_pre(P, Y, vpre, p0, p1, K, G, lam, q);
	if (m != 2) return m == 1;
	return share_strong(P, Y, vpre, p0, p1, K, G, lam, q);
}
// stencil_state (PHY:385-390) from the 8 conserved values
SPB_HD void stencil_state(const Phys& P, const double* c, double* st) {
	const double a1 = c[IA1], m1 = c[IM1], m2 = c[IM2], r = m1 + m2;
	st[0] = a1; st[1] = 1.0 - a1; st[2] = m1; st[3] = m2; st[4] = dnan();
	st[5] = r > 0 ? c[IMU] / r : 0.0; st[6] = r > 0 ? c[IMV] / r : 0.0;
	if (a1 > 0 && a1 < 1 && m1 > 0 && m2 > 0) {
		const double p = p_mix(P, a1, m1, m2, c[IE1] + c[IE2]);
		if (p_adm(P, 0, a1, m1, p) && p_adm(P, 1, 1.0 - a1, m2, p)) st[4] = p;
	}
}

// ------------------------------------------------------------------ p closure (PHY:420-494, PHO:472-548)
// tau: the stage's implicit step a_ii dt (read by the finite-rate closures only)
struct ClIn { double atr0, atr1, mtr0, mtr1, etr0, etr1, s, rhoeC, pguess, Q0, Q1; double tau = 0.0; };
// the closure's phase contents: volume fractions, masses (the transported ones but where the pTg closure moves mass)
// and energies, and the pressure
struct ClOut { double a0, a1, m0, m1, e0, e1, pbar; };

SPB_HD double E_above(const Phys& P, int k, double atr, double mtr, double etr) {    // PHY:420-423
	if (is_mg(P, k)) {
		if (!(atr > 0 && mtr > 0)) return 0.0;
		double pr, dpr, er; mg_ref(P, k, mtr / atr, pr, dpr, er); return smax(0.0, etr - mtr * er); }
	return smax(0.0, etr - mtr * P.q[k] - (atr - mtr * P.b[k]) * P.pi[k]); }

// segment of the piecewise-affine defect share at delta (PHY:428-439): lam, extra
SPB_HD void defect_seg(double delta, double Qb, double Q0, double Q1, double lE0, double lE1,
                       double& lam0, double& lam1, double& ex0, double& ex1) {
	ex0 = 0.0; ex1 = 0.0;
	if (!(delta > 0.0) || !(Qb > 0.0)) { lam0 = lE0; lam1 = lE1; }
	else if (delta < Qb) { lam0 = Q0 / Qb; lam1 = Q1 / Qb; }
	else { lam0 = lE0; lam1 = lE1; ex0 = Q0 - Qb * lE0; ex1 = Q1 - Qb * lE1; }
}
// ---- the p closure with a Mie-Grueneisen phase along the PHO iteration paths (alpha_mg PHO:552-583, closure_p_mg
// PHO:585-628). An MG phase's alpha(p) is the root of f(a) = m er(m / a) + a (p - pr) / G + p a - A - B p, increasing where
// c^2 > 0 (f' = rho c^2 / G): the guess moves up past the c^2 <= 0 side, the root is bracketed by doubling / halving (a
// halving across c^2 = 0 bisects geometrically for a point of the admissible branch with f <= 0), then safeguarded
// Newton from the guess when it lies in the bracket, a converged Newton step taken before the bracket test.
SPB_HD bool alpha_mg(const Phys& P, int k, double m, double A, double B, double p, double guess, double& al, double& dal) {
	const double G = mg_G(P, k);
	auto F = [&](double a, double& df) {
		double pr, dpr, er; const double r = m / a; mg_ref(P, k, r, pr, dpr, er);
		df = ((G + 1.0) * (p - pr) + r * dpr) / G;
		return m * er + a * (p - pr) / G + p * a - A - B * p; };
	double lo = guess, hi = guess, d; double flo = F(lo, d), fhi = flo;
	if (!isfin(flo)) return false;
	if (!(d > 0)) {
		for (int it = 0; it < 200 && !(d > 0); it++) { guess *= 2.0; flo = F(guess, d); if (!isfin(flo)) return false; }
		if (!(d > 0)) return false;
		lo = hi = guess; fhi = flo;
	}
	if (flo < 0) {
		for (int it = 0; it < 200 && fhi < 0; it++) { lo = hi; flo = fhi; hi *= 2.0; fhi = F(hi, d); if (!isfin(fhi)) return false; }
		if (fhi < 0) return false;
	} else {
		for (int it = 0; it < 200 && flo > 0; it++) {
			hi = lo; fhi = flo; lo *= 0.5; flo = F(lo, d); if (!isfin(flo)) return false;
			if (d <= 0) {
				bool found = false;
				for (int b = 0; b < 200; b++) {
					const double mid = sqrt(lo * hi); double dm; const double fm = F(mid, dm);
					if (!isfin(fm)) return false;
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
	double a = (flo == 0) ? lo : (fhi == 0 ? hi : (guess >= lo && guess <= hi ? guess : sqrt(lo * hi)));
	for (int it = 0; it < 200; it++) {
		double df; const double fa = F(a, df);
		if (fa == 0) break;
		if (fa < 0) lo = a; else hi = a;
		double an = (df > 0) ? a - fa / df : -1.0;
		if (an > 0 && fabs(an - a) <= 4e-16 * a) { a = an; break; }
		if (!(an > lo && an < hi)) an = sqrt(lo * hi);
		if (fabs(an - a) <= 4e-16 * a || hi - lo <= 4e-16 * hi) { a = an; break; }
		a = an;
	}
	double df; F(a, df); if (!(df > 0)) return false;
	al = a; dal = -(a / G + a - B) / df; return true;
}
// Phi(p) = sum_k alpha_k(p) - 1 on the guess's defect segment (p_coef, PHY:443-454), each MG phase's alpha from alpha_mg
// warm-started from the last accepted alphas, a NASG phase's in closed form (PHY:455-456); bracketed from the guess by
// doubling steps, then bisection-safeguarded Newton
SPB_HD bool closure_p_mg(const Phys& P, const ClIn& in, ClOut& o) {
	const double D0 = in.rhoeC - in.etr0 - in.etr1, sm = in.mtr0 + in.mtr1;
	double lE0, lE1;
	{ const double E0 = E_above(P, 0, in.atr0, in.mtr0, in.etr0), E1 = E_above(P, 1, in.atr1, in.mtr1, in.etr1);
	  double sE = 0.0; sE += E0; sE += E1;
	  if (sE > 0) { lE0 = E0 / sE; lE1 = E1 / sE; }
	  else { lE0 = sm > 0 ? in.mtr0 / sm : 0.5; lE1 = sm > 0 ? in.mtr1 / sm : 0.5; } }
	double lam0, lam1, ex0, ex1;
	defect_seg(D0 + in.pguess * in.s, in.Q0 + in.Q1, in.Q0, in.Q1, lE0, lE1, lam0, lam1, ex0, ex1);
	const double mt[2] = {in.mtr0, in.mtr1};
	const double A[2] = {in.etr0 - in.mtr0 * P.q[0] + lam0 * D0 + ex0, in.etr1 - in.mtr1 * P.q[1] + lam1 * D0 + ex1};
	const double B[2] = {in.atr0 + lam0 * in.s, in.atr1 + lam1 * in.s};
	const double N[2] = {A[0] - (B[0] - in.mtr0 * P.b[0]) * P.pi[0], A[1] - (B[1] - in.mtr1 * P.b[1]) * P.pi[1]};
	double ag[2] = {in.atr0 / (in.atr0 + in.atr1), in.atr1 / (in.atr0 + in.atr1)};
	auto Phi = [&](double p, double& dPhi, bool upd) -> double {
		double sum = -1.0, an[2]; dPhi = 0.0;
		for (int k = 0; k < 2; k++) {
			double a, da;
			if (is_mg(P, k)) { if (!alpha_mg(P, k, mt[k], A[k], B[k], p, ag[k], a, da)) return dnan(); }
			else {
				if (!(p + P.pi[k] > 0) || !(N[k] > 0)) return dnan();
				a = ((P.g[k] - 1.0) * (A[k] + B[k] * p) + (p + P.g[k] * P.pi[k]) * mt[k] * P.b[k]) / (P.g[k] * (p + P.pi[k]));
				da = -(P.g[k] - 1.0) * N[k] / (P.g[k] * (p + P.pi[k]) * (p + P.pi[k]));
			}
			an[k] = a; sum += a; dPhi += da;
		}
		if (upd) { ag[0] = an[0]; ag[1] = an[1]; }
		return sum; };
	double p = in.pguess, d;
	const double f = Phi(p, d, true);
	double scale = 1e-6 * (fabs(p) + 1e5);
	double lo, hi, flo, fhi;
	if (!isfin(f) || f > 0) {
		lo = p; flo = f; hi = p;
		for (int it = 0; it < 300; it++) {
			hi = p + scale; fhi = Phi(hi, d, false);
			if (isfin(fhi) && fhi <= 0) break;
			if (isfin(fhi)) { lo = hi; flo = fhi; }
			scale *= 2.0;
		}
		if (!(isfin(fhi) && fhi <= 0)) return false;
	} else {
		hi = p; fhi = f; lo = p;
		for (int it = 0; it < 300; it++) {
			lo = p - scale; flo = Phi(lo, d, false);
			if (isfin(flo) && flo >= 0) break;
			if (!isfin(flo)) { scale *= 0.5; if (scale < 1e-300) return false; continue; }
			hi = lo; fhi = flo; scale *= 2.0;
		}
		if (!(isfin(flo) && flo >= 0)) return false;
	}
	p = (isfin(f) && in.pguess >= lo && in.pguess <= hi) ? in.pguess : 0.5 * (lo + hi);
	for (int it = 0; it < 300; it++) {
		double df; const double fp = Phi(p, df, true);
		if (fp == 0.0) break;
		if (!isfin(fp) || fp > 0) lo = p; else hi = p;
		double pn = (isfin(fp) && df < 0) ? p - fp / df : 0.5 * (lo + hi);
		if (isfin(fp) && df < 0 && fabs(pn - p) <= 2e-16 * (fabs(p) + fabs(pn))) { p = pn; break; }
		if (!(pn > lo && pn < hi)) pn = 0.5 * (lo + hi);
		if (fabs(pn - p) <= 2e-16 * (fabs(p) + fabs(pn)) || hi - lo <= 4e-16 * (fabs(hi) + fabs(lo))) { p = pn; break; }
		p = pn;
	}
	double df; if (!isfin(Phi(p, df, true))) return false;
	if (!p_adm(P, 0, ag[0], in.mtr0, p)) return false;
	o.e0 = eps_amp(P, 0, ag[0], in.mtr0, p);
	if (!p_adm(P, 1, ag[1], in.mtr1, p)) return false;
	o.e1 = eps_amp(P, 1, ag[1], in.mtr1, p);
	o.a0 = ag[0]; o.a1 = ag[1]; o.m0 = in.mtr0; o.m1 = in.mtr1; o.pbar = p;
	return isfin(p);
}
// closure_p along the PHO iteration path: safeguarded Newton from the guess, Phi and Phi' together (PHO:472-548).
// Returns false where the reference's closure returns false; on success fills o (alpha, eps per phase and pbar).
// NODELTA (PHY:234): without the energy defect (no defect share, D0 = 0), for the p-pTg activation's prediction only.
template <bool NODELTA = false>
SPB_HD bool closure_p(const Phys& P, const ClIn& in, ClOut& o) {
	// closure() input checks (PHY:787-790; the activation's prediction calls closure_p itself, without them)
	if (!NODELTA) {
		if (!(in.atr0 > 0) || !(in.mtr0 > 0) || !isfin(in.etr0)) return false;
		if (!(in.atr1 > 0) || !(in.mtr1 > 0) || !isfin(in.etr1)) return false;
		if (!isfin(in.rhoeC) || !isfin(in.s)) return false;
		if (!(in.mtr0 + in.mtr1 > 0)) return false;
		if (mg_any(P)) return closure_p_mg(P, in, o);   // PHY:463, PHO:473
	}
	const double plo = smax(-P.pi[0], -P.pi[1]);
	const double D0 = NODELTA ? 0.0 : in.rhoeC - in.etr0 - in.etr1;
	const double sm = in.mtr0 + in.mtr1;
	double lE0, lE1;
	{ const double E0 = E_above(P, 0, in.atr0, in.mtr0, in.etr0), E1 = E_above(P, 1, in.atr1, in.mtr1, in.etr1);
	  double sE = 0.0; sE += E0; sE += E1;
	  if (sE > 0) { lE0 = E0 / sE; lE1 = E1 / sE; }
	  else { lE0 = sm > 0 ? in.mtr0 / sm : 0.5; lE1 = sm > 0 ? in.mtr1 / sm : 0.5; } }
	const double Qb = in.Q0 + in.Q1;
	const double A00 = in.etr0 - in.mtr0 * P.q[0], A01 = in.etr1 - in.mtr1 * P.q[1];
	const double g0 = P.g[0], g1 = P.g[1], pi0 = P.pi[0], pi1 = P.pi[1], b0 = P.b[0], b1 = P.b[1];
	double p = plo + smax(in.pguess - plo, 1e-3 * (fabs(plo) + 1.0));
	double lo = plo, hi = dinf();
	bool conv = false;
	for (int it = 0; it < 400; it++) {
		double lam0, lam1, ex0, ex1;
		if (NODELTA) { lam0 = lam1 = ex0 = ex1 = 0.0; }
		else defect_seg(D0 + p * in.s, Qb, in.Q0, in.Q1, lE0, lE1, lam0, lam1, ex0, ex1);
		const double A0 = A00 + lam0 * D0 + ex0, B0 = in.atr0 + lam0 * in.s, N0 = A0 - (B0 - in.mtr0 * b0) * pi0;
		const double A1 = A01 + lam1 * D0 + ex1, B1 = in.atr1 + lam1 * in.s, N1 = A1 - (B1 - in.mtr1 * b1) * pi1;
		double f = 0.0, df = 0.0;
		f += ((g0 - 1.0) * (A0 + B0 * p) + (p + g0 * pi0) * in.mtr0 * b0) / (g0 * (p + pi0));
		df += -(g0 - 1.0) * N0 / (g0 * (p + pi0) * (p + pi0));
		f += ((g1 - 1.0) * (A1 + B1 * p) + (p + g1 * pi1) * in.mtr1 * b1) / (g1 * (p + pi1));
		df += -(g1 - 1.0) * N1 / (g1 * (p + pi1) * (p + pi1));
		f = f - 1.0;
		if (f == 0.0) { conv = true; break; }
		if (f > 0) lo = p; else hi = p;
		double pn = df < 0 ? p - f / df : dnan();
		if (fabs(pn - p) <= 2e-16 * (fabs(p) + fabs(plo))) { p = pn; conv = true; break; }
		if (!(pn > lo && pn < hi)) {
			if (!isfin(hi)) pn = plo + 4.0 * (p - plo);
			else if (lo == plo) pn = plo + 0.25 * (hi - plo);
			else pn = (hi - plo > 1e3 * (lo - plo + 1e-300) && lo > plo) ? plo + sqrt((lo - plo) * (hi - plo)) : 0.5 * (lo + hi);
		}
		if (fabs(pn - p) <= 2e-16 * (fabs(p) + fabs(plo)) || (hi < dinf() && hi - lo <= 4e-16 * (fabs(hi) + fabs(lo)))) { p = pn; conv = true; break; }
		p = pn;
	}
	if (!conv) return false;
	// outputs (PHY:487-493)
	double lam0, lam1, ex0, ex1;
	if (NODELTA) { lam0 = lam1 = ex0 = ex1 = 0.0; }
	else defect_seg(D0 + p * in.s, Qb, in.Q0, in.Q1, lE0, lE1, lam0, lam1, ex0, ex1);
	const double A0 = in.etr0 - in.mtr0 * P.q[0] + lam0 * D0 + ex0, B0 = in.atr0 + lam0 * in.s;
	const double A1 = in.etr1 - in.mtr1 * P.q[1] + lam1 * D0 + ex1, B1 = in.atr1 + lam1 * in.s;
	o.a0 = ((g0 - 1.0) * (A0 + B0 * p) + (p + g0 * pi0) * in.mtr0 * b0) / (g0 * (p + pi0));
	if (!(o.a0 > in.mtr0 * b0) || !(p + pi0 > 0)) return false;
	o.e0 = eps_amp(P, 0, o.a0, in.mtr0, p);
	o.a1 = ((g1 - 1.0) * (A1 + B1 * p) + (p + g1 * pi1) * in.mtr1 * b1) / (g1 * (p + pi1));
	if (!(o.a1 > in.mtr1 * b1) || !(p + pi1 > 0)) return false;
	o.e1 = eps_amp(P, 1, o.a1, in.mtr1, p);
	o.m0 = in.mtr0; o.m1 = in.mtr1;
	o.pbar = p;
	return isfin(p);
}
// closure_p's Newton function Phi(p) = sum_k alpha_k(p) - 1 (PHO:472-548) at a fixed p, with the inputs X (transported
// alpha_1, alpha_2, m1, m2, eps_1, eps_2), s and (rho e)^C carrying one directional derivative: returns Phi and its
// derivative along that direction, and dphidp = dPhi/dp at fixed inputs (closure_p's df). The defect segment is the one
// closure_p selects at p, so within it lam and the extra share do not depend on p.
template <class T>
SPB_HD T closure_phi(const Phys& P, const T X[6], const T& s, const T& rhoeC, double Q0, double Q1, double p, double& dphidp) {
	const T D0 = rhoeC - X[4] - X[5];
	const T sm = X[2] + X[3];
	const T E0 = smax(X[4] - X[2] * P.q[0] - (X[0] - X[2] * P.b[0]) * P.pi[0], 0.0);   // E_above (PHY:420-423)
	const T E1 = smax(X[5] - X[3] * P.q[1] - (X[1] - X[3] * P.b[1]) * P.pi[1], 0.0);
	const T sE = E0 + E1;
	T lE0, lE1;
	if (sE.v > 0) { lE0 = E0 / sE; lE1 = E1 / sE; }
	else if (sm.v > 0) { lE0 = X[2] / sm; lE1 = X[3] / sm; }
	else { lE0 = T(0.5); lE1 = T(0.5); }
	const double Qb = Q0 + Q1;
	const T delta = D0 + p * s;
	T lam0, lam1, ex0, ex1;   // defect_seg (PHY:428-439)
	if (!(delta.v > 0.0) || !(Qb > 0.0)) { lam0 = lE0; lam1 = lE1; }
	else if (delta.v < Qb) { lam0 = T(Q0 / Qb); lam1 = T(Q1 / Qb); }
	else { lam0 = lE0; lam1 = lE1; ex0 = Q0 - Qb * lE0; ex1 = Q1 - Qb * lE1; }
	const double g0 = P.g[0], g1 = P.g[1], pi0 = P.pi[0], pi1 = P.pi[1], b0 = P.b[0], b1 = P.b[1];
	const T A0 = (X[4] - X[2] * P.q[0]) + lam0 * D0 + ex0, B0 = X[0] + lam0 * s, N0 = A0 - (B0 - X[2] * b0) * pi0;
	const T A1 = (X[5] - X[3] * P.q[1]) + lam1 * D0 + ex1, B1 = X[1] + lam1 * s, N1 = A1 - (B1 - X[3] * b1) * pi1;
	const T f = ((g0 - 1.0) * (A0 + B0 * p) + (p + g0 * pi0) * X[2] * b0) / (g0 * (p + pi0))
	           + ((g1 - 1.0) * (A1 + B1 * p) + (p + g1 * pi1) * X[3] * b1) / (g1 * (p + pi1));
	dphidp = -(g0 - 1.0) * N0.v / (g0 * (p + pi0) * (p + pi0)) - (g1 - 1.0) * N1.v / (g1 * (p + pi1) * (p + pi1));
	return f - 1.0;
}
// round-off floor of the closure pressure (PC:631-638)
SPB_HD double closure_roundoff(const Phys& P, const ClIn& in, double p) {
	if (mg_any(P)) return 0.0;   // NASG / SG closures only (PC:632)
	const double D0 = in.rhoeC - in.etr0 - in.etr1;
	const double sm = in.mtr0 + in.mtr1;
	const double E0 = E_above(P, 0, in.atr0, in.mtr0, in.etr0), E1 = E_above(P, 1, in.atr1, in.mtr1, in.etr1);
	double sE = 0.0; sE += E0; sE += E1;
	double lE0, lE1;
	if (sE > 0) { lE0 = E0 / sE; lE1 = E1 / sE; }
	else { lE0 = sm > 0 ? in.mtr0 / sm : 0.5; lE1 = sm > 0 ? in.mtr1 / sm : 0.5; }
	double lam0, lam1, ex0, ex1;
	defect_seg(D0 + p * in.s, in.Q0 + in.Q1, in.Q0, in.Q1, lE0, lE1, lam0, lam1, ex0, ex1);
	const double A0 = in.etr0 - in.mtr0 * P.q[0] + lam0 * D0 + ex0, B0 = in.atr0 + lam0 * in.s, N0 = A0 - (B0 - in.mtr0 * P.b[0]) * P.pi[0];
	const double A1 = in.etr1 - in.mtr1 * P.q[1] + lam1 * D0 + ex1, B1 = in.atr1 + lam1 * in.s, N1 = A1 - (B1 - in.mtr1 * P.b[1]) * P.pi[1];
	double dphi = 0.0, sens = 0.0;
	dphi += -(P.g[0] - 1.0) * N0 / (P.g[0] * (p + P.pi[0]) * (p + P.pi[0]));
	sens += fabs(lam0) * (P.g[0] - 1.0) / (P.g[0] * (p + P.pi[0]));
	dphi += -(P.g[1] - 1.0) * N1 / (P.g[1] * (p + P.pi[1]) * (p + P.pi[1]));
	sens += fabs(lam1) * (P.g[1] - 1.0) / (P.g[1] * (p + P.pi[1]));
	const double en = 2.220446049250313e-16 * (fabs(in.rhoeC) + fabs(in.etr0) + fabs(in.etr1) + fabs(in.mtr0 * P.q[0]) + fabs(in.mtr1 * P.q[1]));
	return dphi < 0 ? en * sens / -dphi : 0.0;
}

// ------------------------------------------------------------------ relaxation level 1: the pT closure (PHY:183-230, PHY:660-672)
// closure types (PHY:226): mechanical (p), thermal (pT; finite rate theta) and chemical (pTg; finite rate nu) equilibrium
enum { CL_P = 0, CL_MU = 1, CL_PT = 2, CL_THETA = 3, CL_NU = 4, CL_PTG = 5 };
// Brent's method on a sign-changing bracket (PHY:183-204), operation for operation
template <class F>
SPB_HD bool brent_root(F f, double a, double b, double fa, double fb, double tol, double& root) {
	if (!(fa * fb <= 0)) return false;
	if (fa == 0) { root = a; return true; }
	if (fb == 0) { root = b; return true; }
	double c = a, fc = fa, d = b - a, e = d;
	for (int it = 0; it < 200; it++) {
		if (fb * fc > 0) { c = a; fc = fa; d = b - a; e = d; }
		if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
		const double tol1 = 2.0 * 1e-16 * fabs(b) + 0.5 * tol, xm = 0.5 * (c - b);
		if (fabs(xm) <= tol1 || fb == 0) { root = b; return true; }
		if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
			const double sv = fb / fa;
			double pv, qv;
			if (a == c) { pv = 2 * xm * sv; qv = 1 - sv; }
			else { const double qq = fa / fc, r = fb / fc; pv = sv * (2 * xm * qq * (qq - r) - (b - a) * (r - 1)); qv = (qq - 1) * (r - 1) * (sv - 1); }
			if (pv > 0) qv = -qv;
			pv = fabs(pv);
			if (2 * pv < smin(3 * xm * qv - fabs(tol1 * qv), fabs(e * qv))) { e = d; d = pv / qv; } else { d = xm; e = d; }
		} else { d = xm; e = d; }
		a = b; fa = fb;
		b += (fabs(d) > tol1) ? d : (xm > 0 ? tol1 : -tol1);
		fb = f(b); if (!isfin(fb)) return false;
	}
	root = b; return true;
}
// a root of f above plo: bracket search on y = ln(p - plo) from the guess, then Brent (PHY:205-225); every loop bounded
template <class F>
SPB_HD bool root_above(F f, double plo, double guess, double& root) {
	auto Fy = [&](double y) { return f(plo + pexp(y)); };
	const double y0 = plog(smax(guess - plo, 1e-6 * (fabs(plo) + 1.0) + 1e-12));
	const double f0 = Fy(y0);
	if (f0 == 0.0) { root = plo + pexp(y0); return true; }
	double ya = y0, yb = y0, fa = f0, fb = f0, st = 0.1;
	if (!isfin(f0)) {
		for (int k = 0; k < 200 && !isfin(fb); k++) { yb += 1.0; fb = Fy(yb); }
		if (!isfin(fb)) return false;
		ya = yb; fa = fb;
	}
	if (fa > 0) { for (int k = 0; k < 400 && !(fb < 0); k++) { ya = yb; fa = fb; yb += st; st *= 1.6; fb = Fy(yb); if (!isfin(fb)) return false; } }
	else {
		for (int k = 0; k < 400 && !(fa > 0); k++) {
			yb = ya; fb = fa; ya -= st; st *= 1.6; fa = Fy(ya);
			if (!isfin(fa)) {
				ya = yb; fa = fb;
				for (int q = 0; q < 60; q++) {
					const double ym = ya - st / ldexp(1.0, q + 1), fm = Fy(ym);
					if (isfin(fm)) { if (fm > 0) { ya = ym; fa = fm; break; } }
				}
				break;
			}
		}
	}
	if (fa == 0.0) { root = plo + pexp(ya); return true; }
	if (fb == 0.0) { root = plo + pexp(yb); return true; }
	if (!(fa > 0 && fb < 0)) return false;
	double y; if (!brent_root(Fy, ya, yb, fa, fb, 1e-15, y)) return false;
	root = plo + pexp(y); return true;
}
SPB_HD double v_pT(const Phys& P, int k, double p, double T) { return (P.g[k] - 1.0) * P.cv[k] * T / (p + P.pi[k]) + P.b[k]; }   // PHY:78
SPB_HD double e_pT(const Phys& P, int k, double p, double T) { return P.cv[k] * T * (p + P.g[k] * P.pi[k]) / (p + P.pi[k]) + P.q[k]; }   // PHY:79
// the pT closure (B6.3): mechanical and thermal equilibrium at the transported masses, T(p) from the energy, the root of
// sum m_k v_k(p, T(p)) = 1 above max(-pi_k); T_out: the common temperature. closure()'s input checks (PHY:787-790) first.
SPB_HD bool closure_pT(const Phys& P, const ClIn& in, ClOut& o, double* T_out = nullptr) {
	if (!(in.atr0 > 0) || !(in.mtr0 > 0) || !isfin(in.etr0)) return false;
	if (!(in.atr1 > 0) || !(in.mtr1 > 0) || !isfin(in.etr1)) return false;
	if (!isfin(in.rhoeC) || !isfin(in.s)) return false;
	const double m0 = in.mtr0, m1 = in.mtr1;
	const double Eth = in.rhoeC - m0 * P.q[0] - m1 * P.q[1];
	if (!(Eth > 0)) return false;
	auto Tof = [&](double p) {
		double d = 0;
		d += m0 * P.cv[0] * (p + P.g[0] * P.pi[0]) / (p + P.pi[0]);
		d += m1 * P.cv[1] * (p + P.g[1] * P.pi[1]) / (p + P.pi[1]);
		return Eth / d;
	};
	auto Phi = [&](double p) { const double T = Tof(p); return m0 * v_pT(P, 0, p, T) + m1 * v_pT(P, 1, p, T) - 1.0; };
	double p; if (!root_above(Phi, smax(-P.pi[0], -P.pi[1]), in.pguess, p)) return false;
	const double T = Tof(p);
	o.a0 = m0 * v_pT(P, 0, p, T); o.a1 = m1 * v_pT(P, 1, p, T); o.e0 = m0 * e_pT(P, 0, p, T); o.e1 = m1 * e_pT(P, 1, p, T);
	o.m0 = m0; o.m1 = m1;
	o.pbar = p;
	if (T_out) *T_out = T;
	return T > 0 && isfin(o.pbar);
}

// ------------------------------------------------------------------ relaxation levels 2-3: the saturation line and the pTg closure
// (PHY:111-163, PHY:756-778), with the portable log of the device
// the saturation table (built on the host by the same arithmetic, PHY:125-152): n points of ln p on a uniform grid and
// T_sat there, and the critical point (pcrit, vcrit); and the thermal and chemical relaxation rates theta, nu (PHY:33,
// resolved by level: inf = instantaneous)
// the saturation table and the rates; nu_law: Pelanti's state-dependent transfer rate (c_r, p_c; PHY:34-36) with each
// interior cell's rate of the step in cnu (STEP_NU, PC:162)
struct RelaxDev { const double* lp; const double* T; int n; double pcrit, vcrit, theta, nu, mu; int nu_law; double nu_cr, nu_pc; double* cnu; int form; };
SPB_HD double s_pT(const Phys& P, int k, double p, double T) {                      // PHY:80
	return P.cv[k] * (P.g[k] * plog(T) - (P.g[k] - 1.0) * plog(p + P.pi[k])) + P.qp[k]; }
SPB_HD double g_pT(const Phys& P, int k, double p, double T) {                      // PHY:81
	return P.g[k] * P.cv[k] * T + P.b[k] * p + P.q[k] - T * s_pT(P, k, p, T); }
SPB_HD double sat_f(const Phys& P, double p, double T) { return g_pT(P, 0, p, T) - g_pT(P, 1, p, T); }   // PHY:113
// Newton on T for g_1 = g_2 at fixed p, the step kept within [T/2, 2T] (PHY:114-124)
SPB_HD double sat_newton(const Phys& P, double p, double Tg, bool& ok) {
	double T = Tg; ok = false;
	for (int it = 0; it < 60; it++) {
		const double f = sat_f(P, p, T), d = s_pT(P, 1, p, T) - s_pT(P, 0, p, T);
		if (!(d > 0) || !isfin(f)) return T;
		double Tn = T - f / d; if (Tn < 0.5 * T) Tn = 0.5 * T; if (Tn > 2.0 * T) Tn = 2.0 * T;
		if (fabs(Tn - T) < 1e-14 * T) { ok = true; return Tn; }
		T = Tn;
	}
	ok = fabs(sat_f(P, p, T)) < 1e-8 * fabs(g_pT(P, 0, p, T)) + 1e-8; return T;
}
// T_sat(p) (PHY:154-162): Newton from the table's linear interpolation in ln p; NaN outside (0, pcrit) or on failure
SPB_HD double Tsat(const Phys& P, const RelaxDev& S, double p) {
	if (!(p > 0) || !(p < S.pcrit) || S.n < 1) return dnan();
	const double lp = plog(p); double Tg;
	if (lp <= S.lp[0]) Tg = S.T[0];
	else if (lp >= S.lp[S.n - 1]) Tg = S.T[S.n - 1];
	else {
		const double h = S.lp[1] - S.lp[0];
		const int m = (int)((lp - S.lp[0]) / h), n = m < S.n - 2 ? m : S.n - 2;
		const double w = (lp - S.lp[n]) / h; Tg = (1 - w) * S.T[n] + w * S.T[n + 1];
	}
	bool ok; const double T = sat_newton(P, p, Tg, ok); return ok ? T : dnan();
}
constexpr double YMIN = 1e-10;                                                        // PHY:242
// T at p and vapour mass fraction Y from the mixture's specific energy (PHY:701-704)
SPB_HD double T_from_energy(const Phys& P, double p, double Y, double eC) {
	const double d = (1 - Y) * P.cv[0] * (p + P.g[0] * P.pi[0]) / (p + P.pi[0]) + Y * P.cv[1] * (p + P.g[1] * P.pi[1]) / (p + P.pi[1]);
	return (eC - (1 - Y) * P.q[0] - Y * P.q[1]) / d;
}
// the pTg state at p (PHY:757-767): T = T_sat(p) and Y from the energy, clamped to [YMIN, 1 - YMIN] (then T from the
// energy); at p >= pcrit a single phase by v against vcrit
SPB_HD bool pTg_state(const Phys& P, const RelaxDev& S, double p, double rho, double eC, double& Y, double& T) {
	const double Ts = Tsat(P, S, p);
	if (isfin(Ts)) {
		const double el = e_pT(P, 0, p, Ts), ev = e_pT(P, 1, p, Ts);
		Y = (eC - el) / (ev - el);
		if (Y > YMIN && Y < 1 - YMIN) { T = Ts; return true; }
		Y = smin(smax(Y, YMIN), 1 - YMIN);
	} else {
		if (!(p >= S.pcrit)) return false;
		Y = (1.0 / rho < S.vcrit) ? YMIN : 1 - YMIN;
	}
	T = T_from_energy(P, p, Y, eC); return T > 0;
}
// the pTg closure (B6.6, PHY:768-778): mechanical, thermal and chemical equilibrium, the root of rho sum_k Y_k v_k = 1
// above max(-pi_k); the masses move (vapour fraction Y). closure()'s input checks (PHY:787-790) first.
SPB_HD bool closure_pTg(const Phys& P, const RelaxDev& S, const ClIn& in, ClOut& o) {
	if (!(in.atr0 > 0) || !(in.mtr0 > 0) || !isfin(in.etr0)) return false;
	if (!(in.atr1 > 0) || !(in.mtr1 > 0) || !isfin(in.etr1)) return false;
	if (!isfin(in.rhoeC) || !isfin(in.s)) return false;
	const double rho = in.mtr0 + in.mtr1, eC = in.rhoeC / rho;
	auto Phi = [&](double p) {
		double Y, T; if (!pTg_state(P, S, p, rho, eC, Y, T)) return dnan();
		return rho * ((1 - Y) * v_pT(P, 0, p, T) + Y * v_pT(P, 1, p, T)) - 1.0; };
	double p; if (!root_above(Phi, smax(-P.pi[0], -P.pi[1]), in.pguess, p)) return false;
	double Y, T; if (!pTg_state(P, S, p, rho, eC, Y, T)) return false;
	o.m0 = rho * (1 - Y); o.m1 = rho * Y;
	o.a0 = o.m0 * v_pT(P, 0, p, T); o.a1 = o.m1 * v_pT(P, 1, p, T); o.e0 = o.m0 * e_pT(P, 0, p, T); o.e1 = o.m1 * e_pT(P, 1, p, T);
	o.pbar = p;
	return isfin(o.pbar);
}
// ---- finite relaxation rates, backward-Euler form (RELAX_FORM 0, the reference's default)
// the p closure's coefficients of the defect segment at pseg (p_coef, PHY:443-454): A_k, B_k, N_k
SPB_HD bool p_coef_at(const Phys& P, const ClIn& in, double pseg, double A[2], double B[2], double N[2]) {
	if (!(in.mtr0 + in.mtr1 > 0)) return false;
	const double D0 = in.rhoeC - in.etr0 - in.etr1, sm = in.mtr0 + in.mtr1;
	const double E0 = E_above(P, 0, in.atr0, in.mtr0, in.etr0), E1 = E_above(P, 1, in.atr1, in.mtr1, in.etr1);
	double sE = 0.0; sE += E0; sE += E1;
	double lE0, lE1;
	if (sE > 0) { lE0 = E0 / sE; lE1 = E1 / sE; }
	else { lE0 = sm > 0 ? in.mtr0 / sm : 0.5; lE1 = sm > 0 ? in.mtr1 / sm : 0.5; }
	double lam0, lam1, ex0, ex1;
	defect_seg(D0 + pseg * in.s, in.Q0 + in.Q1, in.Q0, in.Q1, lE0, lE1, lam0, lam1, ex0, ex1);
	A[0] = in.etr0 - in.mtr0 * P.q[0] + lam0 * D0 + ex0; B[0] = in.atr0 + lam0 * in.s; N[0] = A[0] - (B[0] - in.mtr0 * P.b[0]) * P.pi[0];
	A[1] = in.etr1 - in.mtr1 * P.q[1] + lam1 * D0 + ex1; B[1] = in.atr1 + lam1 * in.s; N[1] = A[1] - (B[1] - in.mtr1 * P.b[1]) * P.pi[1];
	return true;
}
// finite theta (B6.4, PHY:679-700): at fixed p the phases' temperatures are linear in the exchanged heat, and
// dQ = tau theta (T2 - T1) / (1 + tau theta S), S = 1/Cp1 + 1/Cp2, shifts the p closure's volume fractions; the
// coefficient segment is the one at the guess, the root of sum alpha_k = 1 by the bracket search above max(-pi_k)
SPB_HD bool closure_theta(const Phys& P, double theta, const ClIn& in, ClOut& o, int form = 0) {
	if (!(in.atr0 > 0) || !(in.mtr0 > 0) || !isfin(in.etr0)) return false;
	if (!(in.atr1 > 0) || !(in.mtr1 > 0) || !isfin(in.etr1)) return false;
	if (!isfin(in.rhoeC) || !isfin(in.s)) return false;
	double A[2], B[2], N[2];
	if (!p_coef_at(P, in, in.pguess, A, B, N)) return false;
	const double tt = __builtin_isinf(theta) ? dinf() : in.tau * theta;
	if (tt > 0 && !__builtin_isinf(tt)) { if (!(N[0] > 0) || !(N[1] > 0)) return false; }
	const double m[2] = {in.mtr0, in.mtr1};
	const double Cp0 = P.g[0] * P.cv[0] * m[0], Cp1 = P.g[1] * P.cv[1] * m[1];
	const double S = 1.0 / Cp0 + 1.0 / Cp1;
	auto eval = [&](double p, double& a0, double& a1) {
		double al[2], T0[2];
		for (int k = 0; k < 2; k++) {
			al[k] = ((P.g[k] - 1.0) * (A[k] + B[k] * p) + (p + P.g[k] * P.pi[k]) * m[k] * P.b[k]) / (P.g[k] * (p + P.pi[k]));   // alpha_p
			T0[k] = T_amp(P, k, al[k], m[k], p);
		}
		double dQ;
		if (form == 1) dQ = (T0[1] - T0[0]) * (-pexpm1(-tt * S)) / S;   // the exponential form (PHY:685)
		else if (__builtin_isinf(tt)) dQ = (T0[1] - T0[0]) / S;
		else dQ = tt * (T0[1] - T0[0]) / (1.0 + tt * (1.0 / Cp0 + 1.0 / Cp1));
		a0 = al[0] + (P.g[0] - 1.0) * dQ / (P.g[0] * (p + P.pi[0]));
		a1 = al[1] - (P.g[1] - 1.0) * dQ / (P.g[1] * (p + P.pi[1]));
	};
	auto Phi = [&](double p) { double a0, a1; eval(p, a0, a1); return a0 + a1 - 1.0; };
	double p; if (!root_above(Phi, smax(-P.pi[0], -P.pi[1]), in.pguess, p)) return false;
	eval(p, o.a0, o.a1);
	if (!(o.a0 > m[0] * P.b[0]) || !(o.a1 > m[1] * P.b[1])) return false;
	o.m0 = m[0]; o.m1 = m[1];
	o.e0 = eps_amp(P, 0, o.a0, m[0], p); o.e1 = eps_amp(P, 1, o.a1, m[1], p);
	o.pbar = p;
	return isfin(o.pbar);
}
// finite nu (B6.5, PHY:716-753): at each p the vapour fraction Y solves rho Y - m2 + tau nu (g2 - g1)(p, T(p, Y)) = 0 on
// its admissible interval (clamped at the ends), T from the mixture energy; the root of rho sum_k Y_k v_k = 1
SPB_HD bool closure_nu(const Phys& P, double nu, const ClIn& in, ClOut& o, int form = 0) {
	if (!(in.atr0 > 0) || !(in.mtr0 > 0) || !isfin(in.etr0)) return false;
	if (!(in.atr1 > 0) || !(in.mtr1 > 0) || !isfin(in.etr1)) return false;
	if (!isfin(in.rhoeC) || !isfin(in.s)) return false;
	const double rho = in.mtr0 + in.mtr1, eC = in.rhoeC / rho, tn = __builtin_isinf(nu) ? dinf() : in.tau * nu;
	auto Yof = [&](double p, double& Y, double& T) -> bool {
		auto G = [&](double y) { const double TT = T_from_energy(P, p, y, eC); if (!(TT > 0)) return dnan(); return g_pT(P, 1, p, TT) - g_pT(P, 0, p, TT); };
		auto R = [&](double y) {
			const double TT = T_from_energy(P, p, y, eC); if (!(TT > 0)) return dnan();
			if (form == 1 || __builtin_isinf(tn)) return g_pT(P, 1, p, TT) - g_pT(P, 0, p, TT);
			return rho * y - in.mtr1 + tn * (g_pT(P, 1, p, TT) - g_pT(P, 0, p, TT)); };
		double ya = YMIN, yb = 1.0 - YMIN;
		const double dq = P.q[1] - P.q[0], yT = dq != 0.0 ? (eC - P.q[0]) / dq : dinf();
		if (dq > 0) yb = smin(yb, yT * (1.0 - 1e-12)); else if (dq < 0) ya = smax(ya, yT * (1.0 + 1e-12));
		if (!(ya < yb)) return false;
		const double Y0 = in.mtr1 / rho;
		if (form == 1 && tn == 0.0) { Y = smin(smax(Y0, ya), yb); T = T_from_energy(P, p, Y, eC); return T > 0; }
		const double Ra = R(ya), Rb = R(yb);
		if (!isfin(Ra) || !isfin(Rb)) return false;
		if (Ra >= 0) Y = ya; else if (Rb <= 0) Y = yb; else if (!brent_root(R, ya, yb, Ra, Rb, 1e-16, Y)) return false;
		if (form == 1 && !__builtin_isinf(tn)) {   // the exponential form (PHY:733-745): along the secant D of G from Y_eq
			const double Yeq = Y, Ys = smin(smax(Y0, ya), yb);
			const double D = Ys != Yeq ? G(Ys) / (Ys - Yeq) : 0.0;
			if (D > 0 && isfin(D)) Y = smin(smax(Yeq + (Y0 - Yeq) * pexp(-(tn * D) / rho), ya), yb);
			else if (fabs(Ys - Yeq) <= 1e-14) Y = Yeq;
			else {
				auto Rbe = [&](double y) { const double g = G(y); return isfin(g) ? rho * y - in.mtr1 + tn * g : dnan(); };
				const double ra = Rbe(ya), rb = Rbe(yb); if (!isfin(ra) || !isfin(rb)) return false;
				if (ra >= 0) Y = ya; else if (rb <= 0) Y = yb; else if (!brent_root(Rbe, ya, yb, ra, rb, 1e-16, Y)) return false;
			}
		}
		T = T_from_energy(P, p, Y, eC); return T > 0;
	};
	auto Phi = [&](double p) { double Y, T; if (!Yof(p, Y, T)) return dnan(); return rho * ((1 - Y) * v_pT(P, 0, p, T) + Y * v_pT(P, 1, p, T)) - 1.0; };
	double p; if (!root_above(Phi, smax(-P.pi[0], -P.pi[1]), in.pguess, p)) return false;
	double Y, T; if (!Yof(p, Y, T)) return false;
	o.m0 = rho * (1 - Y); o.m1 = rho * Y;
	o.a0 = o.m0 * v_pT(P, 0, p, T); o.a1 = o.m1 * v_pT(P, 1, p, T); o.e0 = o.m0 * e_pT(P, 0, p, T); o.e1 = o.m1 * e_pT(P, 1, p, T);
	o.pbar = p;
	return isfin(o.pbar);
}
// the closure of type ty (CL_P, CL_PT, CL_THETA, CL_NU, CL_PTG; PHY:786-810)
// finite mu (B6.2, PHY:575-656), backward-Euler form: for a volume transfer R the phase pressures (p1, p2) solve a 2x2
// linear system (each phase's energy with its own pressure work, the defect share on the guess's segment, the frozen
// impedance weights w_k splitting the transfer's work); R is the root of F(R) = R - tau mu (p1 - p2)(R), bracketed from
// F(0) by steps of the first estimate (an inadmissible trial halves back towards the last good point), then Brent.
// pbar = alpha_1 p1 + alpha_2 p2. Z0, Z1: the cell's frozen impedances.
SPB_HD bool closure_mu(const Phys& P, double mu, const ClIn& in, double Z0, double Z1, ClOut& o, int form = 0) {
	if (!(in.atr0 > 0) || !(in.mtr0 > 0) || !isfin(in.etr0) || !(in.atr1 > 0) || !(in.mtr1 > 0) || !isfin(in.etr1)) return false;
	if (!isfin(in.rhoeC) || !isfin(in.s)) return false;
	const double sa = in.atr0 + in.atr1;
	const double om0 = in.atr0 / sa, om1 = in.atr1 / sa, ak0 = in.atr0 / sa, ak1 = in.atr1 / sa;
	const double D0 = in.rhoeC - in.etr0 - in.etr1, sm = in.mtr0 + in.mtr1;
	double lE0, lE1;
	{ const double E0 = E_above(P, 0, in.atr0, in.mtr0, in.etr0), E1 = E_above(P, 1, in.atr1, in.mtr1, in.etr1);
	  double sE = 0.0; sE += E0; sE += E1;
	  if (sE > 0) { lE0 = E0 / sE; lE1 = E1 / sE; }
	  else { lE0 = sm > 0 ? in.mtr0 / sm : 0.5; lE1 = sm > 0 ? in.mtr1 / sm : 0.5; } }
	double lam0, lam1, ex0, ex1;
	defect_seg(D0 + in.pguess * in.s, in.Q0 + in.Q1, in.Q0, in.Q1, lE0, lE1, lam0, lam1, ex0, ex1);
	const double w1 = Z0 + Z1 > 0 ? Z1 / (Z0 + Z1) : 0.5, w2 = 1.0 - w1;
	const double tm = __builtin_isinf(mu) ? dinf() : in.tau * mu;
	double det = 0.0;   // the last solve's determinant
	auto solve = [&](double R, double& p1, double& p2) -> bool {
		const double a1 = ak0 + R, a2 = ak1 - R;
		const double C1 = dedp_am(P, 0, a1, in.mtr0) + om0 * in.s, C2 = dedp_am(P, 1, a2, in.mtr1) + om1 * in.s;
		const double M11 = C1 + R * w1 - lam0 * in.s * om0, M12 = R * w2 - lam0 * in.s * om1;
		const double M21 = -R * w1 - lam1 * in.s * om0, M22 = C2 - R * w2 - lam1 * in.s * om1;
		const double r1 = in.etr0 - in.mtr0 * P.q[0] - P.g[0] * P.pi[0] * dedp_am(P, 0, a1, in.mtr0) + lam0 * D0 + ex0;
		const double r2 = in.etr1 - in.mtr1 * P.q[1] - P.g[1] * P.pi[1] * dedp_am(P, 1, a2, in.mtr1) + lam1 * D0 + ex1;
		det = M11 * M22 - M12 * M21; if (!(fabs(det) > 0)) return false;
		p1 = (r1 * M22 - M12 * r2) / det; p2 = (M11 * r2 - M21 * r1) / det; return isfin(p1) && isfin(p2);
	};
	double R = 0.0, p1, p2;
	// the exponential form first finds the equilibrium transfer R_eq, dp(R_eq) = 0 (PHY:601-643), from a start r0 inside
	// the covolume bounds, its bracket not crossing the 2x2 system's singularity (det keeps its sign)
	const bool eq = form == 1 || __builtin_isinf(tm);
	if (tm > 0) {
		double det0 = 0.0, r0 = 0.0;
		if (eq) {
			const double Rlo = in.mtr0 * P.b[0] - ak0, Rhi = ak1 - in.mtr1 * P.b[1];
			if (!(Rlo < Rhi)) return false;
			if (!(Rlo < 0.0 && 0.0 < Rhi)) r0 = Rlo >= 0.0 ? Rlo + 1e-9 * (Rhi - Rlo) : Rhi - 1e-9 * (Rhi - Rlo);
		}
		auto Fa = [&](double r, double& f) -> bool {
			double q1, q2; if (!solve(r, q1, q2)) return false;
			const double a1 = ak0 + r, a2 = ak1 - r;
			if (!(a1 > in.mtr0 * P.b[0] && a2 > in.mtr1 * P.b[1])) return false;
			if (eq) { if (r != r0 && !(det * det0 > 0)) return false; }
			else if (!(q1 + P.pi[0] > 0 && q2 + P.pi[1] > 0)) return false;
			f = eq ? q2 - q1 : r - tm * (q1 - q2); return isfin(f); };
		double f0; if (!Fa(r0, f0)) return false;
		det0 = det;
		if (eq && fabs(f0) <= 1e-13 * (fabs(in.pguess) + P.g[0] * fabs(P.pi[0]) + P.g[1] * fabs(P.pi[1]))) f0 = 0.0;
		R = r0;
		if (f0 != 0.0) {
			double hi = -f0;
			if (eq) {   // the acoustic linearisation dp(0) / sum_k rho_k c_k^2 / a_k (half the compressed phase's volume if inadmissible)
				double q0, q1, k = 0.0; solve(r0, q0, q1);
				{ const double a = ak0 + r0, r = in.mtr0 / a; k += r * c2_rp(P, 0, r, q0) / a; }
				{ const double a = ak1 - r0, r = in.mtr1 / a; k += r * c2_rp(P, 1, r, q1) / a; }
				hi = -f0 / k;
				if (!(isfin(hi) && hi * -f0 > 0)) hi = f0 < 0 ? 0.5 * (ak1 - r0 - in.mtr1 * P.b[1]) : -0.5 * (ak0 + r0 - in.mtr0 * P.b[0]);
				hi += r0;
			}
			double lo = r0, flo = f0, fhi = dnan(); bool ok = false, exact = false;
			for (int it = 0; it < 200; it++) {
				if (Fa(hi, fhi)) {
					if (eq && fhi == 0.0) { ok = exact = true; break; }
					if (fhi * flo < 0) { ok = true; break; }
					lo = hi; flo = fhi; hi = lo + (lo - r0 == 0 ? -f0 : (lo - r0));
				} else hi = 0.5 * (lo + hi);
			}
			if (!ok) return false;
			auto F = [&](double r) { double f; return Fa(r, f) ? f : dnan(); };
			if (exact) R = hi;
			else if (!brent_root(F, smin(lo, hi), smax(lo, hi), lo < hi ? flo : fhi, lo < hi ? fhi : flo, 1e-16 * (fabs(lo) + fabs(hi)) + 1e-300, R)) return false;
		}
		if (form == 1 && !__builtin_isinf(tm)) {   // PHY:644-647: R = R_eq (1 - exp(-tm D)) along the secant D = dp(0) / R_eq
			const double D = R != 0.0 && r0 == 0.0 ? -f0 / R : 0.0;
			if (D > 0 && isfin(D)) R = -R * pexpm1(-tm * D);
		}
	}
	if (!solve(R, p1, p2)) return false;
	o.a0 = ak0 + R; o.a1 = ak1 - R; o.m0 = in.mtr0; o.m1 = in.mtr1;
	if (!(o.a0 > o.m0 * P.b[0]) || !(p1 + P.pi[0] > 0)) return false;
	o.e0 = eps_amp(P, 0, o.a0, o.m0, p1);
	if (!(o.a1 > o.m1 * P.b[1]) || !(p2 + P.pi[1] > 0)) return false;
	o.e1 = eps_amp(P, 1, o.a1, o.m1, p2);
	o.pbar = o.a0 * p1 + o.a1 * p2;
	return isfin(o.pbar);
}
// the saturation pressure psat(T), the inverse of T_sat on the table (PHY:162-171): bisection in ln p between the table's
// lowest pressure and the critical one; 0 below the table's T, inf above the critical T
SPB_HD double psat(const Phys& P, const RelaxDev& S, double T) {
	if (S.n < 1) return dnan();
	double a = pexp(S.lp[0]), b = S.pcrit * (1 - 1e-12);
	const double Ta = Tsat(P, S, a), Tb = Tsat(P, S, b);
	if (!(T > Ta)) return 0.0;
	if (!(T < Tb)) return dinf();
	for (int it = 0; it < 80; it++) { const double c = sqrt(a * b); if (Tsat(P, S, c) < T) a = c; else b = c; if (b / a - 1 < 1e-14) break; }
	return sqrt(a * b);
}
// Pelanti 2022 eq. (84) (PHY:172-177): nu = c_r alpha_v^0.6 ((psat(T) - p) / (p_c - psat(T)))^1.76 where p < psat(T)
SPB_HD double nu_law(const Phys& P, const RelaxDev& S, double p, double T, double av) {
	const double ps = psat(P, S, T);
	if (!isfin(ps) || !(ps > p) || !(ps < S.nu_pc) || !(av > 0)) return 0.0;
	return S.nu_cr * ppow(av, 0.6) * ppow((ps - p) / (S.nu_pc - ps), 1.76);
}
// nu: the cell's own transfer rate of the step (>= 0 replaces the context's; the nu law)
SPB_HD bool closure_of(const Phys& P, const RelaxDev* S, int ty, const ClIn& in, ClOut& o, double Z0 = 1.0, double Z1 = 1.0, double nu = -1.0) {
	switch (ty) {
	case CL_MU: return closure_mu(P, S->mu, in, Z0, Z1, o, S->form);
	case CL_PT: return closure_pT(P, in, o);
	case CL_THETA: return closure_theta(P, S->theta, in, o, S->form);
	case CL_NU: return closure_nu(P, nu >= 0 ? nu : S->nu, in, o, S->form);
	case CL_PTG: return closure_pTg(P, *S, in, o);
	default: return closure_p(P, in, o);
	}
}
// the cell's closure type of the step at relaxation level `level` (PHY:814-830), from its content and specific energy:
// beyond p only where eps_I <= alpha_1 <= 1 - eps_I; level 1 thermal; level 2 chemical where the pT equilibrium's T
// exceeds T_sat(p), else thermal; level 3 chemical where a phase's T at the mechanical equilibrium without the energy
// defect exceeds T_sat(p), else mechanical
SPB_HD int closure_type_of(const Phys& P, const RelaxDev* S, int level, double eps_i, double a1, double m1, double m2, double e1, double e2, double rhoe) {
	const int mech = __builtin_isinf(S->mu) ? CL_P : CL_MU;   // PHY:815
	if (level == 0) return mech;
	if (!(a1 >= eps_i && a1 <= 1.0 - eps_i)) return mech;
	ClIn in; in.atr0 = a1; in.atr1 = 1 - a1; in.mtr0 = m1; in.mtr1 = m2; in.etr0 = e1; in.etr1 = e2; in.s = 0; in.rhoeC = rhoe;
	in.Q0 = 0.0; in.Q1 = 0.0; in.pguess = smax(p_mix(P, a1, m1, m2, rhoe), 1.0);
	const int chem = __builtin_isinf(S->nu) ? CL_PTG : CL_NU, therm = __builtin_isinf(S->theta) ? CL_PT : CL_THETA;
	if (level == 3) {
		ClOut o; if (!closure_p<true>(P, in, o)) return mech;
		const double Ts = Tsat(P, *S, o.pbar);
		const double T0 = T_amp(P, 0, o.a0, o.m0, o.pbar), T1 = T_amp(P, 1, o.a1, o.m1, o.pbar);
		return (isfin(Ts) && (T0 > Ts || T1 > Ts)) ? chem : mech;
	}
	if (level == 2) {
		ClOut o; double T;
		if (closure_pT(P, in, o, &T)) { const double Ts = Tsat(P, *S, o.pbar); if (isfin(Ts) && T > Ts) return chem; }
	}
	return therm;
}

}   // namespace spb
