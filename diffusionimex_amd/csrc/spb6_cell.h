// spb6_cell.h -- per-cell device functions of the SPB 6+1 IMEX step, shared by the grid kernels (spb6_kernels.hip) and
// the batched 1D stepper (spb6_batch.hip). Each function computes one cell (or face row) from the structure-of-arrays
// fields; the grid kernels map one thread to one cell, the batched stepper loops a workgroup over a problem's cells.
// Layout: field k of a conserved grid at base + k*NC + c, c = i + j*NXT over the ghosted grid (two ghost layers).
#pragma once
#include "spb6_device.h"
#include "spb6_kernels.h"

namespace spb {

// ------------------------------------------------------------------------------------------------ boundaries (PHY:926-957)
// ghost cell dst from source cell src with the wall rule of side (normal momentum mirrored, or 2 rho u_w - rho u with the
// energy fix)
__device__ __forceinline__ void ghost_copy(const Geo& g, double* u, int dst, int src, int side) {
	for (int k = 0; k < NV; k++) u[k * g.NC + dst] = u[k * g.NC + src];
	if (g.bc[side] != BC_WALL) return;
	const int kn = side < 2 ? IMU : IMV;
	const double rho = u[IM1 * g.NC + dst] + u[IM2 * g.NC + dst];
	if (g.bcu[side] == 0.0) { u[kn * g.NC + dst] = -u[kn * g.NC + dst]; return; }
	const double ms = u[kn * g.NC + src];
	const double mn = 2.0 * rho * g.bcu[side] - ms;
	u[IEN * g.NC + dst] += 0.5 * (mn * mn - ms * ms) / rho; u[kn * g.NC + dst] = mn;
}
// x ghost layer q (0 = next to the interior) of row j, side 0 (left) or 1 (right)
__device__ __forceinline__ void bc_x_at(const Geo& g, double* u, int j, int q, int side) {
	const int nx = g.nx;
	if (side == 0) {
		const int iL = GH - 1 - q, sL = g.bc[0] == BC_PERIODIC ? nx + GH - 1 - q : (g.bc[0] == BC_WALL ? GH + q : GH);
		ghost_copy(g, u, iL + j * g.NXT, sL + j * g.NXT, 0);
	} else {
		const int iR = nx + GH + q, sR = g.bc[1] == BC_PERIODIC ? GH + q : (g.bc[1] == BC_WALL ? nx + GH - 1 - q : nx + GH - 1);
		ghost_copy(g, u, iR + j * g.NXT, sR + j * g.NXT, 1);
	}
}
// bound_cond_array (PC:490-511) in x: both ghost layers q of row j
__device__ __forceinline__ void bcarr_x_at(const Geo& g, double* a, int j, int q) {
	const int nx = g.nx;
	auto src = [&](int side) { return g.bc[side] == BC_PERIODIC ? ((side % 2 == 0) ? nx + GH - 1 - q : GH + q)
		: g.bc[side] == BC_WALL ? ((side % 2 == 0) ? GH + q : nx + GH - 1 - q) : ((side % 2 == 0) ? GH : nx + GH - 1); };
	a[(GH - 1 - q) + j * g.NXT] = a[src(0) + j * g.NXT];
	a[(nx + GH + q) + j * g.NXT] = a[src(1) + j * g.NXT];
}

// ------------------------------------------------------------------------------------------------ explicit stage
// reconstruction scales (PC:356-366) of cell c; false where rho <= 0 (the check of Update, PC:255-258). With vu, vv
// given, also the cell velocities mu / rho, mv / rho (the quotients the reconstruction's velocity minmod takes, PC:428-433)
__device__ __forceinline__ bool scales_at(const Geo& g, const Phys& P, const double* __restrict__ Q, double* __restrict__ vsc,
                                          double* __restrict__ esc, int c, double* __restrict__ vu = nullptr, double* __restrict__ vv = nullptr) {
	const int NC = g.NC;
	const double a1 = Q[IA1 * NC + c], m1 = Q[IM1 * NC + c], m2 = Q[IM2 * NC + c], mu = Q[IMU * NC + c], mv = Q[IMV * NC + c];
	const double e1 = Q[IE1 * NC + c], e2 = Q[IE2 * NC + c];
	const double rho = m1 + m2, mm = phypot(mu, mv);
	double cf;
	vsc[c] = (rho > 0 ? mm / rho : 0.0) + (prim_cf(P, a1, m1, m2, e1, e2, cf) ? cf : 0.0);
	esc[c] = fabs(e1) + fabs(e2) + (rho > 0 ? 0.5 * mm * mm / rho : 0.0);
	if (vu) { vu[c] = mu / rho; vv[c] = mv / rho; }
	return rho > 0;
}
// sources of the reconstruction's inputs at a cell: vol(q, cell) the volume field VOL[q] (alpha1, m1, m2, e1, e2),
// vsc / esc the scales, vel(d, cell, rho) the velocity component d (the quotient (rho u_d) / rho)
struct ReconGlobal {   // the conserved grid, velocities divided out here
	const double* __restrict__ Q; const double* __restrict__ vs; const double* __restrict__ es; int NC;
	__device__ double vol(int q, int cell) const { return Q[(q == 0 ? IA1 : q == 1 ? IM1 : q == 2 ? IM2 : q == 3 ? IE1 : IE2) * NC + cell]; }
	__device__ double vsc(int cell) const { return vs[cell]; }
	__device__ double esc(int cell) const { return es[cell]; }
	__device__ double vel(int d, int cell, double rho) const { return Q[(d == 0 ? IMU : IMV) * NC + cell] / rho; }
};
struct ReconGlobalV : ReconGlobal {   // the same with the velocities scales_at stored
	const double* __restrict__ vu; const double* __restrict__ vv;
	__device__ double vel(int d, int cell, double) const { return d == 0 ? vu[cell] : vv[cell]; }
};
// TVD reconstruction of one cell along one direction (PC:392-445) from the inputs of cells ca, cc, cb in S: face values
// L, R (8 vars) and the limiter factor
template <class Src>
__device__ __forceinline__ double recon_cell_s(const Phys& P, int limiter, const Src& S, int ca, int cc, int cb, double* L, double* R) {
	const int VOL[5] = {IA1, IM1, IM2, IE1, IE2};
	double xa[5], xc[5], xb[5];
	#pragma unroll
	for (int q = 0; q < 5; q++) { xa[q] = S.vol(q, ca); xc[q] = S.vol(q, cc); xb[q] = S.vol(q, cb); }
	const double esum = S.esc(ca) + S.esc(cc) + S.esc(cb);
	double phi = 1.0;
	#pragma unroll
	for (int q = 0; q < 5; q++) {
		double sc;
		if (q == 0) sc = 3.0;
		else if (q <= 2) sc = fabs(xa[q]) + fabs(xc[q]) + fabs(xb[q]);
		else sc = esum;
		const double a = xc[q] - xa[q], b = xb[q] - xc[q], tiny = 1e-12 * sc + 1e-300;
		if (!limiter) break;
		if (fabs(a) <= tiny && fabs(b) <= tiny) continue;
		if (a * b <= 0) { phi = 0.0; break; }
		const double cm = 0.5 * (a + b), m = fabs(a) < fabs(b) ? a : b;
		phi = smin(phi, m / cm);
	}
	if (limiter && phi > 0.0) {   // face admissibility cap (PC:400-412)
		#pragma unroll
		for (int ph = 0; ph < 2; ph++) {
			if (is_mg(P, ph)) continue;   // SG / NASG phases only (PC:394)
			double L1[3], L2[3];
			const double av[3] = {ph == 0 ? xa[0] : 1.0 - xa[0], ph == 0 ? xc[0] : 1.0 - xc[0], ph == 0 ? xb[0] : 1.0 - xb[0]};
			const double mv[3] = {ph == 0 ? xa[1] : xa[2], ph == 0 ? xc[1] : xc[2], ph == 0 ? xb[1] : xb[2]};
			const double ev[3] = {ph == 0 ? xa[3] : xa[4], ph == 0 ? xc[3] : xc[4], ph == 0 ? xb[3] : xb[4]};
			#pragma unroll
			for (int n = 0; n < 3; n++) { L1[n] = av[n] - mv[n] * P.b[ph]; L2[n] = ev[n] - mv[n] * P.q[ph] - L1[n] * P.pi[ph]; }
			{ const double dL = fabs(L1[2] - L1[0]); if (L1[1] > 0.0 && dL > 0.0) phi = smin(phi, 2.0 * L1[1] / dL); }
			{ const double dL = fabs(L2[2] - L2[0]); if (L2[1] > 0.0 && dL > 0.0) phi = smin(phi, 2.0 * L2[1] / dL); }
		}
	}
	#pragma unroll
	for (int q = 0; q < 5; q++) {
		const double sig = phi * 0.5 * (xb[q] - xa[q]);
		L[VOL[q]] = xc[q] - 0.5 * sig;
		R[VOL[q]] = xc[q] + 0.5 * sig;
	}
	const double ra = xa[1] + xa[2], rc = xc[1] + xc[2], rb = xb[1] + xb[2];
	const double tiny = 1e-12 * (S.vsc(ca) + S.vsc(cc) + S.vsc(cb)) + 1e-300;
	double vL[2], vR[2], sm = 1.0;
	#pragma unroll
	for (int d = 0; d < 2; d++) {
		const double va = S.vel(d, ca, ra), vc = S.vel(d, cc, rc), vb = S.vel(d, cb, rb), a = vc - va, b = vb - vc;
		double sl = 0.0;
		if (!limiter) sl = 0.5 * (a + b);
		else if (!(fabs(a) <= tiny && fabs(b) <= tiny)) {
			if (a * b > 0) { sl = fabs(a) < fabs(b) ? a : b; sm = smin(sm, sl / (0.5 * (a + b))); }
			else sm = 0.0;
		}
		vL[d] = vc - 0.5 * sl; vR[d] = vc + 0.5 * sl;
	}
	const double rL = L[IM1] + L[IM2], rR = R[IM1] + R[IM2];
	L[IMU] = rL * vL[0]; L[IMV] = rL * vL[1]; L[IEN] = L[IE1] + L[IE2] + 0.5 * rL * (vL[0] * vL[0] + vL[1] * vL[1]);
	R[IMU] = rR * vR[0]; R[IMV] = rR * vR[1]; R[IEN] = R[IE1] + R[IE2] + 0.5 * rR * (vR[0] * vR[0] + vR[1] * vR[1]);
	return smin(phi, sm);
}
__device__ __forceinline__ double recon_cell(const Geo& g, const Phys& P, int limiter, const double* __restrict__ Q,
                                             const double* __restrict__ vsc, const double* __restrict__ esc,
                                             int ca, int cc, int cb, double* L, double* R) {
	return recon_cell_s(P, limiter, ReconGlobal{Q, vsc, esc, g.NC}, ca, cc, cb, L, R);
}
// The face donors the implicit stage reads (m1, m2, e1, e2 of the left and right face states of a ghosted cell) are
// kept per direction as four arrays of (left, right) pairs, pair c of field q at don + 2 (q NC + c): four 16-byte
// stores per cell and direction, contiguous across a wavefront's lanes (half the streams of one array per value).
typedef double dbl2 __attribute__((ext_vector_type(2)));
template <bool NT = true>
__device__ __forceinline__ void put_donors(double* __restrict__ don, int NC, int c, const double* L, const double* R) {
	const int F[4] = {IM1, IM2, IE1, IE2};
	#pragma unroll
	for (int q = 0; q < 4; q++) {
		const dbl2 v = {L[F[q]], R[F[q]]};
		dbl2* p = (dbl2*)(don + 2 * ((long)q * NC + c));
		if constexpr (NT) __builtin_nontemporal_store(v, p); else *p = v;
	}
}
// x-direction face states of ghosted cell c (all rows; cells with both x neighbours, else first order)
__device__ __forceinline__ void recon_x_at(const Geo& g, const Phys& P, int limiter, int order, const double* __restrict__ Q,
                                           const double* __restrict__ vsc, const double* __restrict__ esc,
                                           double* __restrict__ wxL, double* __restrict__ wxR, double* __restrict__ phx, int c,
                                           double* __restrict__ donx = nullptr) {
	const int NC = g.NC, i = c % g.NXT;
	double L[NV], R[NV];
	if (order >= 2 && i >= 1 && i + 1 < g.NXT) {
		phx[c] = recon_cell(g, P, limiter, Q, vsc, esc, c - 1, c, c + 1, L, R);
		for (int k = 0; k < NV; k++) { wxL[k * NC + c] = L[k]; wxR[k * NC + c] = R[k]; }
	} else {
		phx[c] = 0.0;
		for (int k = 0; k < NV; k++) { const double v = Q[k * NC + c]; wxL[k * NC + c] = v; wxR[k * NC + c] = v; L[k] = v; R[k] = v; }
	}
	if (donx) put_donors<false>(donx, NC, c, L, R);
}
// face states L, R and limiter factor of ghosted cell c along the direction of stride d (1: x, NXT: y): the TVD
// reconstruction when second is set (order 2 and both neighbours on the line), else first order (L = R = Q, factor 0)
// (velocities from the arrays vu, vv that scales_at wrote)
__device__ __forceinline__ double recon_dir(const Geo& g, const Phys& P, int limiter, bool second, const double* __restrict__ Q,
                                            const double* __restrict__ vsc, const double* __restrict__ esc, const double* __restrict__ vu,
                                            const double* __restrict__ vv, int c, int d, double* L, double* R) {
	if (second) {
		ReconGlobalV S; S.Q = Q; S.vs = vsc; S.es = esc; S.NC = g.NC; S.vu = vu; S.vv = vv;
		return recon_cell_s(P, limiter, S, c - d, c, c + d, L, R);
	}
	#pragma unroll
	for (int k = 0; k < NV; k++) { const double v = Q[k * g.NC + c]; L[k] = v; R[k] = v; }
	return 0.0;
}
// a store that does not keep its line in the caches (outputs a kernel does not read again), so that the L2 keeps the
// stencil's input lines for the neighbouring tiles
__device__ __forceinline__ void st_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
// a store that is nontemporal when NT (the grid kernels; the batch kernel reads its values back inside the workgroup)
template <bool NT> __device__ __forceinline__ void st_out(double* p, double v) { if constexpr (NT) st_nt(p, v); else *p = v; }
// Rusanov flux with the local material speed between face states QL and QR (PC:224-246), and the face velocity
// ub = (u_L + u_R) / 2 of the path-conservative alpha1 term (PC:300-304)
__device__ __forceinline__ void rusanov_states(const double* QL, const double* QR, int dirk, double* F, double& ub) {
	const double uL = QL[dirk] / (QL[IM1] + QL[IM2]), uR = QR[dirk] / (QR[IM1] + QR[IM2]);
	const double a = smax(fabs(QL[dirk] / (QL[IM1] + QL[IM2])), fabs(QR[dirk] / (QR[IM1] + QR[IM2])));
	#pragma unroll
	for (int k = 0; k < NV; k++) { const double fL = QL[k] * uL, fR = QR[k] * uR; F[k] = 0.5 * (fL + fR) - 0.5 * a * (QR[k] - QL[k]); }
	ub = 0.5 * (uL + uR);
}
// the same between the stored face states: L state at cL (its right face), R state at cR
__device__ __forceinline__ void rusanov_flux(const double* __restrict__ wR_, int cL, const double* __restrict__ wL_, int cR, int NC, int dirk, double* F) {
	double QL[NV], QR[NV], ub;
	#pragma unroll
	for (int k = 0; k < NV; k++) { QL[k] = wR_[k * NC + cL]; QR[k] = wL_[k * NC + cR]; }
	rusanov_states(QL, QR, dirk, F, ub);
}
__device__ __forceinline__ double face_vel(const double* __restrict__ wR_, int cL, const double* __restrict__ wL_, int cR, int NC, int dirk) {
	const double uL = wR_[dirk * NC + cL] / (wR_[IM1 * NC + cL] + wR_[IM2 * NC + cL]);
	const double uR = wL_[dirk * NC + cR] / (wL_[IM1 * NC + cR] + wL_[IM2 * NC + cR]);
	return 0.5 * (uL + uR);
}
// X* = Qs - adt div F(Qe) with the path-conservative alpha1 term (PC:251-326) of interior row q
__device__ __forceinline__ void update_at(const Geo& g, double rxd, double ryd, const double* __restrict__ Qe, const double* __restrict__ Qs,
                                          const double* __restrict__ wxL, const double* __restrict__ wxR, const double* __restrict__ wyL,
                                          const double* __restrict__ wyR, double* __restrict__ Ust, double* __restrict__ a2s,
                                          double* __restrict__ ubx, double* __restrict__ uby, int q) {
	const int NC = g.NC, i = q % g.nx + GH, j = q / g.nx + GH, c = i + j * g.NXT;
	double fp[NV], fm[NV], o[NV];
	rusanov_flux(wxR, c, wxL, c + 1, NC, IMU, fp);
	rusanov_flux(wxR, c - 1, wxL, c, NC, IMU, fm);
	#pragma unroll
	for (int k = 0; k < NV; k++) o[k] = Qs[k * NC + c] - rxd * (fp[k] - fm[k]);
	const double ube = face_vel(wxR, c, wxL, c + 1, NC, IMU), ubw = face_vel(wxR, c - 1, wxL, c, NC, IMU);
	o[IA1] += rxd * Qe[IA1 * NC + c] * (ube - ubw);
	ubx[c] = ube; if (i == GH) ubx[c - 1] = ubw;
	if (g.twoD) {
		double gp[NV], gm[NV];
		rusanov_flux(wyR, c, wyL, c + g.NXT, NC, IMV, gp);
		rusanov_flux(wyR, c - g.NXT, wyL, c, NC, IMV, gm);
		#pragma unroll
		for (int k = 0; k < NV; k++) o[k] -= ryd * (gp[k] - gm[k]);
		const double ubn = face_vel(wyR, c, wyL, c + g.NXT, NC, IMV), ubs = face_vel(wyR, c - g.NXT, wyL, c, NC, IMV);
		o[IA1] += ryd * Qe[IA1 * NC + c] * (ubn - ubs);
		uby[c] = ubn; if (j == GH) uby[c - g.NXT] = ubs;
	}
	#pragma unroll
	for (int k = 0; k < NV; k++) Ust[k * NC + c] = o[k];
	a2s[c] = 1.0 - o[IA1];
}

// ------------------------------------------------------------------------------------------------ implicit stage preparation
__device__ __forceinline__ int map_P(const Geo& g, int i, int j) {   // M_P (PC:688-697): ghosted (i, j) -> pressure row
	int I, J;
	if (i < GH) I = g.xper ? g.nx + i : GH;
	else if (i > g.nx + GH - 1) I = g.xper ? i - g.nx : g.nx + GH - 1;
	else I = i;
	if (j < GH) J = (g.yper && g.ny > 1) ? g.ny + j : GH;
	else if (j > g.ny + GH - 1) J = (g.yper && g.ny > 1) ? j - g.ny : g.ny + GH - 1;
	else J = j;
	return (I - GH) + g.nx * (J - GH);
}
// donor velocities, frozen sound speed, residual scale K + P_E and previous pressure of ghosted cell c (PC:859-900);
// false where the stage state is inadmissible (PC:862)
__device__ __forceinline__ bool prep_cell_at(const Geo& g, const Phys& P, const double* __restrict__ Xs, const double* __restrict__ a2s,
                                             const double* __restrict__ Pst, double* __restrict__ ust, double* __restrict__ vst,
                                             double* __restrict__ cz, double* __restrict__ Kw, double* __restrict__ Pold, int c) {
	const int NC = g.NC, i = c % g.NXT, j = c / g.NXT;
	const double a1 = Xs[IA1 * NC + c], a2 = a2s[c], m1 = Xs[IM1 * NC + c], m2 = Xs[IM2 * NC + c];
	const double e1 = Xs[IE1 * NC + c], e2 = Xs[IE2 * NC + c];
	const double sa = a1 + a2, rho = m1 + m2;
	const bool ok = sa > 0 && rho > 0;
	ust[c] = Xs[IMU * NC + c] / rho; vst[c] = Xs[IMV * NC + c] / rho;
	double czv = dinf(), K = smax(P.pi[0], P.pi[1]), PE = 0.0;
	if (a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0) {
		const double p1 = p_ame(P, 0, a1, m1, e1), p2 = p_ame(P, 1, a2, m2, e2);
		if (p_adm(P, 0, a1, m1, p1) && p_adm(P, 1, a2, m2, p2)) {
			const double c21 = c2_rp(P, 0, m1 / a1, p1), c22 = c2_rp(P, 1, m2 / a2, p2);
			const double c2 = (m1 * c21 + m2 * c22) / (m1 + m2);
			czv = c2 > 0 ? sqrt(c2) : dinf();
			K = 1.0 / ((a1 / sa) / (m1 / a1 * c21) + (a2 / sa) / (m2 / a2 * c22));
		}
		const double dd = dedp_am(P, 0, a1, m1) + dedp_am(P, 1, a2, m2);
		if (dd > 0) PE = (fabs(e1) + eps_round(P, 0, a1, m1) + fabs(e2) + eps_round(P, 1, a2, m2)) / dd;
	}
	cz[c] = czv; Kw[c] = K + PE;
	Pold[c] = Pst[map_P(g, i, j)];
	return ok;
}
__device__ __forceinline__ double mmod(double a, double b) { return a * b <= 0 ? 0.0 : (fabs(a) < fabs(b) ? a : b); }
__device__ __forceinline__ double psi_face(int order, int limiter, double xm, double xL, double xR, double xp) {   // PC:901-905
	if (order < 2) return 1.0;
	const double D = xR - xL; if (D == 0.0) return 1.0;
	if (!limiter) return smin(1.0, smax(0.0, 1.0 - 0.25 * ((xR - xm) + (xp - xL)) / D));
	return 1.0 - 0.5 * (mmod(xL - xm, D) + mmod(D, xp - xR)) / D;
}
// face coefficients of u_f = ubs - kp (P_R - P_L) and p_f = wL P_L + (1 - wL) P_R - pv (set_face, PC:926-950) of the
// faces east (and north) of ghosted cell c
__device__ __forceinline__ void prep_face_at(const Geo& g, int order, int limiter, double tau, const double* __restrict__ Xs, const double* __restrict__ ust,
                                             const double* __restrict__ vst, const double* __restrict__ cz, const double* __restrict__ Pold,
                                             const double* __restrict__ phx, const double* __restrict__ phy,
                                             double* __restrict__ ubsx, double* __restrict__ kpx, double* __restrict__ wLx, double* __restrict__ pvx,
                                             double* __restrict__ ubsy, double* __restrict__ kpy, double* __restrict__ wLy, double* __restrict__ pvy, int c) {
	const int NC = g.NC, NXT = g.NXT, NYT = g.NYT, i = c % NXT, j = c / NXT;
	auto rhoS = [&](int cc) { return Xs[IM1 * NC + cc] + Xs[IM2 * NC + cc]; };
	auto dil_x = [&](int ii, int jj) -> double {
		const int cc = ii + jj * NXT;
		const double d = ust[cc + 1] - ust[cc];
		if (g.ny == 1 || jj < 1 || jj + 1 >= NYT) return d;
		return 0.25 * (2.0 * d + ((ust[cc + 1 + NXT] - ust[cc + NXT]) + (ust[cc + 1 - NXT] - ust[cc - NXT])))
			+ 0.25 * (g.dx / g.dy) * ((vst[cc + NXT] - vst[cc - NXT]) + (vst[cc + 1 + NXT] - vst[cc + 1 - NXT]));
	};
	auto dil_y = [&](int ii, int jj) -> double {
		const int cc = ii + jj * NXT;
		const double d = vst[cc + NXT] - vst[cc];
		if (ii < 1 || ii + 1 >= NXT) return d;
		return 0.25 * (2.0 * d + ((vst[cc + 1 + NXT] - vst[cc + 1]) + (vst[cc - 1 + NXT] - vst[cc - 1])))
			+ 0.25 * (g.dy / g.dx) * ((ust[cc + 1] - ust[cc - 1]) + (ust[cc + 1 + NXT] - ust[cc - 1 + NXT]));
	};
	for (int dir = 0; dir < (g.twoD ? 2 : 1); dir++) {
		const bool ydir = dir == 1;
		if (!ydir && i >= NXT - 1) continue;
		if (ydir && j >= NYT - 1) continue;
		const int di = ydir ? 0 : 1, dj = ydir ? 1 : 0;
		const int cL = c, cR = c + di + dj * NXT;
		const double hh = ydir ? g.dy : g.dx;
		const double* V = ydir ? vst : ust;
		const double rL = rhoS(cL), rR = rhoS(cR), uL = V[cL], uR = V[cR];
		double ubs = 0.5 * (uL + uR), kp = tau / hh / (0.5 * (rL + rR)), wL = rR / (rL + rR), pv = 0.0;
		const double cLz = cz[cL], cRz = cz[cR];
		if (isfin(cLz) && isfin(cRz)) {
			const double phi = smax(0.0, 1.0 - tau * smin(cLz, cRz) / hh);
			if (phi != 0.0) {
				const int im = i - di, jm = j - dj, ip = i + 2 * di, jp = j + 2 * dj;
				double psiP = 1.0, psiU = 1.0;
				if (im >= 0 && jm >= 0 && ip < NXT && jp < NYT) {
					const double* ph = ydir ? phy : phx;
					const double phis = smin(ph[cL], ph[cR]);
					const int cm = im + jm * NXT, cp = ip + jp * NXT;
					psiP = 1.0 - (1.0 - psi_face(order, limiter, Pold[cm], Pold[cL], Pold[cR], Pold[cp])) * phis;
					psiU = psi_face(order, limiter, V[cm], uL, uR, V[cp]);
				}
				const double ZL = rL * cLz, ZR = rR * cRz;
				kp += phi * psiP / (ZL + ZR); pv = phi * psiU * ZL * ZR / (ZL + ZR) * (ydir ? dil_y(i, j) : dil_x(i, j));
			}
		}
		if (!ydir) { ubsx[c] = ubs; kpx[c] = kp; wLx[c] = wL; pvx[c] = pv; }
		else { ubsy[c] = ubs; kpy[c] = kp; wLy[c] = wL; pvy[c] = pv; }
	}
}
// inputs of the shock budget of interior row q on the +-1 cross stencil of X* (PHY:395-418): true for a compressive
// cell with a pressure jump over its stencil, then Y (mass fractions), vpre (pre-shock phase volumes), the low and high
// stencil pressures and the cell's density are set
__device__ __forceinline__ bool shock_inputs(const Geo& g, const Phys& P, const double* __restrict__ Xs, int q, double Y[2], double vpre[2],
                                             double& plo, double& phi, double& rho) {
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	// the compressibility test first: the divergence takes only the four neighbours' velocities (stencil_state's
	// st[5 + d], the same operations in the same order), so a cell that fails it skips the five mixture pressures
	double div = 0.0;
	for (int d = 0; d < (g.twoD ? 2 : 1); d++) for (int s = -1; s <= 1; s += 2) {
		const int cc = c + (d == 0 ? s : s * g.NXT);
		const double r = Xs[IM1 * NC + cc] + Xs[IM2 * NC + cc];
		const double v = r > 0 ? Xs[(d == 0 ? IMU : IMV) * NC + cc] / r : 0.0;
		if (s == 1) div += v; else div -= v;
	}
	if (!(div < 0.0)) return false;
	auto load = [&](int cc, double* u) { for (int k = 0; k < NV; k++) u[k] = Xs[k * NC + cc]; };
	double u[NV], c0[7], st[7], lo[7] = {0, 0, 0, 0, 0, 0, 0};
	plo = 1e300; phi = -1e300;
	load(c, u); stencil_state(P, u, c0);
	for (int d = 0; d < (g.twoD ? 2 : 1); d++) for (int s = -1; s <= 1; s++) {
		if (d == 1 && s == 0) continue;
		load(c + (d == 0 ? s : s * g.NXT), u); stencil_state(P, u, st);
		if (!isfin(st[4])) continue;
		if (st[4] < plo) { plo = st[4]; for (int n = 0; n < 7; n++) lo[n] = st[n]; }
		if (st[4] > phi) phi = st[4];
	}
	if (!(phi > plo)) return false;
	rho = c0[2] + c0[3];
	if (!(rho > 0)) return false;
	for (int k = 0; k < 2; k++) {
		Y[k] = c0[2 + k] / rho;
		vpre[k] = (lo[k] > 0 && lo[2 + k] > 0) ? lo[k] / lo[2 + k] : (c0[2 + k] > 0 ? c0[k] / c0[2 + k] : 0.0);
	}
	return (Y[0] > 0) && (Y[1] > 0) && (vpre[0] > 0) && (vpre[1] > 0);
}
// shock budget Q_k = lam_k rho q of interior row q (PHY:395-418), in one pass
__device__ __forceinline__ void shock_at(const Geo& g, const Phys& P, const double* __restrict__ Xs, double* __restrict__ Qph, int q) {
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	double Q0 = 0.0, Q1 = 0.0, Y[2], vpre[2], plo, phi, rho, lam[2], qq;
	if (shock_inputs(g, P, Xs, q, Y, vpre, plo, phi, rho) && shock_share(P, Y, vpre, plo, phi, lam, qq)) {
		Q0 = lam[0] * rho * qq; Q1 = lam[1] * rho * qq;
	}
	Qph[c] = Q0; Qph[NC + c] = Q1;
}
// first of the two passes of the shock budget: settles every cell without a strong shock (no share, or the weak-shock
// closed form) and returns true for a cell that needs the strong branch (its Qph is then left to shock_strong_at)
__device__ __forceinline__ bool shock_classify_at(const Geo& g, const Phys& P, const double* __restrict__ Xs, double* __restrict__ Qph, int q) {
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	double Q0 = 0.0, Q1 = 0.0, Y[2], vpre[2], plo, phi, rho, lam[2], qq, K[2], G[2];
	if (shock_inputs(g, P, Xs, q, Y, vpre, plo, phi, rho)) {
		const int m = share_pre(P, Y, vpre, plo, phi, K, G, lam, qq);
		if (m == 2) return true;
		if (m == 1) { Q0 = lam[0] * rho * qq; Q1 = lam[1] * rho * qq; }
	}
	Qph[c] = Q0; Qph[NC + c] = Q1;
	return false;
}
// second pass: the strong branch of a cell shock_classify_at returned true for
__device__ __forceinline__ void shock_strong_at(const Geo& g, const Phys& P, const double* __restrict__ Xs, double* __restrict__ Qph, int q) {
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	double Q0 = 0.0, Q1 = 0.0, Y[2], vpre[2], plo, phi, rho, lam[2], qq, K[2], G[2];
	if (shock_inputs(g, P, Xs, q, Y, vpre, plo, phi, rho) && share_pre(P, Y, vpre, plo, phi, K, G, lam, qq) == 2 &&
	    share_strong(P, Y, vpre, plo, phi, K, G, lam, qq)) {
		Q0 = lam[0] * rho * qq; Q1 = lam[1] * rho * qq;
	}
	Qph[c] = Q0; Qph[NC + c] = Q1;
}
// the grid path's first pass with records: settles every cell without a strong shock, including those whose profile
// problem has no share (share_setup false), and returns true for a strong-shock cell with its record: the profile
// problem S, phase 0's Hugoniot volume, rho and the start of the shooting bracket (SHOCK_REC doubles, in this order)
__device__ __forceinline__ bool shock_classify_rec_at(const Geo& g, const Phys& P, const double* __restrict__ Xs, double* __restrict__ Qph, int q,
                                                      double rec[SHOCK_REC]) {
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	double Q0 = 0.0, Q1 = 0.0, Y[2], vpre[2], plo, phi, rho, lam[2], qq, K[2], G[2];
	if (shock_inputs(g, P, Xs, q, Y, vpre, plo, phi, rho)) {
		const int m = share_pre(P, Y, vpre, plo, phi, K, G, lam, qq);
		if (m == 2) {
			Prof S; double vH0, t0;
			if (share_setup(P, Y, vpre, plo, phi, K, G, S, vH0, t0)) {
				rec[0] = S.Y0; rec[1] = S.Y1; rec[2] = S.vk00; rec[3] = S.p0; rec[4] = S.j2; rec[5] = S.v0; rec[6] = S.vend;
				rec[7] = vH0; rec[8] = rho; rec[9] = t0;
				return true;
			}
		}
		if (m == 1) { Q0 = lam[0] * rho * qq; Q1 = lam[1] * rho * qq; }
	}
	Qph[c] = Q0; Qph[NC + c] = Q1;
	return false;
}
// wavefront maximum (xor butterflies)
__device__ __forceinline__ double wave_max(double v) {
	#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) { const double o = __shfl_xor(v, off, 64); v = smax(v, o); }
	return v;
}
// non-negative double max through its bit pattern (monotone for x >= 0): block maximum, then one atomic per block on one
// of 64 striped slots (dst[0..63]); the host takes the maximum of the 64 slots (256-thread blocks)
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* dst, double v) {
	__shared__ double shm[4];
	v = wave_max(v);
	__syncthreads();
	if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) {
		const double m = smax(smax(shm[0], shm[1]), smax(shm[2], shm[3]));
		if (m > 0.0) atomicMax(&dst[blockIdx.x & 63], (unsigned long long)__double_as_longlong(m));
	}
}
// append the cells of a 256-thread block whose flag is set to list, with one atomic per block, and return each flagged
// thread's slot (-1 for the others); every thread of the block calls it. The list order is not deterministic but each
// cell's result does not depend on it. (One device-scope atomic per wavefront on the one counter, over a 16M-cell grid,
// serialises at the memory side: 3.0 ms against 0.8 ms for the strong-shock classification of a 4096 x 4096 grid.)
__device__ __forceinline__ int block_append(bool flag, int q, int* __restrict__ list, int* __restrict__ count) {
	__shared__ int s_cnt[4], s_base;
	const unsigned long long m = __ballot(flag);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) s_cnt[w] = __popcll(m);
	__syncthreads();
	if (threadIdx.x == 0) {
		const int tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
		s_base = tot ? atomicAdd(count, tot) : 0;
	}
	__syncthreads();
	if (!flag) return -1;
	int slot = s_base + __popcll(m & ((1ull << lane) - 1ull));
	for (int k = 0; k < w; k++) slot += s_cnt[k];
	list[slot] = q;
	return slot;
}

// ------------------------------------------------------------------------------------------------ cell residual (PC:584-705)
struct FaceIn { double wl, pv, ubs, kp, ube, cl[4], cr[4], utl, utr, bw; int wall; };
struct CellIn { double xs[NV], a2s, Q0, Q1, Kw; FaceIn fe, fw, fn, fs; };
template <class T> struct FaceValsT { T uf, pf, wL, wR; };
using FaceVals = FaceValsT<double>;
struct ResOut { double a0, a1, m0, m1, e0, e1, mx, my, E, sig, p; };

__device__ __forceinline__ void load_face(const Dev& D, bool ydir, int c, int idx, FaceIn& f) {
	const Geo& g = D.g; const int NC = g.NC, cn = c + (ydir ? g.NXT : 1);
	const double* wLf = ydir ? D.wLy : D.wLx; const double* pvf = ydir ? D.pvy : D.pvx;
	const double* ubsf = ydir ? D.ubsy : D.ubsx; const double* kpf = ydir ? D.kpy : D.kpx; const double* ubef = ydir ? D.uby : D.ubx;
	const double* don = ydir ? D.dony : D.donx;
	const double* ut = ydir ? D.ust : D.vst;
	f.wl = wLf[c]; f.pv = pvf[c]; f.ubs = ubsf[c]; f.kp = kpf[c]; f.ube = ubef[c];
	#pragma unroll
	for (int q = 0; q < 4; q++) {   // right state of c, left state of cn (put_donors' pairs)
		f.cl[q] = don[2 * ((long)q * NC + c) + 1];
		f.cr[q] = don[2 * ((long)q * NC + cn)];
	}
	f.utl = ut[c]; f.utr = ut[cn];
	// wall faces (wall_x / wall_y, PC:594-595); idx is the face's i (x) or j (y) index
	f.wall = -1;
	if (!ydir) { if (idx == GH - 1 && g.bc[0] == BC_WALL) f.wall = 0; else if (idx == g.nx + GH - 1 && g.bc[1] == BC_WALL) f.wall = 1; }
	else { if (idx == GH - 1 && g.bc[2] == BC_WALL) f.wall = 2; else if (idx == g.ny + GH - 1 && g.bc[3] == BC_WALL) f.wall = 3; }
	f.bw = f.wall >= 0 ? g.bcu[f.wall] : 0.0;
}
__device__ __forceinline__ void load_cell(const Dev& D, int q, CellIn& C) {
	const Geo& g = D.g; const int NC = g.NC, i = q % g.nx + GH, j = q / g.nx + GH, c = i + j * g.NXT;
	#pragma unroll
	for (int k = 0; k < NV; k++) C.xs[k] = D.Ust[k * NC + c];
	C.a2s = D.a2s[c]; C.Q0 = D.Qph[c]; C.Q1 = D.Qph[NC + c]; C.Kw = D.Kw[c];
	load_face(D, false, c, i, C.fe); load_face(D, false, c - 1, i - 1, C.fw);
	if (g.twoD) { load_face(D, true, c, j, C.fn); load_face(D, true, c - g.NXT, j - 1, C.fs); }
}
// The face and cell assembly below is templated on the scalar: double for the residual (the reference's operations,
// bit for bit), Dual<2> for a face's derivatives along its two pressures in the analytic Jacobian.
template <class T>
SPB_HD void face_eval(const FaceIn& f, T PL, T PR, FaceValsT<T>& v) {   // face_x / face_y (PC:597-616)
	v.pf = f.wl * PL + (1.0 - f.wl) * PR - f.pv;
	if (f.wall >= 0 && f.bw == 0.0) { v.uf = T(0.0); v.wL = T(0.0); v.wR = T(0.0); return; }
	v.uf = f.wall >= 0 ? T(f.bw) : f.ubs - f.kp * (PR - PL);
	v.wL = smax(v.uf, 0.0) - smax(f.ube, 0.0); v.wR = smin(v.uf, 0.0) - smin(f.ube, 0.0);   // face_weights (PC:592)
}
template <class T>
SPB_HD T finc(const FaceIn& f, const FaceValsT<T>& v, int k) { return f.cl[k] * v.wL + f.cr[k] * v.wR; }   // inc (PC:617), k in m1 m2 e1 e2
// face_flux (PC:620-625): increment mass, tangential momentum and energy fluxes
template <class T>
SPB_HD void fflux(const FaceIn& f, const FaceValsT<T>& v, double P_REF, T& M, T& Mt, T& E) {
	const T ML = (f.cl[0] + f.cl[1]) * v.wL, MR = (f.cr[0] + f.cr[1]) * v.wR;
	M = ML + MR; Mt = f.utl * ML + f.utr * MR;
	E = finc(f, v, 2) + finc(f, v, 3) + 0.5 * v.uf * v.uf * M + 0.5 * (f.utl * f.utl * ML + f.utr * f.utr * MR) + (P_REF + v.pf) * v.uf;
}
// closure inputs of one cell at the given pressures (cell_residual, PC:640-672): X = transported (alpha_1, alpha_2, m1,
// m2, eps_1, eps_2), the volume change s, the momenta and the total energy. face_at(w, f) fills face w (0 east, 1 west,
// 2 north, 3 south) when it is first needed: the x pair is evaluated and folded in before the y pair is read, with every
// sum in the reference's order (each X[k] takes its x term, then its y term).
template <class T, class FaceAt>
SPB_HD void cell_assemble(double P_REF, double rx, double ry, bool twoD, const double xs[NV], FaceAt face_at, T Pc, T PW, T PE,
                          T PS, T PN, T X[6], T& s, T& mx, T& my, T& E) {
	X[0] = T(xs[IA1]); X[2] = T(xs[IM1]); X[3] = T(xs[IM2]); X[4] = T(xs[IE1]); X[5] = T(xs[IE2]);
	T sx;
	{
		FaceIn fa, fb;
		face_at(0, fa); face_at(1, fb);
		FaceValsT<T> fe, fw;
		face_eval(fa, Pc, PE, fe); face_eval(fb, PW, Pc, fw);
		#pragma unroll
		for (int k = 2; k < 6; k++) X[k] -= rx * (finc(fa, fe, k - 2) - finc(fb, fw, k - 2));
		sx = rx * (fe.uf - fw.uf);
		T Me, Mte, Ee, Mw, Mtw, Ew;
		fflux(fa, fe, P_REF, Me, Mte, Ee); fflux(fb, fw, P_REF, Mw, Mtw, Ew);
		mx = xs[IMU] - rx * (fe.pf - fw.pf) - rx * (fe.uf * Me - fw.uf * Mw);
		my = xs[IMV] - rx * (Mte - Mtw);
		E = xs[IEN] - rx * (Ee - Ew);
	}
	if (twoD) {
		FaceIn fa, fb;
		face_at(2, fa); face_at(3, fb);
		FaceValsT<T> fn, fs;
		face_eval(fa, Pc, PN, fn); face_eval(fb, PS, Pc, fs);
		#pragma unroll
		for (int k = 2; k < 6; k++) X[k] -= ry * (finc(fa, fn, k - 2) - finc(fb, fs, k - 2));
		s = sx + ry * (fn.uf - fs.uf);
		T Mn, Mtn, En, Ms, Mts, Es;
		fflux(fa, fn, P_REF, Mn, Mtn, En); fflux(fb, fs, P_REF, Ms, Mts, Es);
		my -= ry * (fn.pf - fs.pf) + ry * (fn.uf * Mn - fs.uf * Ms);
		mx -= ry * (Mtn - Mts);
		E -= ry * (En - Es);
	}
	else s = sx + T(0.0);
	X[1] = 1.0 - X[0];
	X[0] *= 1.0 - s; X[1] *= 1.0 - s;
}
// G = (pbar_cl - P_REF) - Pc of one cell (cell_residual, PC:640-679), faces from face_at as in cell_assemble; out (if
// given) receives the closure's state and, with want_sig, its round-off bound sigma
template <class FaceAt>
SPB_HD bool cell_res(const Phys& P, double P_REF, double rx, double ry, bool twoD, const double xs[NV], double Q0, double Q1,
                     FaceAt face_at, double Pc, double PW, double PE, double PS, double PN, double& G, ResOut* out, bool want_sig = true) {
	double X[6], s, mx, my, E;
	cell_assemble(P_REF, rx, ry, twoD, xs, face_at, Pc, PW, PE, PS, PN, X, s, mx, my, E);
	const double rho = X[2] + X[3];
	if (!(rho > 0)) return false;
	ClIn in; in.atr0 = X[0]; in.atr1 = X[1]; in.mtr0 = X[2]; in.mtr1 = X[3]; in.etr0 = X[4]; in.etr1 = X[5];
	in.s = s; in.rhoeC = E - 0.5 * (mx * mx + my * my) / rho; in.pguess = P_REF + Pc; in.Q0 = Q0; in.Q1 = Q1;
	ClOut o; if (!closure_p(P, in, o)) return false;
	G = (o.pbar - P_REF) - Pc;
	if (out) { out->a0 = o.a0; out->a1 = o.a1; out->m0 = X[2]; out->m1 = X[3]; out->e0 = o.e0; out->e1 = o.e1; out->mx = mx; out->my = my; out->E = E; out->sig = want_sig ? closure_roundoff(P, in, o.pbar) : 0.0; out->p = o.pbar; }
	return isfin(G);
}
// the same on a cell whose data is already loaded (the finite-difference Jacobian evaluates one cell five times)
SPB_HD bool cell_res(const Phys& P, double P_REF, double rx, double ry, bool twoD, const CellIn& C,
                     double Pc, double PW, double PE, double PS, double PN, double& G, ResOut* out) {
	auto face_at = [&](int w, FaceIn& f) { f = w == 0 ? C.fe : w == 1 ? C.fw : w == 2 ? C.fn : C.fs; };
	return cell_res(P, P_REF, rx, ry, twoD, C.xs, C.Q0, C.Q1, face_at, Pc, PW, PE, PS, PN, G, out);
}
// on-demand faces of interior row q (w = 0 east, 1 west, 2 north, 3 south) and its cell values; the scheduling
// barrier keeps each face's loads next to their use instead of all ~64 face doubles live from the top
struct CellFaces {
	const Dev& D; int q, c, i, j;
	__device__ CellFaces(const Dev& D_, int q_) : D(D_), q(q_) { i = q % D.g.nx + GH; j = q / D.g.nx + GH; c = i + j * D.g.NXT; }
	__device__ void operator()(int w, FaceIn& f) const {
		__builtin_amdgcn_sched_barrier(0);
		if (w == 0) load_face(D, false, c, i, f);
		else if (w == 1) load_face(D, false, c - 1, i - 1, f);
		else if (w == 2) load_face(D, true, c, j, f);
		else load_face(D, true, c - D.g.NXT, j - 1, f);
	}
	__device__ void cell(double xs[NV], double& Q0, double& Q1, double& Kw) const {
		#pragma unroll
		for (int k = 0; k < NV; k++) xs[k] = D.Ust[k * D.g.NC + c];
		Q0 = D.Qph[c]; Q1 = D.Qph[D.g.NC + c]; Kw = D.Kw[c];
	}
};
struct Nb { int W, E, S, N; };
// M_P of the four neighbours of interior row q at grid position (ii, jj)
__device__ __forceinline__ Nb neighbours_at(const Geo& g, int q, int ii, int jj) {
	Nb b;
	b.W = ii > 0 ? q - 1 : (g.xper ? q + g.nx - 1 : q);
	b.E = ii < g.nx - 1 ? q + 1 : (g.xper ? q - (g.nx - 1) : q);
	const bool yp = g.yper && g.ny > 1;
	b.S = jj > 0 ? q - g.nx : (yp ? q + (g.ny - 1) * g.nx : q);
	b.N = jj < g.ny - 1 ? q + g.nx : (yp ? q - (g.ny - 1) * g.nx : q);
	return b;
}
__device__ __forceinline__ Nb neighbours(const Geo& g, int q) { const int jj = q / g.nx; return neighbours_at(g, q, q - jj * g.nx, jj); }
// residual of interior row q at P + lam dP (dP may be null): writes G (NaN where the closure fails) and the trial pressure;
// nv = |G| / (|P_REF + P| + K + 1) (PC:770); false where the closure fails
template <bool NT = false>
__device__ __forceinline__ bool residual_at(const Dev& D, const double* __restrict__ P, const double* __restrict__ dP, double lam,
                                            double* __restrict__ Pout, double* __restrict__ G, int q, double& nv) {
	const Nb b = neighbours(D.g, q);
	auto pv = [&](int r) { return dP ? P[r] + lam * dP[r] : P[r]; };
	const double Pc = pv(q), PW = pv(b.W), PE = pv(b.E);
	const double PS = D.g.twoD ? pv(b.S) : 0.0, PN = D.g.twoD ? pv(b.N) : 0.0;
	const CellFaces F(D, q);
	double xs[NV], Q0, Q1, Kw; F.cell(xs, Q0, Q1, Kw);
	double g;
	bool ok = true;
	nv = 0.0;
	if (!cell_res(D.ph, D.P_REF, D.rx, D.ry, D.g.twoD, xs, Q0, Q1, F, Pc, PW, PE, PS, PN, g, nullptr)) { ok = false; g = dnan(); }
	else nv = fabs(g) / (fabs(D.P_REF + Pc) + Kw + 1.0);
	st_out<NT>(G + q, g);
	if (Pout) st_out<NT>(Pout + q, Pc);
	return ok;
}
// P scale s_q = (|P| + K + 1) / (1 + K D) of the finite-difference steps (PC:783-789)
__device__ __forceinline__ void fdscale_at(const Dev& D, const double* __restrict__ P, double* __restrict__ sc, int q) {
	const Geo& g = D.g; const int I = q % g.nx + GH, J = q / g.nx + GH, c = I + J * g.NXT;
	const double K = D.Kw[c];
	double Dd = D.rx * (D.kpx[c] + D.kpx[c - 1]);
	if (g.ny > 1) Dd += D.ry * (D.kpy[c] + D.kpy[c - g.NXT]);
	sc[q] = (fabs(D.P_REF + P[q]) + K + 1.0) / (1.0 + K * Dd);
}
// finite-difference Jacobian row: self, W, E (, S, N) perturbed residuals (PC:790-804); J is 5 arrays of n (3 used in
// 1D); rhs = -G; false where a perturbed closure fails
__device__ __forceinline__ bool jacobian_at(const Dev& D, const double* __restrict__ P, const double* __restrict__ G, const double* __restrict__ sc,
                                            double* __restrict__ J, double* __restrict__ rhs, int q) {
	const int n = D.g.n; const bool twoD = D.g.twoD;
	const Nb b = neighbours(D.g, q);
	CellIn C; load_cell(D, q, C);
	const double Pc = P[q], PW = P[b.W], PE = P[b.E], PS = twoD ? P[b.S] : 0.0, PN = twoD ? P[b.N] : 0.0, Gq = G[q];
	rhs[q] = -Gq;
	const double h = 1e-7 * sc[q];
	double g1; bool ok = cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, C, Pc + h, PW, PE, PS, PN, g1, nullptr);
	double v0 = 0, vW = 0, vE = 0, vS = 0, vN = 0;
	if (ok) {
		v0 = (g1 - Gq) / h;
		double hn = 1e-7 * sc[b.W], g2;
		ok = cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, C, Pc, PW + hn, PE, PS, PN, g2, nullptr); if (ok) vW = (g2 - Gq) / hn;
		if (ok) { hn = 1e-7 * sc[b.E]; ok = cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, C, Pc, PW, PE + hn, PS, PN, g2, nullptr); if (ok) vE = (g2 - Gq) / hn; }
		if (ok && twoD) { hn = 1e-7 * sc[b.S]; ok = cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, C, Pc, PW, PE, PS + hn, PN, g2, nullptr); if (ok) vS = (g2 - Gq) / hn; }
		if (ok && twoD) { hn = 1e-7 * sc[b.N]; ok = cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, C, Pc, PW, PE, PS, PN + hn, g2, nullptr); if (ok) vN = (g2 - Gq) / hn; }
	}
	J[q] = v0; J[n + q] = vW; J[2 * n + q] = vE;
	if (twoD) { J[3 * n + q] = vS; J[4 * n + q] = vN; }
	return ok;
}
// One face's share of the closure inputs of a cell (the terms of cell_assemble), with the derivatives along the face's
// two pressures (PL, PR) scattered into the cell's directions iL, iR (0 self, 1 W, 2 E, 3 S, 4 N). c = -rx (-ry) for the
// cell's east (north) face and +rx (+ry) for its west (south) face; X[2..5], s, the normal and tangential momenta and E
// accumulate.
SPB_HD void jac_face(const FaceIn& f, double PL, double PR, int iL, int iR, double c, double P_REF,
                     Dual<5> X[6], Dual<5>& s, Dual<5>& mn, Dual<5>& mt, Dual<5>& E) {
	using T2 = Dual<2>;
	FaceValsT<T2> v;
	face_eval(f, T2::seed(PL, 0), T2::seed(PR, 1), v);
	T2 M, Mt, Ef;
	fflux(f, v, P_REF, M, Mt, Ef);
	auto add = [&](Dual<5>& acc, const T2& q, double w) { acc.v += w * q.v; acc.d[iL] += w * q.d[0]; acc.d[iR] += w * q.d[1]; };
	#pragma unroll
	for (int k = 0; k < 4; k++) add(X[k + 2], finc(f, v, k), c);
	add(s, v.uf, -c);
	add(mn, v.pf + v.uf * M, c);
	add(mt, Mt, c);
	add(E, Ef, c);
}
// analytic Jacobian row q (the fast forms): the residual's closure is already solved at P, so each entry follows from
// the implicit function theorem at that root, dG/dP_j = -(dPhi/dP_j) / (dPhi/dp) - [j = self]. The closure inputs and
// their derivatives along the five pressures (self, W, E, S, N) are accumulated face by face (each face loaded when it is
// used, its derivatives along its own two pressures in Dual<2>), then the closure function is evaluated once in Dual<5>;
// no closure iterations. Same entries, layout and rhs = -G as jacobian_at; a kinked flux (upwind switch at u_f = 0)
// takes the forward difference's side. False where dPhi/dp is not negative or an entry is not finite.
template <class FaceAt>
SPB_HD bool jacobian_analytic_row(const Phys& Ph, double P_REF, double rx, double ry, bool twoD, const double xs[NV], double Q0, double Q1,
                                  const double Pv[5], double Gq, FaceAt face_at, double J[5]) {
	Dual<5> X[6], s(0.0), mx(xs[IMU]), my(xs[IMV]), E(xs[IEN]);
	X[2] = Dual<5>(xs[IM1]); X[3] = Dual<5>(xs[IM2]); X[4] = Dual<5>(xs[IE1]); X[5] = Dual<5>(xs[IE2]);
	FaceIn f;
	face_at(0, f); jac_face(f, Pv[0], Pv[2], 0, 2, -rx, P_REF, X, s, mx, my, E);   // east: (Pc, PE)
	face_at(1, f); jac_face(f, Pv[1], Pv[0], 1, 0, rx, P_REF, X, s, mx, my, E);    // west: (PW, Pc)
	if (twoD) {
		face_at(2, f); jac_face(f, Pv[0], Pv[4], 0, 4, -ry, P_REF, X, s, my, mx, E);   // north: (Pc, PN)
		face_at(3, f); jac_face(f, Pv[3], Pv[0], 3, 0, ry, P_REF, X, s, my, mx, E);    // south: (PS, Pc)
	}
	X[0] = xs[IA1] * (1.0 - s); X[1] = (1.0 - xs[IA1]) * (1.0 - s);
	const Dual<5> rhoeC = E - 0.5 * (mx * mx + my * my) / (X[2] + X[3]);
	const double p = (Gq + Pv[0]) + P_REF;   // the closure's root at P: G = (pbar - P_REF) - Pc
	double dphidp;
	const Dual<5> phi = closure_phi(Ph, X, s, rhoeC, Q0, Q1, p, dphidp);
	bool ok = dphidp < 0.0;
	#pragma unroll
	for (int j = 0; j < 5; j++) { J[j] = -phi.d[j] / dphidp - (j == 0 ? 1.0 : 0.0); ok = ok && isfin(J[j]); }
	return ok;
}
template <bool NT = false>
__device__ __forceinline__ bool jacobian_analytic_at(const Dev& D, const double* __restrict__ P, const double* __restrict__ G,
                                                     double* __restrict__ J, double* __restrict__ rhs, int q) {
	const Geo& g = D.g; const int n = g.n;
	const bool twoD = g.twoD;
	const Nb b = neighbours(g, q);
	const double Pv[5] = {P[q], P[b.W], P[b.E], twoD ? P[b.S] : 0.0, twoD ? P[b.N] : 0.0};
	const CellFaces F(D, q);
	double xs[NV], Q0, Q1, Kw; F.cell(xs, Q0, Q1, Kw);
	const double Gq = G[q];
	st_out<NT>(rhs + q, -Gq);
	double Jr[5];
	const bool ok = jacobian_analytic_row(D.ph, D.P_REF, D.rx, D.ry, twoD, xs, Q0, Q1, Pv, Gq, F, Jr);
	st_out<NT>(J + q, Jr[0]); st_out<NT>(J + n + q, Jr[1]); st_out<NT>(J + 2 * n + q, Jr[2]);
	if (twoD) { st_out<NT>(J + 3 * n + q, Jr[3]); st_out<NT>(J + 4 * n + q, Jr[4]); }
	return ok;
}
// round-off floor acceptance test of row q (PC:834-842): false where the residual exceeds tol N_i + sigma_i
__device__ __forceinline__ bool floor_ok_at(const Dev& D, const double* __restrict__ P, double tol, int q) {
	const Nb b = neighbours(D.g, q);
	const CellFaces F(D, q);
	double xs[NV], Q0, Q1, Kw; F.cell(xs, Q0, Q1, Kw);
	const bool twoD = D.g.twoD;
	double g; ResOut o;
	if (!cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, xs, Q0, Q1, F, P[q], P[b.W], P[b.E], twoD ? P[b.S] : 0.0, twoD ? P[b.N] : 0.0, g, &o)) return false;
	return !(fabs(g) > tol * (fabs(D.P_REF + P[q]) + Kw + 1.0) + o.sig);
}
// final stage state of row q at the converged pressure (PC:975-987) and the stage derivative Hs = (U_s - Qs) / adt
// (PC:1276-1284); false where the closure fails. T0, T1: the phase temperatures of the closure (PHY:491) where the phase
// is present and T finite -- the reference's T_max statistics, PC:986-989 -- else 0
template <bool NT = false>
__device__ __forceinline__ bool final_at(const Dev& D, const double* __restrict__ P, const double* __restrict__ Qs, double adt, double* __restrict__ H, int q,
                                         double* T0 = nullptr, double* T1 = nullptr) {
	const Geo& g = D.g; const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	const Nb b = neighbours(g, q);
	const CellFaces F(D, q);
	double xs[NV], Q0, Q1, Kw; F.cell(xs, Q0, Q1, Kw);
	const bool twoD = g.twoD;
	double gg; ResOut o;
	if (!cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, xs, Q0, Q1, F, P[q], P[b.W], P[b.E], twoD ? P[b.S] : 0.0, twoD ? P[b.N] : 0.0, gg, &o, false))
		return false;
	const double u[NV] = {o.a0, o.m0, o.m1, o.mx, o.my, o.e0, o.e1, o.E};
	#pragma unroll
	for (int k = 0; k < NV; k++) { st_out<NT>(D.Ust + k * NC + c, u[k]); st_out<NT>(H + k * NC + c, (u[k] - Qs[k * NC + c]) / adt); }
	if (T0) {
		const double t0 = T_amp(D.ph, 0, o.a0, o.m0, o.p), t1 = T_amp(D.ph, 1, o.a1, o.m1, o.p);
		*T0 = (o.a0 > 1e-9 && isfin(t0)) ? smax(t0, 0.0) : 0.0;
		*T1 = (o.a1 > 1e-9 && isfin(t1)) ? smax(t1, 0.0) : 0.0;
	}
	return true;
}
// ------------------------------------------------------------------------------------------------ IMEX variants
// The reference's sibling programs change the implicit stage's cell assembly (cell_residual): TV (SPB6_TV_IMEX.cpp:
// 641-646) transports alpha_1 and alpha_2 with the increments too and takes s = 1 - alpha_1 - alpha_2; the TV-type
// variants (SPB6_TVpure / TVbb / BB_IMEX.cpp:655-675) move the phase energies by the upwind flux eps_k u_f from the
// explicit step's eTV_k and the total energy by the enthalpy flux (sum eps_k + p_f) u_f from its ETV, mass and
// momentum staying with the explicit step; TVinc / TVsplit (SPB6_TVinc_IMEX.cpp:655-687) add the mass increments, the
// momentum terms and the increments' kinetic energy. A variant face also carries its alpha_1 donors (TV).
struct FaceInV : FaceIn { double cla, cra; };
// type: the cell's closure type (CL_P; CL_PT, CL_THETA, CL_NU, CL_PTG at relaxation levels 1-3); rel: the saturation
// table and the finite rates; tau: the stage's implicit step a_ii dt
struct VarCell { double a2s, eT1, eT2, ET; int type; const RelaxDev* rel; double tau; double Z0 = 1.0, Z1 = 1.0, nu = -1.0; };
SPB_HD double var_inc(const FaceInV& f, const FaceVals& v, int k) {   // inc (PC:617) over the 6-vector (alpha1, 1 - alpha1, m1, m2, e1, e2)
	if (k == 0) return f.cla * v.wL + f.cra * v.wR;
	if (k == 1) return (1.0 - f.cla) * v.wL + (1.0 - f.cra) * v.wR;
	return finc(f, v, k - 2);
}
SPB_HD double var_up(const FaceInV& f, const FaceVals& v, int k) { return v.uf > 0.0 ? f.cl[k - 2] : f.cr[k - 2]; }   // up (TVpure:656)
template <class FaceAt>
SPB_HD void cell_assemble_var(int var, double P_REF, double rx, double ry, bool twoD, const double xs[NV], const VarCell& vc, FaceAt face_at,
                              double Pc, double PW, double PE, double PS, double PN, double X[6], double& s, double& mx, double& my, double& E) {
	FaceInV fa, fb, fc, fd;
	FaceVals fe, fw, fn, fs;
	face_at(0, fa); face_at(1, fb);
	face_eval(fa, Pc, PE, fe); face_eval(fb, PW, Pc, fw);
	if (twoD) { face_at(2, fc); face_at(3, fd); face_eval(fc, Pc, PN, fn); face_eval(fd, PS, Pc, fs); }
	const bool tvx = var >= VAR_TVPURE, tvinc = var == VAR_TVINC || var == VAR_TVSPLIT;
	X[0] = xs[IA1]; X[1] = vc.a2s; X[2] = xs[IM1]; X[3] = xs[IM2]; X[4] = tvx ? vc.eT1 : xs[IE1]; X[5] = tvx ? vc.eT2 : xs[IE2];
	if (tvx)
		for (int k = 4; k < 6; k++) {
			X[k] -= rx * (var_up(fa, fe, k) * fe.uf - var_up(fb, fw, k) * fw.uf);
			if (twoD) X[k] -= ry * (var_up(fc, fn, k) * fn.uf - var_up(fd, fs, k) * fs.uf);
		}
	const int k0 = var == VAR_TV ? 0 : 2, k1 = !tvx ? 6 : tvinc ? 4 : 2;
	for (int k = k0; k < k1; k++) {
		X[k] -= rx * (var_inc(fa, fe, k) - var_inc(fb, fw, k));
		if (twoD) X[k] -= ry * (var_inc(fc, fn, k) - var_inc(fd, fs, k));
	}
	if (var == VAR_TV) s = 1.0 - X[0] - X[1];
	else {
		X[1] = 1.0 - X[0];
		s = rx * (fe.uf - fw.uf) + (twoD ? ry * (fn.uf - fs.uf) : 0.0);
		X[0] *= 1.0 - s; X[1] *= 1.0 - s;
	}
	auto hflux = [&](const FaceInV& f, const FaceVals& v) { return (var_up(f, v, 4) + var_up(f, v, 5) + P_REF + v.pf) * v.uf; };
	auto kinc = [&](const FaceInV& f, const FaceVals& v, double M) {
		return 0.5 * v.uf * v.uf * M + 0.5 * (f.utl * f.utl * (f.cl[0] + f.cl[1]) * v.wL + f.utr * f.utr * (f.cr[0] + f.cr[1]) * v.wR); };
	double Me, Mte, Ee, Mw, Mtw, Ew;
	fflux(fa, fe, P_REF, Me, Mte, Ee); fflux(fb, fw, P_REF, Mw, Mtw, Ew);
	const bool mom_inc = !tvx || tvinc;
	if (mom_inc) { mx = xs[IMU] - rx * (fe.pf - fw.pf) - rx * (fe.uf * Me - fw.uf * Mw); my = xs[IMV] - rx * (Mte - Mtw); }
	else { mx = xs[IMU] - rx * (fe.pf - fw.pf); my = xs[IMV]; }
	if (!tvx) E = xs[IEN] - rx * (Ee - Ew);
	else if (tvinc) E = vc.ET - rx * (hflux(fa, fe) + kinc(fa, fe, Me) - hflux(fb, fw) - kinc(fb, fw, Mw));
	else E = vc.ET - rx * (hflux(fa, fe) - hflux(fb, fw));
	if (twoD) {
		double Mn, Mtn, En, Ms, Mts, Es;
		fflux(fc, fn, P_REF, Mn, Mtn, En); fflux(fd, fs, P_REF, Ms, Mts, Es);
		if (mom_inc) { my -= ry * (fn.pf - fs.pf) + ry * (fn.uf * Mn - fs.uf * Ms); mx -= ry * (Mtn - Mts); }
		else my -= ry * (fn.pf - fs.pf);
		if (!tvx) E -= ry * (En - Es);
		else if (tvinc) E -= ry * (hflux(fc, fn) + kinc(fc, fn, Mn) - hflux(fd, fs) - kinc(fd, fs, Ms));
		else E -= ry * (hflux(fc, fn) - hflux(fd, fs));
	}
}
template <class FaceAt>
SPB_HD bool cell_res_var(int var, const Phys& P, double P_REF, double rx, double ry, bool twoD, const double xs[NV], const VarCell& vc, double Q0,
                         double Q1, FaceAt face_at, double Pc, double PW, double PE, double PS, double PN, double& G, ResOut* out, bool want_sig = true) {
	double X[6], s, mx, my, E;
	cell_assemble_var(var, P_REF, rx, ry, twoD, xs, vc, face_at, Pc, PW, PE, PS, PN, X, s, mx, my, E);
	const double rho = X[2] + X[3];
	if (!(rho > 0)) return false;
	ClIn in; in.atr0 = X[0]; in.atr1 = X[1]; in.mtr0 = X[2]; in.mtr1 = X[3]; in.etr0 = X[4]; in.etr1 = X[5];
	in.s = s; in.rhoeC = E - 0.5 * (mx * mx + my * my) / rho; in.pguess = P_REF + Pc; in.Q0 = Q0; in.Q1 = Q1; in.tau = vc.tau;
	ClOut o;
	if (!closure_of(P, vc.rel, vc.type, in, o, vc.Z0, vc.Z1, vc.nu)) return false;
	G = (o.pbar - P_REF) - Pc;
	if (out) {
		out->a0 = o.a0; out->a1 = o.a1; out->m0 = o.m0; out->m1 = o.m1; out->e0 = o.e0; out->e1 = o.e1; out->mx = mx; out->my = my; out->E = E;
		out->sig = want_sig && vc.type == CL_P ? closure_roundoff(P, in, o.pbar) : 0.0;   // PC:631-638 bounds the p closure only
		out->p = o.pbar;
	}
	return isfin(G);
}
// the faces and cell values of interior row q for a variant (TV: the alpha_1 donors too)
struct CellFacesV {
	const Dev& D; int q, c, i, j;
	__device__ CellFacesV(const Dev& D_, int q_) : D(D_), q(q_) { i = q % D.g.nx + GH; j = q / D.g.nx + GH; c = i + j * D.g.NXT; }
	__device__ void operator()(int w, FaceInV& f) const {
		const bool y = w >= 2;
		const int cf = w == 0 ? c : w == 1 ? c - 1 : w == 2 ? c : c - D.g.NXT;
		load_face(D, y, cf, w == 0 ? i : w == 1 ? i - 1 : w == 2 ? j : j - 1, f);
		if (D.variant == VAR_TV) {
			const double* da = y ? D.donay : D.donax;
			f.cla = da[2 * cf + 1]; f.cra = da[2 * (cf + (y ? D.g.NXT : 1))];
		} else { f.cla = 0.0; f.cra = 0.0; }
	}
	__device__ void cell(double xs[NV], VarCell& vc, double& Q0, double& Q1, double& Kw) const {
		#pragma unroll
		for (int k = 0; k < NV; k++) xs[k] = D.Ust[k * D.g.NC + c];
		vc.a2s = D.a2s[c];
		if (D.variant >= VAR_TVPURE) { vc.eT1 = D.eTV1[c]; vc.eT2 = D.eTV2[c]; vc.ET = D.ETV[c]; } else { vc.eT1 = vc.eT2 = vc.ET = 0.0; }
		vc.type = D.ctype ? D.ctype[q] : CL_P; vc.rel = &D.rel; vc.tau = D.tau;
		if (vc.type == CL_MU) { vc.Z0 = D.Z[c]; vc.Z1 = D.Z[D.g.NC + c]; }
		if (D.rel.cnu) vc.nu = D.rel.cnu[q];   // PC:671
		Q0 = D.Qph[c]; Q1 = D.Qph[D.g.NC + c]; Kw = D.Kw[c];
	}
};
// residual_at, jacobian_at, floor_ok_at and final_at of a variant (the same outputs and operation order, the variant's
// cell assembly)
__device__ __forceinline__ bool residual_var_at(const Dev& D, const double* __restrict__ P, const double* __restrict__ dP, double lam,
                                                double* __restrict__ Pout, double* __restrict__ G, int q, double& nv) {
	const Nb b = neighbours(D.g, q);
	auto pv = [&](int r) { return dP ? P[r] + lam * dP[r] : P[r]; };
	const double Pc = pv(q), PW = pv(b.W), PE = pv(b.E);
	const double PS = D.g.twoD ? pv(b.S) : 0.0, PN = D.g.twoD ? pv(b.N) : 0.0;
	const CellFacesV F(D, q);
	double xs[NV], Q0, Q1, Kw; VarCell vc; F.cell(xs, vc, Q0, Q1, Kw);
	double g;
	bool ok = true;
	nv = 0.0;
	if (!cell_res_var(D.variant, D.ph, D.P_REF, D.rx, D.ry, D.g.twoD, xs, vc, Q0, Q1, F, Pc, PW, PE, PS, PN, g, nullptr)) { ok = false; g = dnan(); }
	else nv = fabs(g) / (fabs(D.P_REF + Pc) + Kw + 1.0);
	G[q] = g;
	if (Pout) Pout[q] = Pc;
	return ok;
}
__device__ __forceinline__ bool jacobian_var_at(const Dev& D, const double* __restrict__ P, const double* __restrict__ G, const double* __restrict__ sc,
                                                double* __restrict__ J, double* __restrict__ rhs, int q) {
	const int n = D.g.n; const bool twoD = D.g.twoD;
	const Nb b = neighbours(D.g, q);
	const CellFacesV F(D, q);
	double xs[NV], Q0, Q1, Kw; VarCell vc; F.cell(xs, vc, Q0, Q1, Kw);
	const double Pc = P[q], PW = P[b.W], PE = P[b.E], PS = twoD ? P[b.S] : 0.0, PN = twoD ? P[b.N] : 0.0, Gq = G[q];
	rhs[q] = -Gq;
	auto res = [&](double a, double w, double e, double s_, double nn, double& g) {
		return cell_res_var(D.variant, D.ph, D.P_REF, D.rx, D.ry, twoD, xs, vc, Q0, Q1, F, a, w, e, s_, nn, g, nullptr); };
	const double h = 1e-7 * sc[q];
	double g1; bool ok = res(Pc + h, PW, PE, PS, PN, g1);
	double v0 = 0, vW = 0, vE = 0, vS = 0, vN = 0;
	if (ok) {
		v0 = (g1 - Gq) / h;
		double hn = 1e-7 * sc[b.W], g2;
		ok = res(Pc, PW + hn, PE, PS, PN, g2); if (ok) vW = (g2 - Gq) / hn;
		if (ok) { hn = 1e-7 * sc[b.E]; ok = res(Pc, PW, PE + hn, PS, PN, g2); if (ok) vE = (g2 - Gq) / hn; }
		if (ok && twoD) { hn = 1e-7 * sc[b.S]; ok = res(Pc, PW, PE, PS + hn, PN, g2); if (ok) vS = (g2 - Gq) / hn; }
		if (ok && twoD) { hn = 1e-7 * sc[b.N]; ok = res(Pc, PW, PE, PS, PN + hn, g2); if (ok) vN = (g2 - Gq) / hn; }
	}
	J[q] = v0; J[n + q] = vW; J[2 * n + q] = vE;
	if (twoD) { J[3 * n + q] = vS; J[4 * n + q] = vN; }
	return ok;
}
__device__ __forceinline__ bool floor_ok_var_at(const Dev& D, const double* __restrict__ P, double tol, int q) {
	const Nb b = neighbours(D.g, q);
	const CellFacesV F(D, q);
	double xs[NV], Q0, Q1, Kw; VarCell vc; F.cell(xs, vc, Q0, Q1, Kw);
	const bool twoD = D.g.twoD;
	double g; ResOut o;
	if (!cell_res_var(D.variant, D.ph, D.P_REF, D.rx, D.ry, twoD, xs, vc, Q0, Q1, F, P[q], P[b.W], P[b.E], twoD ? P[b.S] : 0.0, twoD ? P[b.N] : 0.0, g, &o))
		return false;
	return !(fabs(g) > tol * (fabs(D.P_REF + P[q]) + Kw + 1.0) + o.sig);
}
__device__ __forceinline__ bool final_var_at(const Dev& D, const double* __restrict__ P, const double* __restrict__ Qs, double adt, double* __restrict__ H,
                                             int q, double* T0, double* T1) {
	const Geo& g = D.g; const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	const Nb b = neighbours(g, q);
	const CellFacesV F(D, q);
	double xs[NV], Q0, Q1, Kw; VarCell vc; F.cell(xs, vc, Q0, Q1, Kw);
	const bool twoD = g.twoD;
	double gg; ResOut o;
	if (!cell_res_var(D.variant, D.ph, D.P_REF, D.rx, D.ry, twoD, xs, vc, Q0, Q1, F, P[q], P[b.W], P[b.E], twoD ? P[b.S] : 0.0, twoD ? P[b.N] : 0.0,
	                  gg, &o, false))
		return false;
	const double u[NV] = {o.a0, o.m0, o.m1, o.mx, o.my, o.e0, o.e1, o.E};
	#pragma unroll
	for (int k = 0; k < NV; k++) { D.Ust[k * NC + c] = u[k]; H[k * NC + c] = (u[k] - Qs[k * NC + c]) / adt; }
	const double t0 = T_amp(D.ph, 0, o.a0, o.m0, o.p), t1 = T_amp(D.ph, 1, o.a1, o.m1, o.p);
	*T0 = (o.a0 > 1e-9 && isfin(t0)) ? smax(t0, 0.0) : 0.0;
	*T1 = (o.a1 > 1e-9 && isfin(t1)) ? smax(t1, 0.0) : 0.0;
	return true;
}
// caldt terms of row q (PC:477-488): |u| + c and |v| + c (c only when acoustic) for admissible primitives, else 0
__device__ __forceinline__ void caldt_at(const Geo& g, const Phys& P, const double* __restrict__ U, int acoustic, int q, double& ux, double& uy) {
	ux = 0.0; uy = 0.0;
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	const double a1 = U[IA1 * NC + c], m1 = U[IM1 * NC + c], m2 = U[IM2 * NC + c];
	double cf;
	if (prim_cf(P, a1, m1, m2, U[IE1 * NC + c], U[IE2 * NC + c], cf)) {
		const double rho = m1 + m2, u = U[IMU * NC + c] / rho, v = U[IMV * NC + c] / rho;
		const double cc = acoustic ? cf : 0.0;
		ux = fabs(u) + cc; uy = fabs(v) + cc;
	}
}
// Thomas (PC:807-812) on the 3-point rows J[r] (self), J[n + r] (W: column r - 1, or r at the left boundary), J[2n + r]
// (E: column r + 1, or r at the right); serial, one thread. The recurrences are the reference's, operation for
// operation; rows are taken in blocks of eight whose coefficient loads are issued together, so the dependent chain
// waits on memory once per block instead of once per row.
__device__ __forceinline__ void thomas_serial(int n, const double* __restrict__ J, const double* __restrict__ rhs, double* __restrict__ x,
                                              double* __restrict__ cp, double* __restrict__ dp) {
	auto coef = [&](int r, double& lo, double& di, double& up) {
		lo = 0; up = 0; di = 0;
		di += J[r];
		const int cW = r > 0 ? r - 1 : r, cE = r < n - 1 ? r + 1 : r;
		if (cW == r) di += J[n + r]; else lo = J[n + r];
		if (cE == r) di += J[2 * n + r]; else up = J[2 * n + r];
	};
	constexpr int BLK = 8;
	double lo0, di0, up0;
	coef(0, lo0, di0, up0);
	double cpp = up0 / di0, dpp = rhs[0] / di0;
	cp[0] = cpp; dp[0] = dpp;
	for (int r = 1; r < n; r += BLK) {
		const int m = n - r < BLK ? n - r : BLK;
		double lo[BLK], di[BLK], up[BLK], b[BLK];
		#pragma unroll
		for (int k = 0; k < BLK; k++) if (k < m) { coef(r + k, lo[k], di[k], up[k]); b[k] = rhs[r + k]; }
		#pragma unroll
		for (int k = 0; k < BLK; k++) if (k < m) {
			const double mm = di[k] - lo[k] * cpp;
			cpp = up[k] / mm; dpp = (b[k] - lo[k] * dpp) / mm;
			cp[r + k] = cpp; dp[r + k] = dpp;
		}
	}
	double xn = dpp;   // dp[n - 1]
	x[n - 1] = xn;
	for (int r = n - 2; r >= 0; r -= BLK) {
		const int m = r + 1 < BLK ? r + 1 : BLK;
		double c[BLK], d[BLK];
		#pragma unroll
		for (int k = 0; k < BLK; k++) if (k < m) { c[k] = cp[r - k]; d[k] = dp[r - k]; }
		#pragma unroll
		for (int k = 0; k < BLK; k++) if (k < m) { xn = d[k] - c[k] * xn; x[r - k] = xn; }
	}
}

}   // namespace spb
