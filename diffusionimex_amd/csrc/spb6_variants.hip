// spb6_variants.hip -- gfx950 kernels of the reference's IMEX variants, its sibling programs SPB6_TV_IMEX.cpp,
// SPB6_TVpure_IMEX.cpp, SPB6_TVinc_IMEX.cpp, SPB6_TVsplit_IMEX.cpp, SPB6_TVbb_IMEX.cpp and SPB6_BB_IMEX.cpp: the same
// time step with a different explicit update (Update) and cell residual (cell_residual).
//
// The explicit stage runs from the full face states (k_recon, the reference's TVDReconstruction): face velocities, BB's
// alpha_1 step and face rescaling, the variant's update and the implicit stage's face donors, each a pass of its own
// (one thread per ghosted cell, face or interior cell). The implicit stage's kernels are the reference scheme's with the
// variant's cell assembly (cell_assemble_var). Every operation follows the reference's order, so each variant equals
// the oracle's restatement bit for bit.
#include "spb6_device.h"
#include "spb6_kernels.h"
#include "spb6_cell.h"

namespace spb {

namespace {

inline int nb(long n) { return (int)((n + 255) / 256); }

__device__ __forceinline__ void load8(const double* __restrict__ w, int NC, int c, double* v) {
	#pragma unroll
	for (int k = 0; k < NV; k++) v[k] = w[k * NC + c];
}
// the TV-type variants' explicit kinetic-energy flux (SPB6_TVpure_IMEX.cpp:252-258): k = E - eps_1 - eps_2 moved with
// u, local Rusanov dissipation on its jump
__device__ __forceinline__ double kin_flux(const double* QL, const double* QR, int IM) {
	const double rL = QL[IM1] + QL[IM2], rR = QR[IM1] + QR[IM2], uL = QL[IM] / rL, uR = QR[IM] / rR;
	const double kL = QL[IEN] - QL[IE1] - QL[IE2], kR = QR[IEN] - QR[IE1] - QR[IE2];
	const double a = smax(fabs(uL), fabs(uR));
	return 0.5 * (kL * uL + kR * uR) - 0.5 * a * (kR - kL);
}
// phase energy k's explicit dissipation beyond the face velocity's own (SPB6_TVsplit_IMEX.cpp:259-263)
__device__ __forceinline__ double eps_diss(const double* QL, const double* QR, int IM, int k, double ube) {
	const double uL = QL[IM] / (QL[IM1] + QL[IM2]), uR = QR[IM] / (QR[IM1] + QR[IM2]);
	const double a = smax(fabs(uL), fabs(uR));
	return -0.5 * smax(a - fabs(ube), 0.0) * (QR[k] - QL[k]);
}

// explicit face velocities ube = (u_L + u_R) / 2 of every face the update reads (PC:264-279), from the face states as
// the reconstruction left them (BB rescales the states afterwards; the velocities are the unscaled ones)
__global__ void __launch_bounds__(256) k_var_ube(Geo g, const double* __restrict__ wxL, const double* __restrict__ wxR, const double* __restrict__ wyL,
                                                 const double* __restrict__ wyR, double* __restrict__ ubx, double* __restrict__ uby) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int nfx = (g.nx + 1) * g.ny;
	if (t < nfx) {
		const int c = (t % (g.nx + 1) + GH - 1) + (t / (g.nx + 1) + GH) * g.NXT;
		ubx[c] = face_vel(wxR, c, wxL, c + 1, g.NC, IMU);
		return;
	}
	const int u = t - nfx;
	if (!g.twoD || u >= g.nx * (g.ny + 1)) return;
	const int c = (u % g.nx + GH) + (u / g.nx + GH - 1) * g.NXT;
	uby[c] = face_vel(wyR, c, wyL, c + g.NXT, g.NC, IMV);
}
// BB (SPB6_BB_IMEX.cpp:289-308): alpha_1's explicit step from the unscaled faces, dA = alpha_1* - alpha_1^e per
// interior cell (0 elsewhere)
__global__ void __launch_bounds__(256) k_var_da(Geo g, double rxd, double ryd, const double* __restrict__ Qe, const double* __restrict__ Qs,
                                                const double* __restrict__ wxL, const double* __restrict__ wxR, const double* __restrict__ wyL,
                                                const double* __restrict__ wyR, const double* __restrict__ ubx, const double* __restrict__ uby,
                                                double* __restrict__ dA) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const int NC = g.NC, i = c % g.NXT, j = c / g.NXT;
	if (i < GH || i >= g.nx + GH || j < GH || j >= g.ny + GH) { dA[c] = 0.0; return; }
	double fp[NV], fm[NV];
	rusanov_flux(wxR, c, wxL, c + 1, NC, IMU, fp);
	rusanov_flux(wxR, c - 1, wxL, c, NC, IMU, fm);
	const double qe = Qe[IA1 * NC + c];
	double a = Qs[IA1 * NC + c] - rxd * (fp[IA1] - fm[IA1]) + rxd * qe * (ubx[c] - ubx[c - 1]);
	if (g.twoD) {
		rusanov_flux(wyR, c, wyL, c + g.NXT, NC, IMV, fp);
		rusanov_flux(wyR, c - g.NXT, wyL, c, NC, IMV, fm);
		a += -ryd * (fp[IA1] - fm[IA1]) + ryd * qe * (uby[c] - uby[c - g.NXT]);
	}
	dA[c] = a - qe;
}
// BB's face-state rescaling (SPB6_BB_IMEX.cpp:309-323): each face state of a cell moved to its new alpha_1 by scaling
// the phase masses and energies with the volume fraction ratios, the velocity kept
__device__ __forceinline__ void bb_rescale(double* __restrict__ w, int NC, int c, double d) {
	const double a1 = w[IA1 * NC + c], a2 = 1.0 - a1;
	if (!(a1 > 0 && a2 > 0)) return;
	const double r1 = smax(a1 + d, 0.0) / a1, r2 = smax(a2 - d, 0.0) / a2;
	const double rho = w[IM1 * NC + c] + w[IM2 * NC + c], u = w[IMU * NC + c] / rho, v = w[IMV * NC + c] / rho;
	const double m1 = w[IM1 * NC + c] * r1, e1 = w[IE1 * NC + c] * r1, m2 = w[IM2 * NC + c] * r2, e2 = w[IE2 * NC + c] * r2;
	w[IM1 * NC + c] = m1; w[IE1 * NC + c] = e1; w[IM2 * NC + c] = m2; w[IE2 * NC + c] = e2;
	const double rn = m1 + m2;
	w[IMU * NC + c] = rn * u; w[IMV * NC + c] = rn * v; w[IEN * NC + c] = e1 + e2 + 0.5 * rn * (u * u + v * v);
}
__global__ void __launch_bounds__(256) k_var_rescale(Geo g, const double* __restrict__ dA, double* __restrict__ wxL, double* __restrict__ wxR,
                                                     double* __restrict__ wyL, double* __restrict__ wyR) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const double d = dA[c];
	if (d == 0.0) return;
	bb_rescale(wxL, g.NC, c, d); bb_rescale(wxR, g.NC, c, d);
	if (g.twoD) { bb_rescale(wyL, g.NC, c, d); bb_rescale(wyR, g.NC, c, d); }
}
// the variant's explicit update of interior row q (PC:280-326 with the variant's lines: TV:298 and 315, TVpure:306-307
// and 328, TVsplit:313-316 and 337-340): X* = Qs - adt div F(Qe) with alpha_1 path-conservative (TV: alpha_2* by its own
// flux instead), and the TV-type variants' eTV_k, ETV
__global__ void __launch_bounds__(256) k_var_update(Geo g, int var, double rxd, double ryd, const double* __restrict__ Qe, const double* __restrict__ Qs,
                                                    const double* __restrict__ wxL, const double* __restrict__ wxR, const double* __restrict__ wyL,
                                                    const double* __restrict__ wyR, const double* __restrict__ ubx, const double* __restrict__ uby,
                                                    double* __restrict__ Ust, double* __restrict__ a2s, double* __restrict__ eTV1,
                                                    double* __restrict__ eTV2, double* __restrict__ ETV) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n) return;
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	double fp[NV], fm[NV], o[NV];
	rusanov_flux(wxR, c, wxL, c + 1, NC, IMU, fp);
	rusanov_flux(wxR, c - 1, wxL, c, NC, IMU, fm);
	#pragma unroll
	for (int k = 0; k < NV; k++) o[k] = Qs[k * NC + c] - rxd * (fp[k] - fm[k]);
	const bool split = var == VAR_TVSPLIT || var == VAR_TVBB, tvx = var >= VAR_TVPURE;
	double a2 = 0.0, e1 = 0.0, e2 = 0.0, Et = 0.0;
	if (var == VAR_TV) a2 = (1.0 - Qs[IA1 * NC + c]) - rxd * ((ubx[c] - fp[IA1]) - (ubx[c - 1] - fm[IA1]));
	else if (tvx) {
		double Le[NV], Re[NV], Lw[NV], Rw[NV];
		load8(wxR, NC, c, Le); load8(wxL, NC, c + 1, Re); load8(wxR, NC, c - 1, Lw); load8(wxL, NC, c, Rw);
		if (split) {
			const double d1 = eps_diss(Le, Re, IMU, IE1, ubx[c]) - eps_diss(Lw, Rw, IMU, IE1, ubx[c - 1]);
			const double d2 = eps_diss(Le, Re, IMU, IE2, ubx[c]) - eps_diss(Lw, Rw, IMU, IE2, ubx[c - 1]);
			e1 = Qs[IE1 * NC + c] - rxd * d1; e2 = Qs[IE2 * NC + c] - rxd * d2;
			Et = Qs[IEN * NC + c] - rxd * (kin_flux(Le, Re, IMU) - kin_flux(Lw, Rw, IMU) + d1 + d2);
		} else {
			e1 = Qs[IE1 * NC + c]; e2 = Qs[IE2 * NC + c];
			Et = Qs[IEN * NC + c] - rxd * (kin_flux(Le, Re, IMU) - kin_flux(Lw, Rw, IMU));
		}
	}
	if (var != VAR_TV) o[IA1] += rxd * Qe[IA1 * NC + c] * (ubx[c] - ubx[c - 1]);
	if (g.twoD) {
		const int cs = c - g.NXT, cn = c + g.NXT;
		double gp[NV], gm[NV];
		rusanov_flux(wyR, c, wyL, cn, NC, IMV, gp);
		rusanov_flux(wyR, cs, wyL, c, NC, IMV, gm);
		#pragma unroll
		for (int k = 0; k < NV; k++) o[k] -= ryd * (gp[k] - gm[k]);
		if (var == VAR_TV) a2 -= ryd * ((uby[c] - gp[IA1]) - (uby[cs] - gm[IA1]));
		else if (tvx) {
			double Ln[NV], Rn[NV], Ls[NV], Rs[NV];
			load8(wyR, NC, c, Ln); load8(wyL, NC, cn, Rn); load8(wyR, NC, cs, Ls); load8(wyL, NC, c, Rs);
			if (split) {
				const double d1 = eps_diss(Ln, Rn, IMV, IE1, uby[c]) - eps_diss(Ls, Rs, IMV, IE1, uby[cs]);
				const double d2 = eps_diss(Ln, Rn, IMV, IE2, uby[c]) - eps_diss(Ls, Rs, IMV, IE2, uby[cs]);
				e1 -= ryd * d1; e2 -= ryd * d2;
				Et -= ryd * (kin_flux(Ln, Rn, IMV) - kin_flux(Ls, Rs, IMV) + d1 + d2);
			} else Et -= ryd * (kin_flux(Ln, Rn, IMV) - kin_flux(Ls, Rs, IMV));
		}
		if (var != VAR_TV) o[IA1] += ryd * Qe[IA1 * NC + c] * (uby[c] - uby[cs]);
	}
	#pragma unroll
	for (int k = 0; k < NV; k++) Ust[k * NC + c] = o[k];
	a2s[c] = var == VAR_TV ? a2 : 1.0 - o[IA1];
	if (tvx) { eTV1[c] = e1; eTV2[c] = e2; ETV[c] = Et; }
}
// the implicit stage's face donors of every ghosted cell from its face states (PC:866-873): the (left, right) pairs of
// m1, m2, e1, e2 per direction, and TV's alpha_1 pairs
__global__ void __launch_bounds__(256) k_var_donors(Geo g, int var, const double* __restrict__ wxL, const double* __restrict__ wxR,
                                                    const double* __restrict__ wyL, const double* __restrict__ wyR, double* __restrict__ donx,
                                                    double* __restrict__ dony, double* __restrict__ donax, double* __restrict__ donay) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const int NC = g.NC;
	double L[NV], R[NV];
	load8(wxL, NC, c, L); load8(wxR, NC, c, R);
	put_donors<false>(donx, NC, c, L, R);
	if (var == VAR_TV) { donax[2 * c] = L[IA1]; donax[2 * c + 1] = R[IA1]; }
	if (g.twoD) {
		load8(wyL, NC, c, L); load8(wyR, NC, c, R);
		put_donors<false>(dony, NC, c, L, R);
		if (var == VAR_TV) { donay[2 * c] = L[IA1]; donay[2 * c + 1] = R[IA1]; }
	}
}

// implicit stage of a variant: the reference scheme's kernels with the variant's cell residual
__global__ void __launch_bounds__(256) k_residual_var(Dev D, const double* __restrict__ P, const double* __restrict__ dP, double lam,
                                                      double* __restrict__ Pout, double* __restrict__ G, unsigned long long* __restrict__ nrm,
                                                      int* __restrict__ flags, int flag_slot, const int* __restrict__ gate) {
	if (gate && !*gate) return;
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double nv = 0.0;
	if (q < D.g.n && !residual_var_at(D, P, dP, lam, Pout, G, q, nv)) flags[flag_slot] = 1;
	atomic_max_nonneg(nrm, nv);
}
__global__ void __launch_bounds__(256) k_jacobian_var(Dev D, const double* __restrict__ P, const double* __restrict__ G, const double* __restrict__ sc,
                                                      double* __restrict__ J, double* __restrict__ rhs, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	if (!jacobian_var_at(D, P, G, sc, J, rhs, q)) flags[FLAG_JAC] = 1;
}
__global__ void __launch_bounds__(256) k_floor_var(Dev D, const double* __restrict__ P, double tol, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	if (!floor_ok_var_at(D, P, tol, q)) flags[FLAG_FLOOR] = 1;
}
__global__ void __launch_bounds__(256) k_final_var(Dev D, const double* __restrict__ P, const double* __restrict__ Qs, double adt, double* __restrict__ H,
                                                   int* __restrict__ flags, int flag_slot, unsigned long long* __restrict__ tmax) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double T0 = 0.0, T1 = 0.0;
	if (q < D.g.n && !final_var_at(D, P, Qs, adt, H, q, &T0, &T1)) flags[flag_slot] = 1;
	atomic_max_nonneg(tmax, T0);
	atomic_max_nonneg(tmax + 64, T1);
}

// RelaxationActivation at levels 1-3 (PC:513-560): closure_type_of (PHY:814-830) gives each cell's type from its
// content and specific energy (cell_type_U, PC:517-521), and project_cell moves a pT or pTg cell onto that closure's
// equilibrium (s = 0, the guess max(p_mix, 1); the pTg closure moves mass); a cell whose projection fails stays
// mechanical and unprojected
__global__ void __launch_bounds__(256) k_relax_activate(Geo g, Phys P, RelaxDev S, int level, double* __restrict__ U, double eps_i,
                                                       int* __restrict__ ctype) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n) return;
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	const double a1 = U[IA1 * NC + c], m1 = U[IM1 * NC + c], m2 = U[IM2 * NC + c], mu = U[IMU * NC + c], mv = U[IMV * NC + c];
	const double e1 = U[IE1 * NC + c], e2 = U[IE2 * NC + c];
	const double rho = m1 + m2, rhoe = U[IEN * NC + c] - 0.5 * (mu * mu + mv * mv) / rho;
	int ty = closure_type_of(P, &S, level, eps_i, a1, m1, m2, e1, e2, rhoe);
	if (ty == CL_PT || ty == CL_PTG) {
		ClIn in;
		in.atr0 = a1; in.atr1 = 1.0 - a1; in.mtr0 = m1; in.mtr1 = m2; in.etr0 = e1; in.etr1 = e2; in.s = 0.0; in.rhoeC = rhoe;
		in.pguess = smax(p_mix(P, in.atr0, in.mtr0, in.mtr1, in.rhoeC), 1.0); in.Q0 = 0.0; in.Q1 = 0.0;
		ClOut o;
		if (closure_of(P, &S, ty, in, o)) {
			U[IA1 * NC + c] = o.a0; U[IM1 * NC + c] = o.m0; U[IM2 * NC + c] = o.m1; U[IE1 * NC + c] = o.e0; U[IE2 * NC + c] = o.e1;
		} else ty = __builtin_isinf(S.mu) ? CL_P : CL_MU;   // PC:544
	}
	ctype[q] = ty;
	if (S.nu_law) {   // PC:546-556: the chemical cell's rate from its pT equilibrium (p, T, vapour fraction)
		double nu = 0.0;
		if (ty == CL_NU && a1 > 0 && 1.0 - a1 > 0 && m1 > 0 && m2 > 0 && isfin(e1) && isfin(e2) && isfin(rhoe)) {
			ClIn in;
			in.atr0 = a1; in.atr1 = 1.0 - a1; in.mtr0 = m1; in.mtr1 = m2; in.etr0 = e1; in.etr1 = e2; in.s = 0.0; in.rhoeC = rhoe;
			in.pguess = smax(p_mix(P, in.atr0, in.mtr0, in.mtr1, in.rhoeC), 1.0); in.Q0 = 0.0; in.Q1 = 0.0;
			ClOut o; double T;
			if (closure_pT(P, in, o, &T)) nu = nu_law(P, S, o.pbar, T, o.a1);
		}
		S.cnu[q] = nu;
	}
}

}   // namespace

void launch_relax_activate(const Geo& g, const Phys& P, const RelaxDev& S, int level, double* U, double eps_i, int* ctype, hipStream_t st) {
	hipLaunchKernelGGL(k_relax_activate, dim3(nb(g.n)), dim3(256), 0, st, g, P, S, level, U, eps_i, ctype);
	launch_bc_grid(g, U, st);
}

// (after k_scales, which launch_explicit runs first)
void launch_explicit_var(const Dev& D, int limiter, int order, const double* Qe, const double* Qs, double adt, hipStream_t st) {
	const Geo& g = D.g;
	const double rxd = adt / g.dx, ryd = adt / g.dy;
	launch_recon_full(D, limiter, order, Qe, D.wxL, D.wxR, D.wyL, D.wyR, st);
	const long nf = (long)(g.nx + 1) * g.ny + (g.twoD ? (long)g.nx * (g.ny + 1) : 0);
	hipLaunchKernelGGL(k_var_ube, dim3(nb(nf)), dim3(256), 0, st, g, (const double*)D.wxL, (const double*)D.wxR, (const double*)D.wyL,
	                   (const double*)D.wyR, D.ubx, D.uby);
	if (D.variant == VAR_BB) {
		hipLaunchKernelGGL(k_var_da, dim3(nb(g.NC)), dim3(256), 0, st, g, rxd, ryd, Qe, Qs, (const double*)D.wxL, (const double*)D.wxR,
		                   (const double*)D.wyL, (const double*)D.wyR, (const double*)D.ubx, (const double*)D.uby, D.dA);
		hipLaunchKernelGGL(k_var_rescale, dim3(nb(g.NC)), dim3(256), 0, st, g, (const double*)D.dA, D.wxL, D.wxR, D.wyL, D.wyR);
	}
	hipLaunchKernelGGL(k_var_update, dim3(nb(g.n)), dim3(256), 0, st, g, D.variant, rxd, ryd, Qe, Qs, (const double*)D.wxL, (const double*)D.wxR,
	                   (const double*)D.wyL, (const double*)D.wyR, (const double*)D.ubx, (const double*)D.uby, D.Ust, D.a2s, D.eTV1, D.eTV2, D.ETV);
	hipLaunchKernelGGL(k_var_donors, dim3(nb(g.NC)), dim3(256), 0, st, g, D.variant, (const double*)D.wxL, (const double*)D.wxR,
	                   (const double*)D.wyL, (const double*)D.wyR, D.donx, D.dony, D.donax, D.donay);
	launch_bc_grid(g, D.Ust, st);
	launch_bc_array(g, D.a2s, st);
}
void launch_residual_var(const Dev& D, const double* P, const double* dP, double lam, double* Pout, double* G, unsigned long long* nrm, int flag_slot,
                         hipStream_t st, const int* gate) {
	hipLaunchKernelGGL(k_residual_var, dim3(nb(D.g.n)), dim3(256), 0, st, D, P, dP, lam, Pout, G, nrm, D.flags, flag_slot, gate);
}
void launch_jacobian_var(const Dev& D, const double* P, const double* G, const double* sc, double* J, double* rhs, hipStream_t st) {
	hipLaunchKernelGGL(k_jacobian_var, dim3(nb(D.g.n)), dim3(256), 0, st, D, P, G, (const double*)sc, J, rhs, D.flags);
}
void launch_floor_var(const Dev& D, const double* P, double tol, hipStream_t st) {
	hipLaunchKernelGGL(k_floor_var, dim3(nb(D.g.n)), dim3(256), 0, st, D, P, tol, D.flags);
}
void launch_final_var(const Dev& D, const double* P, const double* Qs, double adt, double* H, unsigned long long* tmax, hipStream_t st, int flag_slot) {
	hipLaunchKernelGGL(k_final_var, dim3(nb(D.g.n)), dim3(256), 0, st, D, P, Qs, adt, H, D.flags, flag_slot, tmax);
}

}   // namespace spb
