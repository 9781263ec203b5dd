// spb6_kernels.hip -- gfx950 kernels of the SPB 6+1 semi-implicit IMEX-RK step (FP64, structure of arrays).
//
// Device layout: every cell field is its own array over the ghosted grid, index c = i + j*NXT (NXT = nx+4, NYT = ny+4,
// two ghost layers); conserved grids hold 8 such arrays back to back (field k at base + k*NC). Pressure-system vectors
// (P, G, Krylov vectors, Jacobian diagonals) are interior-only, row r = ii + jj*nx. One thread per cell or face row;
// 256-thread blocks (4 wavefronts of 64). Reductions use one fixed tree order (see block_sum256 / grid_total), so every
// sum is bit-reproducible and equal to the CPU oracle's "device order".
#include "spb6_device.h"
#include "spb6_kernels.h"

namespace spb {

// ------------------------------------------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_bfly_sum(double v) {
	#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
	return v;
}
__device__ __forceinline__ double wave_max(double v) {
	#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) { const double o = __shfl_xor(v, off, 64); v = smax(v, o); }
	return v;
}
// 256-thread block total in the fixed order: xor butterflies in each wavefront, then (w0 + w1) + (w2 + w3)
__device__ __forceinline__ double block_sum256(double v, double* sh) {
	v = wave_bfly_sum(v);
	__syncthreads();
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
	__syncthreads();
	return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}
// non-negative double max through its bit pattern (monotone for x >= 0): block maximum, then one atomic per block on one
// of 64 striped slots (dst[0..63]); the host takes the maximum of the 64 slots
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

// ------------------------------------------------------------------------------------------------ boundaries (PHY:926-957)
// x pass over all rows (ghost columns), then y pass over all columns; the second pass reads the corners the first wrote
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
__global__ void __launch_bounds__(256) k_bc_x(Geo g, double* u) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;   // (row j, layer q, side)
	if (t >= g.NYT * GH * 2) return;
	const int side = t & 1, q = (t >> 1) % GH, j = (t >> 1) / GH, nx = g.nx;
	if (side == 0) {
		const int iL = GH - 1 - q, sL = g.bc[0] == BC_PERIODIC ? nx + GH - 1 - q : (g.bc[0] == BC_WALL ? GH + q : GH);
		ghost_copy(g, u, iL + j * g.NXT, sL + j * g.NXT, 0);
	} else {
		const int iR = nx + GH + q, sR = g.bc[1] == BC_PERIODIC ? GH + q : (g.bc[1] == BC_WALL ? nx + GH - 1 - q : nx + GH - 1);
		ghost_copy(g, u, iR + j * g.NXT, sR + j * g.NXT, 1);
	}
}
__global__ void __launch_bounds__(256) k_bc_y(Geo g, double* u) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;   // (column i, layer q, side)
	if (t >= g.NXT * GH * 2) return;
	const int side = t & 1, q = (t >> 1) % GH, i = (t >> 1) / GH, ny = g.ny;
	if (ny > 1 || g.bc[2] == BC_PERIODIC) {
		if (side == 0) {
			const int jL = GH - 1 - q, sL = g.bc[2] == BC_PERIODIC ? ny + GH - 1 - q : (g.bc[2] == BC_WALL ? GH + q : GH);
			ghost_copy(g, u, i + jL * g.NXT, i + sL * g.NXT, 2);
		} else {
			const int jR = ny + GH + q, sR = g.bc[3] == BC_PERIODIC ? GH + q : (g.bc[3] == BC_WALL ? ny + GH - 1 - q : ny + GH - 1);
			ghost_copy(g, u, i + jR * g.NXT, i + sR * g.NXT, 3);
		}
	} else {   // 1D: ghost rows are copies of the interior row
		const int dst = side == 0 ? GH - 1 - q : GH + 1 + q;
		for (int k = 0; k < NV; k++) u[k * g.NC + i + dst * g.NXT] = u[k * g.NC + i + GH * g.NXT];
	}
}
// array ghost fill with unit signs (bound_cond_array, PC:490-511)
__global__ void __launch_bounds__(256) k_bcarr_x(Geo g, double* a) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= g.NYT * GH) return;
	const int q = t % GH, j = t / GH, nx = g.nx;
	auto src = [&](int side) { return g.bc[side] == BC_PERIODIC ? ((side % 2 == 0) ? nx + GH - 1 - q : GH + q)
		: g.bc[side] == BC_WALL ? ((side % 2 == 0) ? GH + q : nx + GH - 1 - q) : ((side % 2 == 0) ? GH : nx + GH - 1); };
	a[(GH - 1 - q) + j * g.NXT] = a[src(0) + j * g.NXT];
	a[(nx + GH + q) + j * g.NXT] = a[src(1) + j * g.NXT];
}
__global__ void __launch_bounds__(256) k_bcarr_y(Geo g, double* a) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= g.NXT * GH) return;
	const int q = t % GH, i = t / GH, ny = g.ny;
	int sL, sR;
	if (ny == 1 && g.bc[2] != BC_PERIODIC) { sL = GH; sR = GH; }
	else {
		auto src = [&](int side) { return g.bc[side] == BC_PERIODIC ? ((side % 2 == 0) ? ny + GH - 1 - q : GH + q)
			: g.bc[side] == BC_WALL ? ((side % 2 == 0) ? GH + q : ny + GH - 1 - q) : ((side % 2 == 0) ? GH : ny + GH - 1); };
		sL = src(2); sR = src(3);
	}
	a[i + (GH - 1 - q) * g.NXT] = a[i + sL * g.NXT];
	a[i + (ny + GH + q) * g.NXT] = a[i + sR * g.NXT];
}

// ------------------------------------------------------------------------------------------------ stage combination (PC:1255-1264)
__global__ void __launch_bounds__(256) k_combine(Geo g, const double* __restrict__ Un, const double* __restrict__ H0, const double* __restrict__ H1,
                          int nr, double ae0, double ae1, double ai0, double ai1, double* __restrict__ Qe, double* __restrict__ Qs) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n) return;
	const int c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	for (int k = 0; k < NV; k++) {
		const int o = k * g.NC + c;
		double e = Un[o], s = Un[o];
		e = e + ae0 * H0[o]; s = s + ai0 * H0[o];
		if (nr > 1) { e = e + ae1 * H1[o]; s = s + ai1 * H1[o]; }
		Qe[o] = e; Qs[o] = s;
	}
}

// ------------------------------------------------------------------------------------------------ explicit stage
// reconstruction scales (PC:356-366) and the rho > 0 check of Update (PC:255-258) over all cells of Qe
__global__ void __launch_bounds__(256) k_scales(Geo g, Phys P, const double* __restrict__ Q, double* __restrict__ vsc, double* __restrict__ esc, int* __restrict__ flags) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const int NC = g.NC;
	const double a1 = Q[IA1 * NC + c], m1 = Q[IM1 * NC + c], m2 = Q[IM2 * NC + c], mu = Q[IMU * NC + c], mv = Q[IMV * NC + c];
	const double e1 = Q[IE1 * NC + c], e2 = Q[IE2 * NC + c];
	const double rho = m1 + m2, mm = phypot(mu, mv);
	double cf;
	vsc[c] = (rho > 0 ? mm / rho : 0.0) + (prim_cf(P, a1, m1, m2, e1, e2, cf) ? cf : 0.0);
	esc[c] = fabs(e1) + fabs(e2) + (rho > 0 ? 0.5 * mm * mm / rho : 0.0);
	if (!(rho > 0)) flags[FLAG_RHO] = 1;
}
// TVD reconstruction of one cell along one direction (PC:392-445): face values L, R (8 vars) and the limiter factor
__device__ __forceinline__ double recon_cell(const Geo& g, const Phys& P, int limiter, const double* __restrict__ Q,
                                             const double* __restrict__ vsc, const double* __restrict__ esc,
                                             int ca, int cc, int cb, double* L, double* R) {
	const int NC = g.NC;
	const int VOL[5] = {IA1, IM1, IM2, IE1, IE2};
	double xa[5], xc[5], xb[5];
	#pragma unroll
	for (int q = 0; q < 5; q++) { xa[q] = Q[VOL[q] * NC + ca]; xc[q] = Q[VOL[q] * NC + cc]; xb[q] = Q[VOL[q] * NC + cb]; }
	const double esum = esc[ca] + esc[cc] + esc[cb];
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
	const double tiny = 1e-12 * (vsc[ca] + vsc[cc] + vsc[cb]) + 1e-300;
	double vL[2], vR[2], sm = 1.0;
	#pragma unroll
	for (int d = 0; d < 2; d++) {
		const int k = d == 0 ? IMU : IMV;
		const double va = Q[k * NC + ca] / ra, vc = Q[k * NC + cc] / rc, vb = Q[k * NC + cb] / rb, a = vc - va, b = vb - vc;
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
__global__ void __launch_bounds__(256) k_recon(Geo g, Phys P, int limiter, int order, const double* __restrict__ Q, const double* __restrict__ vsc,
                        const double* __restrict__ esc, double* __restrict__ wxL, double* __restrict__ wxR,
                        double* __restrict__ wyL, double* __restrict__ wyR, double* __restrict__ phx, double* __restrict__ phy) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const int NC = g.NC, i = c % g.NXT, j = c / g.NXT;
	double L[NV], R[NV];
	// x direction (all rows; cells with both x neighbours)
	if (order >= 2 && i >= 1 && i + 1 < g.NXT) {
		phx[c] = recon_cell(g, P, limiter, Q, vsc, esc, c - 1, c, c + 1, L, R);
		for (int k = 0; k < NV; k++) { wxL[k * NC + c] = L[k]; wxR[k * NC + c] = R[k]; }
	} else {
		phx[c] = 0.0;
		for (int k = 0; k < NV; k++) { const double v = Q[k * NC + c]; wxL[k * NC + c] = v; wxR[k * NC + c] = v; }
	}
	if (!g.twoD) return;
	if (order >= 2 && j >= 1 && j + 1 < g.NYT) {
		phy[c] = recon_cell(g, P, limiter, Q, vsc, esc, c - g.NXT, c, c + g.NXT, L, R);
		for (int k = 0; k < NV; k++) { wyL[k * NC + c] = L[k]; wyR[k * NC + c] = R[k]; }
	} else {
		phy[c] = 0.0;
		for (int k = 0; k < NV; k++) { const double v = Q[k * NC + c]; wyL[k * NC + c] = v; wyR[k * NC + c] = v; }
	}
}
// Rusanov flux with the local material speed between face states L (cell cL's side) and R (PC:224-246)
__device__ __forceinline__ void rusanov_flux(const double* __restrict__ wR_, int cL, const double* __restrict__ wL_, int cR, int NC, int dirk, double* F) {
	double QL[NV], QR[NV];
	#pragma unroll
	for (int k = 0; k < NV; k++) { QL[k] = wR_[k * NC + cL]; QR[k] = wL_[k * NC + cR]; }
	const double uL = QL[dirk] / (QL[IM1] + QL[IM2]), uR = QR[dirk] / (QR[IM1] + QR[IM2]);
	const double a = smax(fabs(QL[dirk] / (QL[IM1] + QL[IM2])), fabs(QR[dirk] / (QR[IM1] + QR[IM2])));
	#pragma unroll
	for (int k = 0; k < NV; k++) { const double fL = QL[k] * uL, fR = QR[k] * uR; F[k] = 0.5 * (fL + fR) - 0.5 * a * (QR[k] - QL[k]); }
}
__device__ __forceinline__ double face_vel(const double* __restrict__ wR_, int cL, const double* __restrict__ wL_, int cR, int NC, int dirk) {
	const double uL = wR_[dirk * NC + cL] / (wR_[IM1 * NC + cL] + wR_[IM2 * NC + cL]);
	const double uR = wL_[dirk * NC + cR] / (wL_[IM1 * NC + cR] + wL_[IM2 * NC + cR]);
	return 0.5 * (uL + uR);
}
// X* = Qs - adt div F(Qe) with the path-conservative alpha1 term (PC:251-326); interior cells
__global__ void __launch_bounds__(256) k_update(Geo g, double rxd, double ryd, const double* __restrict__ Qe, const double* __restrict__ Qs,
                         const double* __restrict__ wxL, const double* __restrict__ wxR, const double* __restrict__ wyL,
                         const double* __restrict__ wyR, double* __restrict__ Ust, double* __restrict__ a2s,
                         double* __restrict__ ubx, double* __restrict__ uby) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n) return;
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
// per cell: donor velocities, frozen sound speed, residual scale K + P_E, previous pressure (PC:859-900)
__global__ void __launch_bounds__(256) k_prep_cell(Geo g, Phys P, const double* __restrict__ Xs, const double* __restrict__ a2s, const double* __restrict__ Pst,
                            double* __restrict__ ust, double* __restrict__ vst, double* __restrict__ cz, double* __restrict__ Kw,
                            double* __restrict__ Pold, int* __restrict__ flags) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const int NC = g.NC, i = c % g.NXT, j = c / g.NXT;
	const double a1 = Xs[IA1 * NC + c], a2 = a2s[c], m1 = Xs[IM1 * NC + c], m2 = Xs[IM2 * NC + c];
	const double e1 = Xs[IE1 * NC + c], e2 = Xs[IE2 * NC + c];
	const double sa = a1 + a2, rho = m1 + m2;
	if (!(sa > 0 && rho > 0)) flags[FLAG_STAGE] = 1;
	ust[c] = Xs[IMU * NC + c] / rho; vst[c] = Xs[IMV * NC + c] / rho;
	double czv = dinf(), K = smax(P.pi[0], P.pi[1]), PE = 0.0;
	if (a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0) {
		const double p1 = p_ame(P, 0, a1, m1, e1), p2 = p_ame(P, 1, a2, m2, e2);
		if (p_adm(P, 0, p1) && p_adm(P, 1, p2)) {
			const double c21 = c2_rp(P, 0, m1 / a1, p1), c22 = c2_rp(P, 1, m2 / a2, p2);
			const double c2 = (m1 * c21 + m2 * c22) / (m1 + m2);
			czv = c2 > 0 ? sqrt(c2) : dinf();
			K = 1.0 / ((a1 / sa) / (m1 / a1 * c21) + (a2 / sa) / (m2 / a2 * c22));
		}
		const double dd = dedp_am(P, 0, a1, m1) + dedp_am(P, 1, a2, m2);
		if (dd > 0) PE = (fabs(e1) + fabs(m1 * P.q[0]) + fabs(e2) + fabs(m2 * P.q[1])) / dd;
	}
	cz[c] = czv; Kw[c] = K + PE;
	Pold[c] = Pst[map_P(g, i, j)];
}
__device__ __forceinline__ double mmod(double a, double b) { return a * b <= 0 ? 0.0 : (fabs(a) < fabs(b) ? a : b); }
__device__ __forceinline__ double psi_face(int order, int limiter, double xm, double xL, double xR, double xp) {   // PC:901-905
	if (order < 2) return 1.0;
	const double D = xR - xL; if (D == 0.0) return 1.0;
	if (!limiter) return smin(1.0, smax(0.0, 1.0 - 0.25 * ((xR - xm) + (xp - xL)) / D));
	return 1.0 - 0.5 * (mmod(xL - xm, D) + mmod(D, xp - xR)) / D;
}
// face coefficients of u_f = ubs - kp (P_R - P_L) and p_f = wL P_L + (1 - wL) P_R - pv (set_face, PC:926-950)
__global__ void __launch_bounds__(256) k_prep_face(Geo g, int order, int limiter, double tau, const double* __restrict__ Xs, const double* __restrict__ ust,
                            const double* __restrict__ vst, const double* __restrict__ cz, const double* __restrict__ Pold,
                            const double* __restrict__ phx, const double* __restrict__ phy,
                            double* __restrict__ ubsx, double* __restrict__ kpx, double* __restrict__ wLx, double* __restrict__ pvx,
                            double* __restrict__ ubsy, double* __restrict__ kpy, double* __restrict__ wLy, double* __restrict__ pvy) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
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
	// x face (i, j) = i + 1/2
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
// shock budget of each interior cell on the +-1 cross stencil of X* (PHY:395-418)
__global__ void __launch_bounds__(64) k_shock(Geo g, Phys P, const double* __restrict__ Xs, double* __restrict__ Qph) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n) return;
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	auto load = [&](int cc, double* u) { for (int k = 0; k < NV; k++) u[k] = Xs[k * NC + cc]; };
	double Q0 = 0.0, Q1 = 0.0;
	double u[NV], c0[7], st[7], lo[7] = {0, 0, 0, 0, 0, 0, 0}, plo = 1e300, phi = -1e300, div = 0.0;
	load(c, u); stencil_state(P, u, c0);
	for (int d = 0; d < (g.twoD ? 2 : 1); d++) for (int s = -1; s <= 1; s++) {
		if (d == 1 && s == 0) continue;
		load(c + (d == 0 ? s : s * g.NXT), u); stencil_state(P, u, st);
		if (s == 1) div += st[5 + d]; else if (s == -1) div -= st[5 + d];
		if (!isfin(st[4])) continue;
		if (st[4] < plo) { plo = st[4]; for (int n = 0; n < 7; n++) lo[n] = st[n]; }
		if (st[4] > phi) phi = st[4];
	}
	if ((div < 0.0) && (phi > plo)) {
		const double rho = c0[2] + c0[3];
		if (rho > 0) {
			double Y[2], vpre[2], lam[2], qq;
			for (int k = 0; k < 2; k++) {
				Y[k] = c0[2 + k] / rho;
				vpre[k] = (lo[k] > 0 && lo[2 + k] > 0) ? lo[k] / lo[2 + k] : (c0[2 + k] > 0 ? c0[k] / c0[2 + k] : 0.0);
			}
			if ((Y[0] > 0) && (Y[1] > 0) && (vpre[0] > 0) && (vpre[1] > 0) && shock_share(P, Y, vpre, plo, phi, lam, qq)) {
				Q0 = lam[0] * rho * qq; Q1 = lam[1] * rho * qq;
			}
		}
	}
	Qph[c] = Q0; Qph[NC + c] = Q1;
}

// ------------------------------------------------------------------------------------------------ cell residual (PC:584-705)
struct FaceIn { double wl, pv, ubs, kp, ube, cl[4], cr[4], utl, utr, bw; int wall; };
struct CellIn { double xs[NV], a2s, Q0, Q1, Kw; FaceIn fe, fw, fn, fs; };
struct FaceVals { double uf, pf, wL, wR; };
struct ResOut { double a0, a1, m0, m1, e0, e1, mx, my, E, sig; };

__device__ __forceinline__ void load_face(const Dev& D, bool ydir, int c, int idx, FaceIn& f) {
	const Geo& g = D.g; const int NC = g.NC, cn = c + (ydir ? g.NXT : 1);
	const double* wLf = ydir ? D.wLy : D.wLx; const double* pvf = ydir ? D.pvy : D.pvx;
	const double* ubsf = ydir ? D.ubsy : D.ubsx; const double* kpf = ydir ? D.kpy : D.kpx; const double* ubef = ydir ? D.uby : D.ubx;
	const double* wR = ydir ? D.wyR : D.wxR; const double* wLs = ydir ? D.wyL : D.wxL;
	const double* ut = ydir ? D.ust : D.vst;
	f.wl = wLf[c]; f.pv = pvf[c]; f.ubs = ubsf[c]; f.kp = kpf[c]; f.ube = ubef[c];
	f.cl[0] = wR[IM1 * NC + c]; f.cl[1] = wR[IM2 * NC + c]; f.cl[2] = wR[IE1 * NC + c]; f.cl[3] = wR[IE2 * NC + c];
	f.cr[0] = wLs[IM1 * NC + cn]; f.cr[1] = wLs[IM2 * NC + cn]; f.cr[2] = wLs[IE1 * NC + cn]; f.cr[3] = wLs[IE2 * NC + cn];
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
__device__ __forceinline__ void face_eval(const FaceIn& f, double PL, double PR, FaceVals& v) {   // face_x / face_y (PC:597-616)
	v.pf = f.wl * PL + (1.0 - f.wl) * PR - f.pv;
	if (f.wall >= 0 && f.bw == 0.0) { v.uf = 0.0; v.wL = 0.0; v.wR = 0.0; return; }
	v.uf = f.wall >= 0 ? f.bw : f.ubs - f.kp * (PR - PL);
	v.wL = smax(v.uf, 0.0) - smax(f.ube, 0.0); v.wR = smin(v.uf, 0.0) - smin(f.ube, 0.0);   // face_weights (PC:592)
}
__device__ __forceinline__ double finc(const FaceIn& f, const FaceVals& v, int k) { return f.cl[k] * v.wL + f.cr[k] * v.wR; }   // inc (PC:617), k in m1 m2 e1 e2
// face_flux (PC:620-625): increment mass, tangential momentum and energy fluxes
__device__ __forceinline__ void fflux(const FaceIn& f, const FaceVals& v, double P_REF, double& M, double& Mt, double& E) {
	const double ML = (f.cl[0] + f.cl[1]) * v.wL, MR = (f.cr[0] + f.cr[1]) * v.wR;
	M = ML + MR; Mt = f.utl * ML + f.utr * MR;
	E = finc(f, v, 2) + finc(f, v, 3) + 0.5 * v.uf * v.uf * M + 0.5 * (f.utl * f.utl * ML + f.utr * f.utr * MR) + (P_REF + v.pf) * v.uf;
}
// G = (pbar_cl - P_REF) - Pc of one cell (cell_residual, PC:640-679)
__device__ __forceinline__ bool cell_res(const Phys& P, double P_REF, double rx, double ry, bool twoD, const CellIn& C,
                                         double Pc, double PW, double PE, double PS, double PN, double& G, ResOut* out) {
	FaceVals fe, fw, fn, fs;
	face_eval(C.fe, Pc, PE, fe); face_eval(C.fw, PW, Pc, fw);
	if (twoD) { face_eval(C.fn, Pc, PN, fn); face_eval(C.fs, PS, Pc, fs); }
	double X[6] = {C.xs[IA1], C.a2s, C.xs[IM1], C.xs[IM2], C.xs[IE1], C.xs[IE2]};
	#pragma unroll
	for (int k = 2; k < 6; k++) {
		X[k] -= rx * (finc(C.fe, fe, k - 2) - finc(C.fw, fw, k - 2));
		if (twoD) X[k] -= ry * (finc(C.fn, fn, k - 2) - finc(C.fs, fs, k - 2));
	}
	X[1] = 1.0 - X[0];
	const double s = rx * (fe.uf - fw.uf) + (twoD ? ry * (fn.uf - fs.uf) : 0.0);
	X[0] *= 1.0 - s; X[1] *= 1.0 - s;
	double Me, Mte, Ee, Mw, Mtw, Ew;
	fflux(C.fe, fe, P_REF, Me, Mte, Ee); fflux(C.fw, fw, P_REF, Mw, Mtw, Ew);
	double mx = C.xs[IMU] - rx * (fe.pf - fw.pf) - rx * (fe.uf * Me - fw.uf * Mw);
	double my = C.xs[IMV] - rx * (Mte - Mtw);
	double E = C.xs[IEN] - rx * (Ee - Ew);
	if (twoD) {
		double Mn, Mtn, En, Ms, Mts, Es;
		fflux(C.fn, fn, P_REF, Mn, Mtn, En); fflux(C.fs, fs, P_REF, Ms, Mts, Es);
		my -= ry * (fn.pf - fs.pf) + ry * (fn.uf * Mn - fs.uf * Ms);
		mx -= ry * (Mtn - Mts);
		E -= ry * (En - Es);
	}
	const double rho = X[2] + X[3];
	if (!(rho > 0)) return false;
	ClIn in; in.atr0 = X[0]; in.atr1 = X[1]; in.mtr0 = X[2]; in.mtr1 = X[3]; in.etr0 = X[4]; in.etr1 = X[5];
	in.s = s; in.rhoeC = E - 0.5 * (mx * mx + my * my) / rho; in.pguess = P_REF + Pc; in.Q0 = C.Q0; in.Q1 = C.Q1;
	ClOut o; if (!closure_p(P, in, o)) return false;
	G = (o.pbar - P_REF) - Pc;
	if (out) { out->a0 = o.a0; out->a1 = o.a1; out->m0 = X[2]; out->m1 = X[3]; out->e0 = o.e0; out->e1 = o.e1; out->mx = mx; out->my = my; out->E = E; out->sig = closure_roundoff(P, in, o.pbar); }
	return isfin(G);
}
struct Nb { int W, E, S, N; };
__device__ __forceinline__ Nb neighbours(const Geo& g, int q) {   // M_P of the four neighbours of interior row q
	const int ii = q % g.nx, jj = q / g.nx; Nb b;
	b.W = ii > 0 ? q - 1 : (g.xper ? q + g.nx - 1 : q);
	b.E = ii < g.nx - 1 ? q + 1 : (g.xper ? q - (g.nx - 1) : q);
	const bool yp = g.yper && g.ny > 1;
	b.S = jj > 0 ? q - g.nx : (yp ? q + (g.ny - 1) * g.nx : q);
	b.N = jj < g.ny - 1 ? q + g.nx : (yp ? q - (g.ny - 1) * g.nx : q);
	return b;
}
// residual of every interior cell at P + lam dP (dP may be null); writes the trial pressure and G, the max-norm
// |G| / (|P_REF + P| + K + 1) (PC:770) and the closure-failure flag
__global__ void __launch_bounds__(256) k_residual(Dev D, const double* __restrict__ P, const double* __restrict__ dP, double lam,
                           double* __restrict__ Pout, double* __restrict__ G, unsigned long long* __restrict__ nrm, int* __restrict__ flags, int flag_slot) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double nv = 0.0;
	if (q < D.g.n) {
		const Nb b = neighbours(D.g, q);
		auto pv = [&](int r) { return dP ? P[r] + lam * dP[r] : P[r]; };
		const double Pc = pv(q), PW = pv(b.W), PE = pv(b.E);
		const double PS = D.g.twoD ? pv(b.S) : 0.0, PN = D.g.twoD ? pv(b.N) : 0.0;
		CellIn C; load_cell(D, q, C);
		double g;
		if (!cell_res(D.ph, D.P_REF, D.rx, D.ry, D.g.twoD, C, Pc, PW, PE, PS, PN, g, nullptr)) { flags[flag_slot] = 1; g = dnan(); }
		else nv = fabs(g) / (fabs(D.P_REF + Pc) + C.Kw + 1.0);
		G[q] = g;
		if (Pout) Pout[q] = Pc;
	}
	atomic_max_nonneg(nrm, nv);
}
// P scale s_q = (|P| + K + 1) / (1 + K D) of the finite-difference steps (PC:783-789)
__global__ void __launch_bounds__(256) k_fdscale(Dev D, const double* __restrict__ P, double* __restrict__ sc) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	const Geo& g = D.g; const int I = q % g.nx + GH, J = q / g.nx + GH, c = I + J * g.NXT;
	const double K = D.Kw[c];
	double Dd = D.rx * (D.kpx[c] + D.kpx[c - 1]);
	if (g.ny > 1) Dd += D.ry * (D.kpy[c] + D.kpy[c - g.NXT]);
	sc[q] = (fabs(D.P_REF + P[q]) + K + 1.0) / (1.0 + K * Dd);
}
// finite-difference Jacobian row: self, W, E (, S, N) perturbed residuals (PC:790-804); J is 5 arrays of n; rhs = -G
__global__ void __launch_bounds__(256) k_jacobian(Dev D, const double* __restrict__ P, const double* __restrict__ G, const double* __restrict__ sc,
                           double* __restrict__ J, double* __restrict__ rhs, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
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
	if (!ok) flags[FLAG_JAC] = 1;
	J[q] = v0; J[n + q] = vW; J[2 * n + q] = vE; J[3 * n + q] = vS; J[4 * n + q] = vN;
}
// Newton step criterion max |dP| / (|P| + s) (PC:819-823)
__global__ void __launch_bounds__(256) k_stepcrit(int n, const double* __restrict__ dP, const double* __restrict__ P, const double* __restrict__ sc, unsigned long long* __restrict__ nrm) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double v = 0.0;
	if (q < n) v = fabs(dP[q]) / (fabs(P[q]) + sc[q]);
	atomic_max_nonneg(nrm, v);
}
// round-off floor acceptance test (PC:834-842): flags a cell whose residual exceeds tol N_i + sigma_i
__global__ void __launch_bounds__(256) k_floor(Dev D, const double* __restrict__ P, double tol, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	const Nb b = neighbours(D.g, q);
	CellIn C; load_cell(D, q, C);
	const bool twoD = D.g.twoD;
	double g; ResOut o;
	if (!cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, C, P[q], P[b.W], P[b.E], twoD ? P[b.S] : 0.0, twoD ? P[b.N] : 0.0, g, &o)) { flags[FLAG_FLOOR] = 1; return; }
	if (fabs(g) > tol * (fabs(D.P_REF + P[q]) + C.Kw + 1.0) + o.sig) flags[FLAG_FLOOR] = 1;
}
// final stage state at the converged pressure (PC:975-987) and the stage derivative Hs = (U_s - Qs) / adt (PC:1276-1284)
__global__ void __launch_bounds__(256) k_final(Dev D, const double* __restrict__ P, const double* __restrict__ Qs, double adt, double* __restrict__ H, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	const Geo& g = D.g; const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	const Nb b = neighbours(g, q);
	CellIn C; load_cell(D, q, C);
	const bool twoD = g.twoD;
	double gg; ResOut o;
	if (!cell_res(D.ph, D.P_REF, D.rx, D.ry, twoD, C, P[q], P[b.W], P[b.E], twoD ? P[b.S] : 0.0, twoD ? P[b.N] : 0.0, gg, &o)) { flags[FLAG_FINAL] = 1; return; }
	const double u[NV] = {o.a0, o.m0, o.m1, o.mx, o.my, o.e0, o.e1, o.E};
	#pragma unroll
	for (int k = 0; k < NV; k++) { D.Ust[k * NC + c] = u[k]; H[k * NC + c] = (u[k] - Qs[k * NC + c]) / adt; }
}
// caldt maxima (PC:477-488): max |u| + c and |v| + c over interior cells with admissible primitives
__global__ void __launch_bounds__(256) k_caldt(Geo g, Phys P, const double* __restrict__ U, int acoustic, unsigned long long* __restrict__ nrm) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double ux = 0.0, uy = 0.0;
	if (q < g.n) {
		const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
		const double a1 = U[IA1 * NC + c], m1 = U[IM1 * NC + c], m2 = U[IM2 * NC + c];
		double cf;
		if (prim_cf(P, a1, m1, m2, U[IE1 * NC + c], U[IE2 * NC + c], cf)) {
			const double rho = m1 + m2, u = U[IMU * NC + c] / rho, v = U[IMV * NC + c] / rho;
			const double cc = acoustic ? cf : 0.0;
			ux = fabs(u) + cc; uy = fabs(v) + cc;
		}
	}
	atomic_max_nonneg(&nrm[0], ux);
	atomic_max_nonneg(&nrm[64], uy);
}

// ------------------------------------------------------------------------------------------------ Thomas (PC:807-812), one thread
__global__ void __launch_bounds__(256) k_thomas(int n, const double* __restrict__ J, const double* __restrict__ rhs, double* __restrict__ x, double* __restrict__ cp, double* __restrict__ dp) {
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	// row r entries: self J[r], W J[n+r] (column r-1, or r at the left boundary), E J[2n+r] (column r+1, or r at the right)
	auto coef = [&](int r, double& lo, double& di, double& up) {
		lo = 0; up = 0; di = 0;
		di += J[r];
		const int cW = r > 0 ? r - 1 : r, cE = r < n - 1 ? r + 1 : r;
		if (cW == r) di += J[n + r]; else lo = J[n + r];
		if (cE == r) di += J[2 * n + r]; else up = J[2 * n + r];
	};
	double lo, di, up;
	coef(0, lo, di, up);
	cp[0] = up / di; dp[0] = rhs[0] / di;
	for (int r = 1; r < n; r++) { coef(r, lo, di, up); const double m = di - lo * cp[r - 1]; cp[r] = up / m; dp[r] = (rhs[r] - lo * dp[r - 1]) / m; }
	x[n - 1] = dp[n - 1];
	for (int r = n - 2; r >= 0; r--) x[r] = dp[r] - cp[r] * x[r + 1];
}

// ------------------------------------------------------------------------------------------------ BiCGStab (PC:725-758)
// Right-preconditioned BiCGStab with Jacobi + constant coarse space, five fused passes per iteration. No atomics and no
// grid fences: each pass writes one partial sum per block, and every block of the next pass re-reduces those partials
// itself in the fixed order (the kernel boundary makes them visible), so all blocks derive the same loop scalars. Scalars
// that outlive a pass (rho, alpha, omega, cc) are kept in KS, double-buffered by the parity of the iteration index the
// host passes in, so no pass reads a value that another block of the same pass writes. Every pass returns at once when
// the solve has finished, so the host launches iterations in chunks and reads KS::done once per chunk.
// Rows: block b, thread t owns rows b*256*items + it*256 + t (it = 0..items-1), the fixed order of the sums.
template <typename T> struct KV { T *x, *r, *r0, *p, *v, *s, *t; const T* J; };
struct KP { double* part; int stride; int nb; int items; };
enum { SL_RR = 0, SL_R0R, SL_SP, SL_R0V, SL_SS, SL_SSUM, SL_TS, SL_TT, SL_E, SL_SB, SL_BB, NSLOTS };

template <typename T>
__device__ __forceinline__ T dinv_of(const T* __restrict__ J, int r) { const T d = J[r]; return d != T(0) ? T(1) / d : T(1); }
// y_r = sum_q A_rq (dinv_c z_c + cc) over the row's entries in the order self, W, E, S, N
template <typename T>
__device__ __forceinline__ T spmv_prec(const Geo& g, const T* __restrict__ J, const T* __restrict__ z, T cc, int r) {
	const int n = g.n; const Nb b = neighbours(g, r);
	auto ph = [&](int c) { return dinv_of(J, c) * z[c] + cc; };
	T s = T(0);
	s += J[r] * ph(r);
	s += J[n + r] * ph(b.W);
	s += J[2 * n + r] * ph(b.E);
	if (g.twoD) { s += J[3 * n + r] * ph(b.S); s += J[4 * n + r] * ph(b.N); }
	return s;
}
template <typename T>
__device__ __forceinline__ T spmv_plain(const Geo& g, const T* __restrict__ J, const T* __restrict__ z, int r) {
	const int n = g.n; const Nb b = neighbours(g, r);
	T s = T(0);
	s += J[r] * z[r];
	s += J[n + r] * z[b.W];
	s += J[2 * n + r] * z[b.E];
	if (g.twoD) { s += J[3 * n + r] * z[b.S]; s += J[4 * n + r] * z[b.N]; }
	return s;
}
#define ROWS_BEGIN for (int it_ = 0; it_ < P.items; it_++) { const int r = blockIdx.x * 256 * P.items + it_ * 256 + threadIdx.x; if (r >= n) continue;
#define ROWS_END }
// this block's partial of one slot (thread 0 writes)
__device__ __forceinline__ void put_partial(const KP& P, int slot, double v, double* sh) {
	const double tot = block_sum256(v, sh);
	if (threadIdx.x == 0) P.part[slot * P.stride + blockIdx.x] = tot;
}
// grid total of one slot written by the previous pass, in the fixed order (identical in every block)
__device__ __forceinline__ double grid_total(const KP& P, int slot, double* sh) {
	const double* q = P.part + slot * P.stride;
	double s = 0.0;
	for (int c = threadIdx.x; c < P.nb; c += 256) s += q[c];
	return block_sum256(s, sh);
}

// setup 1: partials of E = sum of all entries, sum(b), b.b (PC:726-738)
template <typename T>
__global__ void __launch_bounds__(256) kb_setup1(Geo g, const T* __restrict__ J, const T* __restrict__ b, KP P) {
	__shared__ double sh[4];
	const int n = g.n; const int nent = g.twoD ? 5 : 3;
	double sE = 0.0, sb = 0.0, sbb = 0.0;
	ROWS_BEGIN
		T e = T(0); for (int qq = 0; qq < nent; qq++) e += J[qq * n + r];
		sE += (double)e; sb += (double)b[r]; sbb += (double)(b[r] * b[r]);
	ROWS_END
	put_partial(P, SL_E, sE, sh); put_partial(P, SL_SB, sb, sh); put_partial(P, SL_BB, sbb, sh);
}
// setup 2: Einv, cc_b, bn; x = M^-1 b, p = v = 0
template <typename T>
__global__ void __launch_bounds__(256) kb_setup2(Geo g, const T* __restrict__ J, const T* __restrict__ b, KV<T> V, KS* ks, KP P) {
	__shared__ double sh[4];
	const int n = g.n;
	const double E = grid_total(P, SL_E, sh), Sb = grid_total(P, SL_SB, sh), bb = grid_total(P, SL_BB, sh);
	const T Et = (T)E;
	const T Einv = (Et != T(0) && isfin((double)Et)) ? T(1) / Et : T(0);
	const T cc = (T)Sb * Einv;
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		ks->Einv = (double)Einv; ks->bn = (double)((T)sqrt((double)(T)bb) + (T)1e-300);
		ks->rho[0] = 1; ks->alpha[0] = 1; ks->omega[0] = 1;
		ks->done = 0; ks->iters = 0; ks->early = 0; ks->status = 0;
	}
	ROWS_BEGIN V.x[r] = dinv_of(J, r) * b[r] + cc; V.p[r] = T(0); V.v[r] = T(0); ROWS_END
}
// setup 3: r = b - A x, r0 = r; partials of r.r, r0.r
template <typename T>
__global__ void __launch_bounds__(256) kb_setup3(Geo g, const T* __restrict__ b, KV<T> V, KP P) {
	__shared__ double sh[4];
	const int n = g.n;
	double srr = 0.0, sr0r = 0.0;
	ROWS_BEGIN
		const T ax = spmv_plain(g, V.J, (const T*)V.x, r);
		const T rv = b[r] - ax; V.r[r] = rv; V.r0[r] = rv;
		srr += (double)(rv * rv); sr0r += (double)(rv * rv);
	ROWS_END
	put_partial(P, SL_RR, srr, sh); put_partial(P, SL_R0R, sr0r, sh);
}
// I1 (iteration it): loop top (PC:741-744) from r.r and r0.r, then p = r + beta (p - omega v); partials of sum(p)
template <typename T>
__global__ void __launch_bounds__(256) kb_p(Geo g, KV<T> V, KS* ks, KP P, int it, double tol, int maxit) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n, cur = it & 1, nxt = (it + 1) & 1;
	const double rr = grid_total(P, SL_RR, sh), r0r = grid_total(P, SL_R0R, sh);
	const T bn = (T)ks->bn;
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if ((T)sqrt((double)(T)rr) <= (T)tol * bn) { if (lead) { ks->done = 1; ks->status = 0; } return; }
	if (it >= maxit) { if (lead) { ks->done = 1; ks->status = ((T)sqrt((double)(T)rr) <= (T)1e3 * (T)tol * bn) ? 0 : 1; } return; }
	const T rho1 = (T)r0r;
	if (rho1 == T(0)) { if (lead) { ks->done = 1; ks->status = 1; } return; }
	const T beta = (rho1 / (T)ks->rho[cur]) * ((T)ks->alpha[cur] / (T)ks->omega[cur]);
	const T omega = (T)ks->omega[cur];
	if (lead) ks->rho[nxt] = (double)rho1;
	double sp = 0.0;
	ROWS_BEGIN const T pv = V.r[r] + beta * (V.p[r] - omega * V.v[r]); V.p[r] = pv; sp += (double)pv; ROWS_END
	put_partial(P, SL_SP, sp, sh);
}
// I2: cc_p = sum(p) Einv; v = A M^-1 p; partials of r0.v
template <typename T>
__global__ void __launch_bounds__(256) kb_v(Geo g, KV<T> V, KS* ks, KP P, int it) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n;
	const T cc = (T)grid_total(P, SL_SP, sh) * (T)ks->Einv;
	if (blockIdx.x == 0 && threadIdx.x == 0) ks->cc_p = (double)cc;
	double s = 0.0;
	ROWS_BEGIN const T vv = spmv_prec(g, V.J, (const T*)V.p, cc, r); V.v[r] = vv; s += (double)(V.r0[r] * vv); ROWS_END
	put_partial(P, SL_R0V, s, sh);
}
// I3: alpha = rho / r0.v; s = r - alpha v; partials of s.s and sum(s)
template <typename T>
__global__ void __launch_bounds__(256) kb_s(Geo g, KV<T> V, KS* ks, KP P, int it) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n, nxt = (it + 1) & 1;
	const T alpha = (T)ks->rho[nxt] / (T)grid_total(P, SL_R0V, sh);
	if (blockIdx.x == 0 && threadIdx.x == 0) ks->alpha[nxt] = (double)alpha;
	double sss = 0.0, ssum = 0.0;
	ROWS_BEGIN const T sv = V.r[r] - alpha * V.v[r]; V.s[r] = sv; sss += (double)(sv * sv); ssum += (double)sv; ROWS_END
	put_partial(P, SL_SS, sss, sh); put_partial(P, SL_SSUM, ssum, sh);
}
// I4: early exit if |s| <= tol |b| (x += alpha M^-1 p, PC:751); otherwise cc_s = sum(s) Einv, t = A M^-1 s; partials of
// t.s and t.t
template <typename T>
__global__ void __launch_bounds__(256) kb_t(Geo g, KV<T> V, KS* ks, KP P, int it, double tol) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n, nxt = (it + 1) & 1;
	const double s2 = grid_total(P, SL_SS, sh), Ss = grid_total(P, SL_SSUM, sh);
	const T alpha = (T)ks->alpha[nxt];
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if ((T)sqrt((double)(T)s2) <= (T)tol * (T)ks->bn) {
		const T ccp = (T)ks->cc_p;
		ROWS_BEGIN V.x[r] += alpha * (dinv_of(V.J, r) * V.p[r] + ccp); ROWS_END
		if (lead) { ks->early = 1; ks->iters = it + 1; ks->status = 0; }
		return;
	}
	const T cc = (T)Ss * (T)ks->Einv;
	if (lead) ks->cc_s = (double)cc;
	double sts = 0.0, stt = 0.0;
	ROWS_BEGIN const T tv = spmv_prec(g, V.J, (const T*)V.s, cc, r); V.t[r] = tv; sts += (double)(tv * V.s[r]); stt += (double)(tv * tv); ROWS_END
	put_partial(P, SL_TS, sts, sh); put_partial(P, SL_TT, stt, sh);
}
// I5: omega = t.s / t.t; x += alpha M^-1 p + omega M^-1 s; r = s - omega t; partials of r.r and r0.r
template <typename T>
__global__ void __launch_bounds__(256) kb_xr(Geo g, KV<T> V, KS* ks, KP P, int it) {
	if (ks->done) return;
	const int nxt = (it + 1) & 1;
	if (ks->early) { if (blockIdx.x == 0 && threadIdx.x == 0) ks->done = 1; return; }
	__shared__ double sh[4];
	const int n = g.n;
	const double ts = grid_total(P, SL_TS, sh), tt = grid_total(P, SL_TT, sh);
	const T omega = (T)ts / (T)tt;
	const T alpha = (T)ks->alpha[nxt], ccp = (T)ks->cc_p, ccs = (T)ks->cc_s;
	if (blockIdx.x == 0 && threadIdx.x == 0) { ks->omega[nxt] = (double)omega; ks->iters = it + 1; }
	double srr = 0.0, sr0r = 0.0;
	ROWS_BEGIN
		const T di = dinv_of(V.J, r);
		const T ph = di * V.p[r] + ccp, shv = di * V.s[r] + ccs;
		V.x[r] += alpha * ph + omega * shv;
		const T rv = V.s[r] - omega * V.t[r]; V.r[r] = rv;
		srr += (double)(rv * rv); sr0r += (double)(V.r0[r] * rv);
	ROWS_END
	put_partial(P, SL_RR, srr, sh); put_partial(P, SL_R0R, sr0r, sh);
}

// FP64 defect of the linear system r = b - J x, with partials of r.r (slot SL_RR) and b.b (slot SL_BB) per block
__global__ void __launch_bounds__(256) k_linres(Geo g, const double* __restrict__ J, const double* __restrict__ b, const double* __restrict__ x,
                                               double* __restrict__ r, double* __restrict__ part, int stride) {
	__shared__ double sh[4];
	const int n = g.n;
	const int q = blockIdx.x * 256 + threadIdx.x;
	double rr = 0.0, bb = 0.0;
	if (q < n) { const double v = b[q] - spmv_plain(g, J, x, q); r[q] = v; rr = v * v; bb = b[q] * b[q]; }
	const double t1 = block_sum256(rr, sh), t2 = block_sum256(bb, sh);
	if (threadIdx.x == 0) { part[blockIdx.x] = t1; part[stride + blockIdx.x] = t2; }
}
// x += (double) d
__global__ void __launch_bounds__(256) k_add_f32(long n, const float* __restrict__ d, double* __restrict__ x) {
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) x[i] += (double)d[i];
}
void launch_linres(const Geo& g, const double* J, const double* b, const double* x, double* r, double* part, int stride, hipStream_t st) {
	hipLaunchKernelGGL(k_linres, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, st, g, J, b, x, r, part, stride);
}
void launch_add_f32(long n, const float* d, double* x, hipStream_t st) { hipLaunchKernelGGL(k_add_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, d, x); }

// precision conversions around the FP32 Krylov solve (the Newton system stays FP64)
__global__ void __launch_bounds__(256) k_to_f32(long n, const double* __restrict__ a, float* __restrict__ b) {
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) b[i] = (float)a[i];
}
__global__ void __launch_bounds__(256) k_to_f64(long n, const float* __restrict__ a, double* __restrict__ b) {
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) b[i] = (double)a[i];
}
void launch_to_f32(long n, const double* a, float* b, hipStream_t st) { hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b); }
void launch_to_f64(long n, const float* a, double* b, hipStream_t st) { hipLaunchKernelGGL(k_to_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b); }

// rows per thread of the Krylov passes: about a thousand blocks whatever n (4 .. 32 rows per thread)
int krylov_items(int n) {
	int it = 4;
	while (it < 32 && (long)n > (long)256 * it * 2048) it *= 2;
	return it;
}

template <typename T>
static void krylov_setup_t(const Geo& g, const KrylovBufsT<T>& B, const T* b, hipStream_t st) {
	const int items = krylov_items(g.n), nb = (g.n + 256 * items - 1) / (256 * items);
	KP P{B.part, B.stride, nb, items};
	KV<T> V{B.x, B.r, B.r0, B.p, B.v, B.s, B.t, B.J};
	hipLaunchKernelGGL(kb_setup1<T>, dim3(nb), dim3(256), 0, st, g, B.J, b, P);
	hipLaunchKernelGGL(kb_setup2<T>, dim3(nb), dim3(256), 0, st, g, B.J, b, V, B.ks, P);
	hipLaunchKernelGGL(kb_setup3<T>, dim3(nb), dim3(256), 0, st, g, b, V, P);
}
template <typename T>
static void krylov_iter_t(const Geo& g, const KrylovBufsT<T>& B, hipStream_t st, int it, double tol, int maxit) {
	const int items = krylov_items(g.n), nb = (g.n + 256 * items - 1) / (256 * items);
	KP P{B.part, B.stride, nb, items};
	KV<T> V{B.x, B.r, B.r0, B.p, B.v, B.s, B.t, B.J};
	hipLaunchKernelGGL(kb_p<T>, dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it, tol, maxit);
	hipLaunchKernelGGL(kb_v<T>, dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
	hipLaunchKernelGGL(kb_s<T>, dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
	hipLaunchKernelGGL(kb_t<T>, dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it, tol);
	hipLaunchKernelGGL(kb_xr<T>, dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
}
void launch_krylov_setup(const Geo& g, const KrylovBufs& B, const double* b, hipStream_t st) { krylov_setup_t<double>(g, B, b, st); }
void launch_krylov_iter(const Geo& g, const KrylovBufs& B, hipStream_t st, int it, double tol, int maxit) { krylov_iter_t<double>(g, B, st, it, tol, maxit); }
void launch_krylov_setup32(const Geo& g, const KrylovBufs32& B, const float* b, hipStream_t st) { krylov_setup_t<float>(g, B, b, st); }
void launch_krylov_iter32(const Geo& g, const KrylovBufs32& B, hipStream_t st, int it, double tol, int maxit) { krylov_iter_t<float>(g, B, st, it, tol, maxit); }

// ------------------------------------------------------------------------------------------------ host launchers
static inline int nblk(long n, int b = 256) { return (int)((n + b - 1) / b); }

void launch_bc_grid(const Geo& g, double* u, hipStream_t st) {
	hipLaunchKernelGGL(k_bc_x, dim3(nblk(g.NYT * GH * 2)), dim3(256), 0, st, g, u);
	hipLaunchKernelGGL(k_bc_y, dim3(nblk(g.NXT * GH * 2)), dim3(256), 0, st, g, u);
}
void launch_bc_array(const Geo& g, double* a, hipStream_t st) {
	hipLaunchKernelGGL(k_bcarr_x, dim3(nblk(g.NYT * GH)), dim3(256), 0, st, g, a);
	hipLaunchKernelGGL(k_bcarr_y, dim3(nblk(g.NXT * GH)), dim3(256), 0, st, g, a);
}
void launch_combine(const Geo& g, const double* Un, const double* H0, const double* H1, int nr, double ae0, double ae1, double ai0, double ai1, double* Qe, double* Qs, hipStream_t st) {
	hipLaunchKernelGGL(k_combine, dim3(nblk(g.n)), dim3(256), 0, st, g, Un, H0, H1, nr, ae0, ae1, ai0, ai1, Qe, Qs);
}
void launch_explicit(const Dev& D, int limiter, int order, const double* Qe, const double* Qs, double adt, hipStream_t st) {
	const Geo& g = D.g;
	hipLaunchKernelGGL(k_scales, dim3(nblk(g.NC)), dim3(256), 0, st, g, D.ph, Qe, D.vsc, D.esc, D.flags);
	hipLaunchKernelGGL(k_recon, dim3(nblk(g.NC)), dim3(256), 0, st, g, D.ph, limiter, order, Qe, D.vsc, D.esc, D.wxL, D.wxR, D.wyL, D.wyR, D.phx, D.phy);
	hipLaunchKernelGGL(k_update, dim3(nblk(g.n)), dim3(256), 0, st, g, adt / g.dx, adt / g.dy, Qe, Qs, D.wxL, D.wxR, D.wyL, D.wyR, D.Ust, D.a2s, D.ubx, D.uby);
	launch_bc_grid(g, D.Ust, st);
	launch_bc_array(g, D.a2s, st);
}
void launch_prep(const Dev& D, int limiter, int order, double tau, const double* Pst, hipStream_t st) {
	const Geo& g = D.g;
	hipLaunchKernelGGL(k_prep_cell, dim3(nblk(g.NC)), dim3(256), 0, st, g, D.ph, D.Ust, D.a2s, Pst, D.ust, D.vst, D.cz, D.Kw, D.Pold, D.flags);
	hipLaunchKernelGGL(k_prep_face, dim3(nblk(g.NC)), dim3(256), 0, st, g, order, limiter, tau, D.Ust, D.ust, D.vst, D.cz, D.Pold, D.phx, D.phy,
	                   D.ubsx, D.kpx, D.wLx, D.pvx, D.ubsy, D.kpy, D.wLy, D.pvy);
}
void launch_shock(const Dev& D, hipStream_t st) {
	hipLaunchKernelGGL(k_shock, dim3(nblk(D.g.n, 64)), dim3(64), 0, st, D.g, D.ph, D.Ust, D.Qph);
}
void launch_residual(const Dev& D, const double* P, const double* dP, double lam, double* Pout, double* G, unsigned long long* nrm, int flag_slot, hipStream_t st) {
	hipLaunchKernelGGL(k_residual, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, dP, lam, Pout, G, nrm, D.flags, flag_slot);
}
void launch_jacobian(const Dev& D, const double* P, const double* G, double* sc, double* J, double* rhs, hipStream_t st) {
	hipLaunchKernelGGL(k_fdscale, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, sc);
	hipLaunchKernelGGL(k_jacobian, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, G, (const double*)sc, J, rhs, D.flags);
}
void launch_fdscale(const Dev& D, const double* P, double* sc, hipStream_t st) {
	hipLaunchKernelGGL(k_fdscale, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, sc);
}
void launch_stepcrit(int n, const double* dP, const double* P, const double* sc, unsigned long long* nrm, hipStream_t st) {
	hipLaunchKernelGGL(k_stepcrit, dim3(nblk(n)), dim3(256), 0, st, n, dP, P, sc, nrm);
}
void launch_floor(const Dev& D, const double* P, double tol, hipStream_t st) {
	hipLaunchKernelGGL(k_floor, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, tol, D.flags);
}
void launch_final(const Dev& D, const double* P, const double* Qs, double adt, double* H, hipStream_t st) {
	hipLaunchKernelGGL(k_final, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, Qs, adt, H, D.flags);
}
void launch_caldt(const Geo& g, const Phys& P, const double* U, int acoustic, unsigned long long* nrm, hipStream_t st) {
	hipLaunchKernelGGL(k_caldt, dim3(nblk(g.n)), dim3(256), 0, st, g, P, U, acoustic, nrm);
}
void launch_thomas(int n, const double* J, const double* rhs, double* x, double* cp, double* dp, hipStream_t st) {
	hipLaunchKernelGGL(k_thomas, dim3(1), dim3(64), 0, st, n, J, rhs, x, cp, dp);
}
}   // namespace spb
