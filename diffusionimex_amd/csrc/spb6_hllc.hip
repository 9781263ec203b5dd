// spb6_hllc.hip -- gfx950 kernels of the reference's fully explicit scheme (src/SPB6_explicit.cpp, "EX") at relaxation
// level 0: minmod reconstruction of the primitive variables, HLLC face fluxes with the phase Hugoniot states in the
// star region, the Godunov update with the non-conservative alpha_1 and phase-energy terms, and the per-cell mechanical
// relaxation by the p closure with the shock budget's defect split. It runs on the stepper's ghosted structure-of-arrays
// grids; the host side (spb6_api.hip) strings the passes into Heun's scheme and the reject-and-halve time loop.
//
// One thread per ghosted cell (primitives), per face (fluxes) or per interior cell (update, relaxation). The face pass
// recomputes the two face states of its face from the cell primitives of a 4-cell stencil, which is the reference's
// per-cell reconstruction (each face state depends only on its cell's 3-cell stencil), so the arithmetic is the
// reference's operation for operation and the results equal the oracle bit for bit.
//
// Every pass returns at once once a failure flag of the attempt is set (FLAG_STAGE: a primitive state or an updated
// state out of range; FLAG_FINAL: a closure failure): the reference stops the attempt at the first failure
// (EX:363-372, EX:546-556), so later passes must not count closure failures it never reached.
#include "spb6_device.h"
#include "spb6_kernels.h"
#include "spb6_cell.h"

namespace spb {

namespace {

// cell primitives of prim_of (PHY:837-847) as nine arrays over the ghosted grid: a1 r1 r2 u v p1 p2 rho cf
enum { HP_A1 = 0, HP_R1, HP_R2, HP_U, HP_V, HP_P1, HP_P2, HP_RHO, HP_CF, HP_N };
// HLLC face values (EX:145): F (m1, m2, normal momentum, tangential momentum, E), u*, (u alpha1)*, (eps_k u)*
enum { HF_F0 = 0, HF_F1, HF_F2, HF_F3, HF_F4, HF_US, HF_UA, HF_UE0, HF_UE1, HF_N };

struct HPrim { double a1, r1, r2, u, v, p1, p2, rho, cf; };

__device__ __forceinline__ bool failed(const int* __restrict__ flags) { return (flags[FLAG_STAGE] | flags[FLAG_FINAL]) != 0; }

__device__ __forceinline__ HPrim load_prim(const double* __restrict__ W, int NC, int c) {
	HPrim w;
	w.a1 = W[HP_A1 * NC + c]; w.r1 = W[HP_R1 * NC + c]; w.r2 = W[HP_R2 * NC + c]; w.u = W[HP_U * NC + c]; w.v = W[HP_V * NC + c];
	w.p1 = W[HP_P1 * NC + c]; w.p2 = W[HP_P2 * NC + c]; w.rho = W[HP_RHO * NC + c]; w.cf = W[HP_CF * NC + c];
	return w;
}

// slope limiter of the primitive reconstruction (EX:238): minmod, or the central slope without a limiter
__device__ __forceinline__ double hx_mm(double a, double b, int limiter) {
	if (!limiter) return 0.5 * (a + b);
	return a * b > 0 ? (fabs(a) < fabs(b) ? a : b) : 0.0;
}
// face state of cell c on its low (side 0) or high (side 1) face from the cells a (below c) and b (above c) (EX:241-253):
// the seven primitives moved by half the limited slope, then rho and the frozen sound speed from them
__device__ __forceinline__ HPrim face_state(const Phys& P, const HPrim& a, const HPrim& c, const HPrim& b, int side, int limiter) {
	const double cv[7] = {c.a1, c.r1, c.r2, c.u, c.v, c.p1, c.p2}, av[7] = {a.a1, a.r1, a.r2, a.u, a.v, a.p1, a.p2};
	const double bv[7] = {b.a1, b.r1, b.r2, b.u, b.v, b.p1, b.p2};
	double f[7];
	#pragma unroll
	for (int q = 0; q < 7; q++) {
		const double sl = hx_mm(cv[q] - av[q], bv[q] - cv[q], limiter);
		f[q] = side == 0 ? cv[q] - 0.5 * sl : cv[q] + 0.5 * sl;
	}
	HPrim w;
	w.a1 = f[0]; w.r1 = f[1]; w.r2 = f[2]; w.u = f[3]; w.v = f[4]; w.p1 = f[5]; w.p2 = f[6];
	const double m1 = w.a1 * w.r1, m2 = (1 - w.a1) * w.r2;
	w.rho = m1 + m2;
	w.cf = sqrt((m1 * c2_rp(P, 0, w.r1, w.p1) + m2 * c2_rp(P, 1, w.r2, w.p2)) / w.rho);
	return w;
}

// phase Hugoniot pressure behind a shock from volume v to vs (EX:147-157, SG / NASG)
__device__ __forceinline__ double hugoniot_p(const Phys& P, int k, double p, double v, double vs) {
	const double g = P.g[k], pi = P.pi[k], b = P.b[k];
	const double num = (p + g * pi) * (v - b) - g * pi * (vs - b) - 0.5 * (g - 1.0) * p * (vs - v);
	const double den = (vs - b) + 0.5 * (g - 1.0) * (vs - v);
	return num / den;
}
__device__ __forceinline__ double total_energy(const Phys& P, const HPrim& w) {
	const double a2 = 1 - w.a1;
	return eps_amp(P, 0, w.a1, w.a1 * w.r1, w.p1) + eps_amp(P, 1, a2, a2 * w.r2, w.p2) + 0.5 * w.rho * (w.u * w.u + w.v * w.v);
}
// HLLC flux of the face between the states L and R, normal direction dir (EX:159-192)
__device__ __forceinline__ void hllc_face(const Phys& P, const HPrim& L, const HPrim& R, int dir, double f[HF_N]) {
	const double uL = dir == 0 ? L.u : L.v, uR = dir == 0 ? R.u : R.v;
	const double pL = L.a1 * L.p1 + (1 - L.a1) * L.p2, pR = R.a1 * R.p1 + (1 - R.a1) * R.p2;
	const double SL = smin(uL - L.cf, uR - R.cf), SR = smax(uL + L.cf, uR + R.cf);
	const double SM = (pR - pL + L.rho * uL * (SL - uL) - R.rho * uR * (SR - uR)) / (L.rho * (SL - uL) - R.rho * (SR - uR));
	const bool left = SM >= 0;
	const HPrim& K = left ? L : R;
	const double uK = left ? uL : uR, SK = left ? SL : SR, pK = left ? pL : pR, tK = dir == 0 ? K.v : K.u;
	const double EKv = total_energy(P, K);
	double FK[5];
	{
		const double p = K.a1 * K.p1 + (1 - K.a1) * K.p2;
		FK[0] = K.a1 * K.r1 * uK; FK[1] = (1 - K.a1) * K.r2 * uK; FK[2] = K.rho * uK * uK + p; FK[3] = K.rho * uK * tK;
		FK[4] = (EKv + p) * uK;
	}
	if ((left && SL >= 0) || (!left && SR <= 0)) {
		#pragma unroll
		for (int q = 0; q < 5; q++) f[q] = FK[q];
		f[HF_US] = uK; f[HF_UA] = uK * K.a1;
		f[HF_UE0] = eps_amp(P, 0, K.a1, K.a1 * K.r1, K.p1) * uK;
		f[HF_UE1] = eps_amp(P, 1, 1 - K.a1, (1 - K.a1) * K.r2, K.p2) * uK;
		return;
	}
	const double fac = (SK - uK) / (SK - SM);
	const double rs = K.rho * fac;
	const double Es = fac * (EKv + (SM - uK) * (K.rho * SM + pK / (SK - uK)));
	const double UK[5] = {K.a1 * K.r1, (1 - K.a1) * K.r2, K.rho * uK, K.rho * tK, EKv};
	const double Us[5] = {UK[0] * fac, UK[1] * fac, rs * SM, rs * tK, Es};
	#pragma unroll
	for (int q = 0; q < 5; q++) f[q] = FK[q] + SK * (Us[q] - UK[q]);
	f[HF_US] = SM; f[HF_UA] = SM * K.a1;
	const double ak[2] = {K.a1, 1 - K.a1}, rk[2] = {K.r1, K.r2}, pk[2] = {K.p1, K.p2};
	#pragma unroll
	for (int k = 0; k < 2; k++) {
		const double v0 = 1.0 / rk[k], vs = 1.0 / (rk[k] * fac);
		double ps = hugoniot_p(P, k, pk[k], v0, vs);
		if (!isfin(ps) || !p_adm_sg(P, k, ps)) ps = pk[k];
		f[HF_UE0 + k] = eps_amp(P, k, ak[k], ak[k] * rk[k] * fac, ps) * SM;
	}
}

// ---------------------------------------------------------------------------------------------------------- kernels
// primitives of every ghosted cell (EX:231-232); a cell prim_of rejects fails the attempt
__global__ void __launch_bounds__(256) k_hx_prim(Geo g, Phys P, const double* __restrict__ U, double* __restrict__ W, int* __restrict__ flags) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC || failed(flags)) return;
	const int NC = g.NC;
	const double a1 = U[IA1 * NC + c], a2 = 1 - a1, m1 = U[IM1 * NC + c], m2 = U[IM2 * NC + c];
	if (!(a1 > 0 && a2 > 0 && m1 > 0 && m2 > 0)) { flags[FLAG_STAGE] = 1; return; }
	const double rho = m1 + m2, r1 = m1 / a1, r2 = m2 / a2, u = U[IMU * NC + c] / rho, v = U[IMV * NC + c] / rho;
	const double p1 = p_ame(P, 0, a1, m1, U[IE1 * NC + c]), p2 = p_ame(P, 1, a2, m2, U[IE2 * NC + c]);
	if (!(p_adm_sg(P, 0, p1) && p_adm_sg(P, 1, p2))) { flags[FLAG_STAGE] = 1; return; }
	const double cf = sqrt((m1 * c2_rp(P, 0, r1, p1) + m2 * c2_rp(P, 1, r2, p2)) / rho);
	W[HP_A1 * NC + c] = a1; W[HP_R1 * NC + c] = r1; W[HP_R2 * NC + c] = r2; W[HP_U * NC + c] = u; W[HP_V * NC + c] = v;
	W[HP_P1 * NC + c] = p1; W[HP_P2 * NC + c] = p2; W[HP_RHO * NC + c] = rho; W[HP_CF * NC + c] = cf;
}
// HLLC flux of every face the update reads (EX:196-202), stored at the face's low cell: x faces (i + 1/2, j) for
// i = GH-1 .. nx+GH-1 on the interior rows, y faces (i, j + 1/2) for j = GH-1 .. ny+GH-1 on the interior columns
template <int DIR>
__global__ void __launch_bounds__(256) k_hx_flux(Geo g, Phys P, const double* __restrict__ W, int order, int limiter, double* __restrict__ F,
                                                 const int* __restrict__ flags) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int nfi = DIR == 0 ? g.nx + 1 : g.nx, nfj = DIR == 0 ? g.ny : g.ny + 1;
	if (t >= nfi * nfj || failed(flags)) return;
	const int i = t % nfi + (DIR == 0 ? GH - 1 : GH), j = t / nfi + (DIR == 0 ? GH : GH - 1);
	const int NC = g.NC, s = DIR == 0 ? 1 : g.NXT, c = i + j * g.NXT;
	const HPrim wc = load_prim(W, NC, c), wn = load_prim(W, NC, c + s);
	HPrim L = wc, R = wn;
	if (order >= 2) {
		const HPrim wb = load_prim(W, NC, c - s), wa = load_prim(W, NC, c + 2 * s);
		L = face_state(P, wb, wc, wn, 1, limiter);
		R = face_state(P, wc, wn, wa, 0, limiter);
	}
	double f[HF_N];
	hllc_face(P, L, R, DIR, f);
	#pragma unroll
	for (int q = 0; q < HF_N; q++) F[q * NC + c] = f[q];
}
// Godunov update of the interior cells (EX:203-224): Uout = U + dU; alpha_1 out of (0, 1) or a non-positive phase mass
// fails the attempt
__global__ void __launch_bounds__(256) k_hx_update(Geo g, const double* __restrict__ U, const double* __restrict__ W, const double* __restrict__ FX,
                                                   const double* __restrict__ FY, double rx, double ry, double* __restrict__ Uo, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n || failed(flags)) return;
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	const double a1 = W[HP_A1 * NC + c], ak[2] = {a1, 1 - a1}, pk[2] = {W[HP_P1 * NC + c], W[HP_P2 * NC + c]};
	auto fx = [&](int k, int cc) { return FX[k * NC + cc]; };
	double dU[NV];
	const double dus = fx(HF_US, c) - fx(HF_US, c - 1);
	dU[IM1] = -rx * (fx(HF_F0, c) - fx(HF_F0, c - 1)); dU[IM2] = -rx * (fx(HF_F1, c) - fx(HF_F1, c - 1));
	dU[IMU] = -rx * (fx(HF_F2, c) - fx(HF_F2, c - 1)); dU[IMV] = -rx * (fx(HF_F3, c) - fx(HF_F3, c - 1));
	dU[IEN] = -rx * (fx(HF_F4, c) - fx(HF_F4, c - 1));
	dU[IA1] = -rx * ((fx(HF_UA, c) - fx(HF_UA, c - 1)) - a1 * dus);
	#pragma unroll
	for (int k = 0; k < 2; k++) dU[IE1 + k] = -rx * ((fx(HF_UE0 + k, c) - fx(HF_UE0 + k, c - 1)) + ak[k] * pk[k] * dus);
	if (g.twoD) {
		const int cs = c - g.NXT;
		auto fy = [&](int k, int cc) { return FY[k * NC + cc]; };
		const double dvs = fy(HF_US, c) - fy(HF_US, cs);
		dU[IM1] -= ry * (fy(HF_F0, c) - fy(HF_F0, cs)); dU[IM2] -= ry * (fy(HF_F1, c) - fy(HF_F1, cs));
		dU[IMV] -= ry * (fy(HF_F2, c) - fy(HF_F2, cs)); dU[IMU] -= ry * (fy(HF_F3, c) - fy(HF_F3, cs));
		dU[IEN] -= ry * (fy(HF_F4, c) - fy(HF_F4, cs));
		dU[IA1] -= ry * ((fy(HF_UA, c) - fy(HF_UA, cs)) - a1 * dvs);
		#pragma unroll
		for (int k = 0; k < 2; k++) dU[IE1 + k] -= ry * ((fy(HF_UE0 + k, c) - fy(HF_UE0 + k, cs)) + ak[k] * pk[k] * dvs);
	}
	double Q[NV];
	#pragma unroll
	for (int k = 0; k < NV; k++) { Q[k] = U[k * NC + c] + dU[k]; Uo[k * NC + c] = Q[k]; }
	if (!(Q[IA1] > 0 && Q[IA1] < 1 && Q[IM1] > 0 && Q[IM2] > 0)) flags[FLAG_STAGE] = 1;
}
// mechanical relaxation of the interior cells (EX:305-343, relaxation level 0): a phase whose transported state is
// inadmissible restarts from the mixture pressure, then the p closure (s = 0, the cell's shock budget Qph) sets the
// volume fractions and phase energies; a closure failure fails the attempt and is counted (*fails)
__global__ void __launch_bounds__(256) k_hx_relax(Geo g, Phys P, double* __restrict__ U, const double* __restrict__ Qph, int* __restrict__ flags,
                                                  int* __restrict__ fails) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n || failed(flags)) return;
	const int NC = g.NC, c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	const double a1 = U[IA1 * NC + c], m1 = U[IM1 * NC + c], m2 = U[IM2 * NC + c], mu = U[IMU * NC + c], mv = U[IMV * NC + c];
	const double E = U[IEN * NC + c], e1o = U[IE1 * NC + c], e2o = U[IE2 * NC + c];
	const double rho = m1 + m2, rhoe = E - 0.5 * (mu * mu + mv * mv) / rho;
	double ek[2] = {e1o, e2o};
	const double ak[2] = {a1, 1 - a1}, mk[2] = {m1, m2};
	#pragma unroll
	for (int k = 0; k < 2; k++)
		if (!p_adm_sg(P, k, p_ame(P, k, ak[k], mk[k], ek[k]))) {
			const double pm = p_mix(P, a1, m1, m2, rhoe);
			if (p_adm_sg(P, k, pm)) ek[k] = eps_amp(P, k, ak[k], mk[k], pm);
		}
	const double pbar = a1 * p_ame(P, 0, a1, m1, e1o) + (1 - a1) * p_ame(P, 1, 1 - a1, m2, e2o);
	ClIn in;
	in.atr0 = a1; in.atr1 = 1 - a1; in.mtr0 = m1; in.mtr1 = m2; in.etr0 = ek[0]; in.etr1 = ek[1];
	in.s = 0.0; in.rhoeC = rhoe; in.pguess = smax(pbar, 1.0); in.Q0 = Qph[c]; in.Q1 = Qph[NC + c];
	ClOut o;
	if (!closure_p(P, in, o)) { flags[FLAG_FINAL] = 1; atomicAdd(fails, 1); return; }
	U[IA1 * NC + c] = o.a0; U[IE1 * NC + c] = o.e0; U[IE2 * NC + c] = o.e1;
}
// Heun's average (EX:552) over the whole ghosted grid: V = (V + W) / 2
__global__ void __launch_bounds__(256) k_hx_average(long n, double* __restrict__ V, const double* __restrict__ W, const int* __restrict__ flags) {
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n || failed(flags)) return;
	V[t] = 0.5 * (V[t] + W[t]);
}

inline int nb(long n) { return (int)((n + 255) / 256); }

}   // namespace

int hllc_prim_fields() { return HP_N; }
int hllc_face_fields() { return HF_N; }

void launch_hllc_prim(const Geo& g, const Phys& P, const double* U, double* W, int* flags, hipStream_t st) {
	hipLaunchKernelGGL(k_hx_prim, dim3(nb(g.NC)), dim3(256), 0, st, g, P, U, W, flags);
}
void launch_hllc_flux(const Geo& g, const Phys& P, const double* W, int order, int limiter, double* FX, double* FY, const int* flags, hipStream_t st) {
	hipLaunchKernelGGL(k_hx_flux<0>, dim3(nb((long)(g.nx + 1) * g.ny)), dim3(256), 0, st, g, P, W, order, limiter, FX, flags);
	if (g.twoD) hipLaunchKernelGGL(k_hx_flux<1>, dim3(nb((long)g.nx * (g.ny + 1))), dim3(256), 0, st, g, P, W, order, limiter, FY, flags);
}
void launch_hllc_update(const Geo& g, const double* U, const double* W, const double* FX, const double* FY, double rx, double ry, double* Uo,
                        int* flags, hipStream_t st) {
	hipLaunchKernelGGL(k_hx_update, dim3(nb(g.n)), dim3(256), 0, st, g, U, W, FX, FY, rx, ry, Uo, flags);
}
void launch_hllc_relax(const Geo& g, const Phys& P, double* U, const double* Qph, int* flags, int* fails, hipStream_t st) {
	hipLaunchKernelGGL(k_hx_relax, dim3(nb(g.n)), dim3(256), 0, st, g, P, U, Qph, flags, fails);
}
void launch_hllc_average(long n, double* V, const double* W, const int* flags, hipStream_t st) {
	hipLaunchKernelGGL(k_hx_average, dim3(nb(n)), dim3(256), 0, st, n, V, W, flags);
}

}   // namespace spb
