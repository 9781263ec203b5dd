// spb6_kernels.hip -- gfx950 kernels of the SPB 6+1 semi-implicit IMEX-RK step (FP64, structure of arrays).
//
// Device layout: every cell field is its own array over the ghosted grid, index c = i + j*NXT (NXT = nx+4, NYT = ny+4,
// two ghost layers); conserved grids hold 8 such arrays back to back (field k at base + k*NC). Pressure-system vectors
// (P, G, Krylov vectors, Jacobian diagonals) are interior-only, row r = ii + jj*nx. One thread per cell or face row;
// 256-thread blocks (4 wavefronts of 64). Reductions use one fixed tree order (see block_sum256 / grid_total), so every
// sum is bit-reproducible and equal to the CPU oracle's "device order".
#include "spb6_device.h"
#include "spb6_kernels.h"
#include "spb6_cell.h"
#include <algorithm>
#include <type_traits>

namespace spb {

// ------------------------------------------------------------------------------------------------ workgroup order
// Workgroups go to the 8 XCDs round robin (workgroup b to XCD b mod 8). xcd_band(b, nb) is a bijection of [0, nb) that
// gives XCD x the contiguous band x q + min(x, m) ... (q = nb / 8, m = nb mod 8) in dispatch order, so that blocks
// working on neighbouring parts of the grid share their XCD's L2 for the rows they read beyond their own.
__device__ __forceinline__ int xcd_band(int b, int nb) {
	const int x = b & 7, q = nb >> 3, m = nb & 7;
	return x * q + (x < m ? x : m) + (b >> 3);
}

// ------------------------------------------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_bfly_sum(double v) {
	#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
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

// ------------------------------------------------------------------------------------------------ boundaries (PHY:926-957)
// x pass over all rows (ghost columns), then y pass over all columns; the second pass reads the corners the first wrote
__global__ void __launch_bounds__(256) k_bc_x(Geo g, double* u) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;   // (row j, layer q, side)
	if (t >= g.NYT * GH * 2) return;
	bc_x_at(g, u, (t >> 1) / GH, (t >> 1) % GH, t & 1);
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
	bcarr_x_at(g, a, t / GH, t % GH);
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
// (and the cell velocities into vu, vv)
__global__ void __launch_bounds__(256) k_scales(Geo g, Phys P, const double* __restrict__ Q, double* __restrict__ vsc, double* __restrict__ esc,
                                                double* __restrict__ vu, double* __restrict__ vv, int* __restrict__ flags) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	if (!scales_at(g, P, Q, vsc, esc, c, vu, vv)) flags[FLAG_RHO] = 1;
}
__global__ void __launch_bounds__(256) k_recon(Geo g, Phys P, int limiter, int order, const double* __restrict__ Q, const double* __restrict__ vsc,
                        const double* __restrict__ esc, double* __restrict__ wxL, double* __restrict__ wxR,
                        double* __restrict__ wyL, double* __restrict__ wyR, double* __restrict__ phx, double* __restrict__ phy) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const int NC = g.NC, j = c / g.NXT;
	recon_x_at(g, P, limiter, order, Q, vsc, esc, wxL, wxR, phx, c);   // x direction (all rows)
	if (!g.twoD) return;
	double L[NV], R[NV];
	if (order >= 2 && j >= 1 && j + 1 < g.NYT) {
		phy[c] = recon_cell(g, P, limiter, Q, vsc, esc, c - g.NXT, c, c + g.NXT, L, R);
		for (int k = 0; k < NV; k++) { wyL[k * NC + c] = L[k]; wyR[k * NC + c] = R[k]; }
	} else {
		phy[c] = 0.0;
		for (int k = 0; k < NV; k++) { const double v = Q[k * NC + c]; wyL[k * NC + c] = v; wyR[k * NC + c] = v; }
	}
}
// TVDReconstruction + Update fused (PC:349-446, PC:251-326) on 1D grids: TX cells of one grid row per block, one cell
// per thread. Each thread reconstructs its cell; the left face states go through LDS to the west neighbour, so each
// face's Rusanov flux is computed once (by the cell west of it; the face west of the block by thread 64, which
// reconstructs the cell west of the block, thread 65 the cell east of it for its left state) and handed back through
// LDS. Face states never reach HBM: the kernel writes the limiter factors, the face donors the implicit stage reads
// (put_donors), X*, alpha2* and the face velocities, each by the operations of recon_x_at / update_at.
template <int TX>
__global__ void __launch_bounds__(TX) k_explicit_1d(Geo g, Phys P, int limiter, int order, double rxd,
		const double* __restrict__ Qe, const double* __restrict__ Qs, const double* __restrict__ vsc, const double* __restrict__ esc,
		const double* __restrict__ vu, const double* __restrict__ vv, double* __restrict__ donx,
		double* __restrict__ phx, double* __restrict__ Ust, double* __restrict__ a2s, double* __restrict__ ubx, int tiles_y) {
	constexpr int HX = 64;
	static_assert(HX + 2 <= TX, "halo threads");
	__shared__ double sh[std::max(NV * (TX + 2), (NV + 1) * (TX + 1))];
	const int NXT = g.NXT, NYT = g.NYT, NC = g.NC, tx = threadIdx.x;
	const int tile = xcd_band(blockIdx.x, gridDim.x), i0 = (tile / tiles_y) * TX, j = tile % tiles_y;
	const int i = i0 + tx, c = i + j * NXT;
	const bool valid = i < NXT && j < NYT;
	const bool inner = i >= GH && i < g.nx + GH && j >= GH && j < g.ny + GH;
	double L[NV], R[NV], F[NV], hR[NV], hF[NV], ub = 0.0, hub = 0.0;
	// states [k][1 + column] (east halo at column TX + 1), then fluxes [k][face] (face f west of column f)
	if (valid) {
		phx[c] = recon_dir(g, P, limiter, order >= 2 && i >= 1 && i + 1 < NXT, Qe, vsc, esc, vu, vv, c, 1, L, R);
		put_donors(donx, NC, c, L, R);
		#pragma unroll
		for (int k = 0; k < NV; k++) sh[k * (TX + 2) + tx + 1] = L[k];
	}
	const int hx = tx - HX;
	bool hw = false;
	if (hx == 0 || hx == 1) {
		const int hi = hx ? i0 + TX : i0 - 1;
		if (hi >= 0 && hi < NXT && j < NYT) {
			double hL[NV];
			recon_dir(g, P, limiter, order >= 2 && hi >= 1 && hi + 1 < NXT, Qe, vsc, esc, vu, vv, hi + j * NXT, 1, hL, hR);
			if (hx) {
				#pragma unroll
				for (int k = 0; k < NV; k++) sh[k * (TX + 2) + TX + 1] = hL[k];
			} else hw = true;
		}
	}
	__syncthreads();
	const bool hasF = valid && i + 1 < NXT;
	if (hasF) {
		double QR[NV];
		#pragma unroll
		for (int k = 0; k < NV; k++) QR[k] = sh[k * (TX + 2) + tx + 2];
		rusanov_states(R, QR, IMU, F, ub);
	}
	if (hw) {
		double QR[NV];
		#pragma unroll
		for (int k = 0; k < NV; k++) QR[k] = sh[k * (TX + 2) + 1];
		rusanov_states(hR, QR, IMU, hF, hub);
	}
	__syncthreads();
	if (hasF) {
		#pragma unroll
		for (int k = 0; k < NV; k++) sh[k * (TX + 1) + tx + 1] = F[k];
		sh[NV * (TX + 1) + tx + 1] = ub;
	}
	if (hw) {
		#pragma unroll
		for (int k = 0; k < NV; k++) sh[k * (TX + 1)] = hF[k];
		sh[NV * (TX + 1)] = hub;
	}
	__syncthreads();
	if (inner) {
		double o[NV];
		const double ubw = sh[NV * (TX + 1) + tx];
		#pragma unroll
		for (int k = 0; k < NV; k++) o[k] = Qs[k * NC + c] - rxd * (F[k] - sh[k * (TX + 1) + tx]);
		o[IA1] += rxd * Qe[IA1 * NC + c] * (ub - ubw);
		ubx[c] = ub; if (i == GH) ubx[c - 1] = ubw;
		#pragma unroll
		for (int k = 0; k < NV; k++) Ust[k * NC + c] = o[k];
		a2s[c] = 1.0 - o[IA1];
	}
}

// The 2D explicit stage with one barrier per direction: each thread reconstructs its cell (halo threads the cells
// beside the tile) and puts both face states in LDS; after the barrier each cell takes its west / south neighbour's right
// state and its east / north neighbour's left state from LDS and computes both of its face fluxes (each interior face
// is computed by the two cells beside it, by the same operations, so both get the same value). Only the update is
// held in registers across a barrier. Halo threads: the 2 TY x-halo cells on threads 64.., the 2 TX y-halo cells on the
// last 2 TX threads.
template <int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_explicit_2d(Geo g, Phys P, int limiter, int order, double rxd, double ryd,
		const double* __restrict__ Qe, const double* __restrict__ Qs, const double* __restrict__ vsc, const double* __restrict__ esc,
		const double* __restrict__ vu, const double* __restrict__ vv, double* __restrict__ donx,
		double* __restrict__ dony, double* __restrict__ phx, double* __restrict__ phy, double* __restrict__ Ust,
		double* __restrict__ a2s, double* __restrict__ ubx, double* __restrict__ uby, int tiles_y) {
	constexpr int NT = TX * TY, HX = 64, HY = NT - 2 * TX;
	static_assert(HY >= HX + 2 * TY, "halo threads");
	// x: states [side][k][row][1 + column] (columns 0 and TX + 1 the halo); y: [side][k][1 + row][column]
	constexpr int SX = NV * TY * (TX + 2), SY = NV * (TY + 2) * TX;
	__shared__ double sh[2 * std::max(SX, SY)];
	const int NXT = g.NXT, NYT = g.NYT, NC = g.NC, tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
	const int tile = xcd_band(blockIdx.x, gridDim.x), i0 = (tile / tiles_y) * TX, j0 = (tile % tiles_y) * TY;
	const int i = i0 + tx, j = j0 + ty, c = i + j * NXT;
	const bool valid = i < NXT && j < NYT;
	const bool inner = i >= GH && i < g.nx + GH && j >= GH && j < g.ny + GH;
	double o[NV];
	// ---- x
	{
		auto at = [&](int side, int k, int row, int col) -> double& { return sh[side * SX + (k * TY + row) * (TX + 2) + col]; };
		if (valid) {
			double L[NV], R[NV];
			st_nt(phx + c, recon_dir(g, P, limiter, order >= 2 && i >= 1 && i + 1 < NXT, Qe, vsc, esc, vu, vv, c, 1, L, R));
			put_donors(donx, NC, c, L, R);
			#pragma unroll
			for (int k = 0; k < NV; k++) { at(0, k, ty, tx + 1) = L[k]; at(1, k, ty, tx + 1) = R[k]; }
		}
		const int hx = tid - HX;
		if (hx >= 0 && hx < 2 * TY) {
			const int hr = hx >> 1, e = hx & 1, hi = e ? i0 + TX : i0 - 1, hj = j0 + hr;
			if (hi >= 0 && hi < NXT && hj < NYT) {
				double L[NV], R[NV];
				recon_dir(g, P, limiter, order >= 2 && hi >= 1 && hi + 1 < NXT, Qe, vsc, esc, vu, vv, hi + hj * NXT, 1, L, R);
				#pragma unroll
				for (int k = 0; k < NV; k++) { at(0, k, hr, e ? TX + 1 : 0) = L[k]; at(1, k, hr, e ? TX + 1 : 0) = R[k]; }
			}
		}
		__syncthreads();
		if (inner) {
			double A[NV], B[NV], Fw[NV], Fe[NV], ubw, ube;
			#pragma unroll
			for (int k = 0; k < NV; k++) { A[k] = at(1, k, ty, tx); B[k] = at(0, k, ty, tx + 1); }
			rusanov_states(A, B, IMU, Fw, ubw);
			#pragma unroll
			for (int k = 0; k < NV; k++) { A[k] = at(1, k, ty, tx + 1); B[k] = at(0, k, ty, tx + 2); }
			rusanov_states(A, B, IMU, Fe, ube);
			#pragma unroll
			for (int k = 0; k < NV; k++) o[k] = Qs[k * NC + c] - rxd * (Fe[k] - Fw[k]);
			o[IA1] += rxd * Qe[IA1 * NC + c] * (ube - ubw);
			st_nt(ubx + c, ube); if (i == GH) st_nt(ubx + c - 1, ubw);
		}
		__syncthreads();
	}
	// ---- y
	{
		auto at = [&](int side, int k, int row, int col) -> double& { return sh[side * SY + (k * (TY + 2) + row) * TX + col]; };
		if (valid) {
			double L[NV], R[NV];
			st_nt(phy + c, recon_dir(g, P, limiter, order >= 2 && j >= 1 && j + 1 < NYT, Qe, vsc, esc, vu, vv, c, NXT, L, R));
			put_donors(dony, NC, c, L, R);
			#pragma unroll
			for (int k = 0; k < NV; k++) { at(0, k, ty + 1, tx) = L[k]; at(1, k, ty + 1, tx) = R[k]; }
		}
		const int hy = tid - HY;
		if (hy >= 0) {
			const int n = hy >= TX, hc = hy % TX, hi = i0 + hc, hj = n ? j0 + TY : j0 - 1;
			if (hi < NXT && hj >= 0 && hj < NYT) {
				double L[NV], R[NV];
				recon_dir(g, P, limiter, order >= 2 && hj >= 1 && hj + 1 < NYT, Qe, vsc, esc, vu, vv, hi + hj * NXT, NXT, L, R);
				#pragma unroll
				for (int k = 0; k < NV; k++) { at(0, k, n ? TY + 1 : 0, hc) = L[k]; at(1, k, n ? TY + 1 : 0, hc) = R[k]; }
			}
		}
		__syncthreads();
		if (inner) {
			double A[NV], B[NV], Fs[NV], Fn[NV], ubs, ubn;
			#pragma unroll
			for (int k = 0; k < NV; k++) { A[k] = at(1, k, ty, tx); B[k] = at(0, k, ty + 1, tx); }
			rusanov_states(A, B, IMV, Fs, ubs);
			#pragma unroll
			for (int k = 0; k < NV; k++) { A[k] = at(1, k, ty + 1, tx); B[k] = at(0, k, ty + 2, tx); }
			rusanov_states(A, B, IMV, Fn, ubn);
			#pragma unroll
			for (int k = 0; k < NV; k++) o[k] -= ryd * (Fn[k] - Fs[k]);
			o[IA1] += ryd * Qe[IA1 * NC + c] * (ubn - ubs);
			st_nt(uby + c, ubn); if (j == GH) st_nt(uby + c - NXT, ubs);
			#pragma unroll
			for (int k = 0; k < NV; k++) st_nt(Ust + k * NC + c, o[k]);
			st_nt(a2s + c, 1.0 - o[IA1]);
		}
	}
}

// ------------------------------------------------------------------------------------------------ implicit stage preparation
// per cell: donor velocities, frozen sound speed, residual scale K + P_E, previous pressure (PC:859-900)
__global__ void __launch_bounds__(256) k_prep_cell(Geo g, Phys P, const double* __restrict__ Xs, const double* __restrict__ a2s, const double* __restrict__ Pst,
                            double* __restrict__ ust, double* __restrict__ vst, double* __restrict__ cz, double* __restrict__ Kw,
                            double* __restrict__ Pold, int* __restrict__ flags) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	if (!prep_cell_at(g, P, Xs, a2s, Pst, ust, vst, cz, Kw, Pold, c)) flags[FLAG_STAGE] = 1;
}
// face coefficients of u_f = ubs - kp (P_R - P_L) and p_f = wL P_L + (1 - wL) P_R - pv (set_face, PC:926-950)
__global__ void __launch_bounds__(256) k_prep_face(Geo g, int order, int limiter, double tau, const double* __restrict__ Xs, const double* __restrict__ ust,
                            const double* __restrict__ vst, const double* __restrict__ cz, const double* __restrict__ Pold,
                            const double* __restrict__ phx, const double* __restrict__ phy,
                            double* __restrict__ ubsx, double* __restrict__ kpx, double* __restrict__ wLx, double* __restrict__ pvx,
                            double* __restrict__ ubsy, double* __restrict__ kpy, double* __restrict__ wLy, double* __restrict__ pvy) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	prep_face_at(g, order, limiter, tau, Xs, ust, vst, cz, Pold, phx, phy, ubsx, kpx, wLx, pvx, ubsy, kpy, wLy, pvy, c);
}
// shock budget of each interior cell on the +-1 cross stencil of X* (PHY:395-418), in two passes: every cell is
// classified and all but the strong-shock cells are settled, each strong-shock cell's profile problem going to a record
// of the list (SoA, field k of slot i at rec[k * cap + i]); the second pass then runs the strong cells' shootings
__global__ void __launch_bounds__(256) k_shock_classify(Geo g, Phys P, const double* __restrict__ Xs, double* __restrict__ Qph,
                                                       int* __restrict__ list, int* __restrict__ count, double* __restrict__ rec, long cap) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double r[SHOCK_REC];
	const bool strong = q < g.n && shock_classify_rec_at(g, P, Xs, Qph, q, r);
	const int slot = block_append(strong, q, list, count);
	if (slot >= 0) {
		#pragma unroll
		for (int k = 0; k < SHOCK_REC; k++) rec[k * cap + slot] = r[k];
	}
}
// second pass: persistent lanes, each running one strong cell's shooting at a time. A shooting is a sequence of
// adaptive RK45 profile integrations (share_strong's loop) whose lengths vary from cell to cell (on a cavitating flow
// ~10 RK steps per integration and ~3.6 integrations per cell on average, up to ~300 steps per cell), so a lane
// advances its integration by one step per pass of the loop and the lanes of a wavefront stay busy until the list is
// empty instead of waiting for the longest cell of a fixed group. Between integrations a lane needs service: the
// shooting's next point (shoot_next_sel), the logistic of the shooting parameter, the next integration's start or the
// cell's share, and a new cell when its cell is done. Service runs for all waiting lanes of the wavefront together,
// in a pass where at least SHOCK_SERVE lanes wait (or no lane is integrating), and the lanes it starts take their
// first step in the same pass: each pass then carries one RK step for most lanes, and the service's divergent code
// runs every few passes for many lanes instead of every pass for a few. A wavefront takes cells from a chunk of
// SHOCK_CHUNK consecutive list slots and claims the next chunk with one atomic on the cursor (an atomic per refill,
// from every wavefront on the one cursor, was measured to serialise: 2.5x slower). The shooting state, read only at
// service, lives in LDS. FUSED takes the profile integration in its fused form (prof_step_fused, one division per
// right-hand side). Every cell runs the same operations as share_strong (or, FUSED, as share_strong with the fused
// integration), so the result does not depend on the order or the launch shape. Every lane's work is bounded (4000 steps per integration, 100 Illinois iterations), so
// every wavefront reaches the exit.
constexpr int SHOCK_CHUNK = 128;
constexpr int SHOCK_SERVE = 16;
template <bool FUSED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_shock_strong(Geo g, Phys P, const double* __restrict__ rec, long cap,
                                                     double* __restrict__ Qph, const int* __restrict__ list, const int* __restrict__ count,
                                                     int* __restrict__ cursor, unsigned long long* __restrict__ rk_steps) {
	__shared__ Shoot Hs[256];
	__shared__ double vH0s[256], rhos[256];
	__shared__ int slots[256];
	Shoot& H = Hs[threadIdx.x];
	double& vH0 = vH0s[threadIdx.x];
	double& rho = rhos[threadIdx.x];
	int& slot = slots[threadIdx.x];
	const int m = *count;
	const int lane = threadIdx.x & 63;
	bool more = true;    // list slots are left (wavefront-uniform, as are qn and qe)
	bool has = false;    // the lane holds a cell whose shooting is not finished
	bool run = false;    // its profile integration is in progress
	bool ended = false;  // the integration has ended with r_end and the shooting has not taken the result yet
	int r_end = 0;
	int qn = 0, qe = 0;  // the wavefront's chunk of list slots not yet taken, [qn, qe)
	int nsteps = 0;      // RK45 steps this lane took (added to *rk_steps at the exit: the pass's work count)
	const ShockRcp rc = shock_rcp(P);
	Prof S; ProfRun R; ProfF F;
	for (;;) {
		const unsigned long long running = __ballot(run), waiting = __ballot(ended || (!has && more));
		if (running == 0ull && waiting == 0ull) break;
		if (waiting != 0ull && (running == 0ull || __popcll(waiting) >= SHOCK_SERVE)) {
			double t = 0.0;
			if (ended) { t = shoot_next_sel(H, share_miss(R, r_end, vH0, S.vk00), R.y2); ended = false; }
			if (has && !run) {   // the shooting's next integration, or its end: the share (shoot_result: lam = (c, 1 - c), q)
				const bool fin = H.phase >= SH_DONE;
				const double c = share_c(fin ? H.t : t);
				if (!fin) {
					prof_begin(S, c, R);
					if (FUSED) proff_begin(P, S, R, F);
					run = true;
				} else {
					double Q0 = 0.0, Q1 = 0.0;
					if (H.phase == SH_DONE && H.qt > 0) { Q0 = c * rho * H.qt; Q1 = (1.0 - c) * rho * H.qt; }
					const int q = list[slot], cc = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
					Qph[cc] = Q0; Qph[g.NC + cc] = Q1;
					has = false;
				}
			}
			if (more) {   // the idle lanes take the next cells of the chunk
				const unsigned long long idle = __ballot(!has);
				bool fresh = false;
				if (idle) {
					if (qn >= qe) {
						int nb = 0;
						if (lane == 0) nb = atomicAdd(cursor, SHOCK_CHUNK);
						nb = __shfl(nb, 0, 64);
						if (nb >= m) more = false;
						else { qn = nb; qe = min(m, nb + SHOCK_CHUNK); }
					}
					if (qn < qe) {
						const int avail = qe - qn, rank = __popcll(idle & ((1ull << lane) - 1ull));
						if (!has && rank < avail) {
							const int i = qn + rank;
							S.Y0 = rec[i]; S.Y1 = rec[cap + i]; S.vk00 = rec[2 * cap + i]; S.p0 = rec[3 * cap + i]; S.j2 = rec[4 * cap + i];
							S.v0 = rec[5 * cap + i]; S.vend = rec[6 * cap + i]; vH0 = rec[7 * cap + i]; rho = rec[8 * cap + i];
							t = shoot_begin(H, rec[9 * cap + i]);
							slot = i; has = true; fresh = true;
						}
						qn += min(__popcll(idle), avail);
					}
				}
				if (fresh) {
					prof_begin(S, share_c(t), R);
					if (FUSED) { proff_cell(S, F); proff_begin(P, S, R, F); }
					run = true;
				}
			}
		}
		if (run) {
			nsteps++;
			int r;
			if (FUSED) r = prof_step_fused(P, S, F, R);
			else if (mg_any(P)) r = prof_step(P, S, R);   // Mie-Grueneisen phases: the plain right-hand side
			else {
				r = prof_step<true>(P, S, R, &rc);
				if (r == 2) r = prof_step(P, S, R);   // (wavefront-uniform) a quotient's operand out of the shortcut's range
			}
			if (r != 0) { run = false; ended = true; r_end = r; }
		}
	}
	#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) nsteps += __shfl_xor(nsteps, off, 64);
	if (lane == 0) atomicAdd(rk_steps, (unsigned long long)nsteps);
}

// ------------------------------------------------------------------------------------------------ cell residual (PC:584-705)
// residual of every interior cell at P + lam dP (dP may be null); writes the trial pressure and G, the max-norm
// |G| / (|P_REF + P| + K + 1) (PC:770) and the closure-failure flag
__global__ void __launch_bounds__(256) k_residual(Dev D, const double* __restrict__ P, const double* __restrict__ dP, double lam,
                           double* __restrict__ Pout, double* __restrict__ G, unsigned long long* __restrict__ nrm, int* __restrict__ flags, int flag_slot,
                           const int* __restrict__ gate) {
	if (gate && !*gate) return;   // a speculative evaluation the device decided against (the same value in every block)
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double nv = 0.0;
	if (q < D.g.n && !residual_at<true>(D, P, dP, lam, Pout, G, q, nv)) flags[flag_slot] = 1;
	atomic_max_nonneg(nrm, nv);
}
// the same residual on the listed rows only (list[0, *count)), each thread taking rows at the grid's stride: the learned
// guess's residual where it can differ from the plain guess's (the other rows come from k_guess_tiles)
__global__ void __launch_bounds__(256) k_residual_list(Dev D, const double* __restrict__ P, const int* __restrict__ list, const int* __restrict__ count,
                                                      double* __restrict__ G, unsigned long long* __restrict__ nrm, int* __restrict__ flags, int flag_slot) {
	const int m = *count;
	double vmax = 0.0;
	bool bad = false;
	for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
		double nv;
		if (!residual_at(D, P, nullptr, 0.0, nullptr, G, list[i], nv)) bad = true;
		vmax = smax(vmax, nv);
	}
	if (bad) flags[flag_slot] = 1;
	atomic_max_nonneg(nrm, vmax);
}
// P scale s_q = (|P| + K + 1) / (1 + K D) of the finite-difference steps (PC:783-789)
// block 0 also zeroes clear[0, nclear) (a Newton step's norm slots and flags, in the pass that starts the step)
__device__ __forceinline__ void clear_ints(int* __restrict__ clear, int nclear) {
	if (blockIdx.x == 0) for (int k = threadIdx.x; k < nclear; k += blockDim.x) clear[k] = 0;
}
__global__ void __launch_bounds__(256) k_fdscale(Dev D, const double* __restrict__ P, double* __restrict__ sc, int* __restrict__ clear, int nclear) {
	clear_ints(clear, nclear);
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	fdscale_at(D, P, sc, q);
}
// finite-difference Jacobian row: self, W, E (, S, N) perturbed residuals (PC:790-804); J is 5 arrays of n; rhs = -G
__global__ void __launch_bounds__(256) k_jacobian(Dev D, const double* __restrict__ P, const double* __restrict__ G, const double* __restrict__ sc,
                           double* __restrict__ J, double* __restrict__ rhs, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	if (!jacobian_at(D, P, G, sc, J, rhs, q)) flags[FLAG_JAC] = 1;
}
// analytic Jacobian rows (the fast forms): one thread per row, no closure iterations. Two waves per SIMD (a few
// registers spilled) measured 252 us at 2M rows against 289 us at the compiler's default of one. The row's P scale
// stays a pass of its own (k_fdscale): computed here it measured 11 us slower than the two passes together.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_jacobian_an(Dev D, const double* __restrict__ P, const double* __restrict__ G,
                           double* __restrict__ J, double* __restrict__ rhs, int* __restrict__ flags) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	if (!jacobian_analytic_at<true>(D, P, G, J, rhs, q)) flags[FLAG_JAC] = 1;
}
// right-hand side -G of a chord step (the faster forms' Newton step on the previous Jacobian) and the row's P scale
__global__ void __launch_bounds__(256) k_chord_rhs(Dev D, const double* __restrict__ P, const double* __restrict__ G, double* __restrict__ sc,
                                                   double* __restrict__ rhs, int* __restrict__ clear, int nclear) {
	clear_ints(clear, nclear);
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= D.g.n) return;
	if (sc) fdscale_at(D, P, sc, q);   // a reused Jacobian (jreuse) keeps the scale of its build
	rhs[q] = -G[q];
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
	if (!floor_ok_at(D, P, tol, q)) flags[FLAG_FLOOR] = 1;
}
// final stage state at the converged pressure (PC:975-987) and the stage derivative Hs = (U_s - Qs) / adt (PC:1276-1284)
__global__ void __launch_bounds__(256) k_final(Dev D, const double* __restrict__ P, const double* __restrict__ Qs, double adt, double* __restrict__ H,
                                              int* __restrict__ flags, int flag_slot, unsigned long long* __restrict__ tmax) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double T0 = 0.0, T1 = 0.0;
	if (q < D.g.n && !final_at(D, P, Qs, adt, H, q, &T0, &T1)) flags[flag_slot] = 1;
	atomic_max_nonneg(tmax, T0);      // the running maxima of the phase temperatures (slots 0..63 and 64..127)
	atomic_max_nonneg(tmax + 64, T1);
}
// write_profile's cell columns (cell_diag, PHY:1105-1116) of every interior cell: out[q * 15 + k]
__global__ void __launch_bounds__(256) k_cell_diag(Geo g, Phys P, const double* __restrict__ U, double* __restrict__ out) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= g.n) return;
	const int c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT;
	double Q[NV], o[15];
	#pragma unroll
	for (int k = 0; k < NV; k++) Q[k] = U[k * g.NC + c];
	cell_diag(P, Q, o);
	#pragma unroll
	for (int k = 0; k < 15; k++) out[(size_t)q * 15 + k] = o[k];
}
// caldt maxima (PC:477-488): max |u| + c and |v| + c over interior cells with admissible primitives
__global__ void __launch_bounds__(256) k_caldt(Geo g, Phys P, const double* __restrict__ U, int acoustic, unsigned long long* __restrict__ nrm) {
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	double ux = 0.0, uy = 0.0;
	if (q < g.n) caldt_at(g, P, U, acoustic, q, ux, uy);
	atomic_max_nonneg(&nrm[0], ux);
	atomic_max_nonneg(&nrm[64], uy);
}

// ------------------------------------------------------------------------------------------------ Thomas (PC:807-812), one thread
__global__ void __launch_bounds__(256) k_thomas(int n, const double* __restrict__ J, const double* __restrict__ rhs, double* __restrict__ x, double* __restrict__ cp, double* __restrict__ dp) {
	if (blockIdx.x != 0 || threadIdx.x != 0) return;
	thomas_serial(n, J, rhs, x, cp, dp);
}

// ------------------------------------------------------------------------------------------------ BiCGStab (PC:725-758)
// Right-preconditioned BiCGStab with Jacobi + constant coarse space, five fused passes per iteration. No atomics and no
// grid fences: each pass writes one partial sum per block, and every block of the next pass re-reduces those partials
// itself in the fixed order (the kernel boundary makes them visible), so all blocks derive the same loop scalars. Scalars
// that outlive a pass (rho, alpha, omega, cc) are kept in KS, double-buffered by the parity of the iteration index the
// host passes in, so no pass reads a value that another block of the same pass writes. Every pass returns at once when
// the solve has finished, so the host launches iterations in chunks and reads KS::done once per chunk.
// Rows: block b, thread t owns rows b*256*IT + it*256 + t (it = 0..IT-1), the fixed order of the sums. IT is a
// compile-time count, and each thread loads the inputs of a chunk of its rows into registers before it computes them
// (the first chunk before the partial re-reduction that opens the pass), so several loads per thread are in flight.
template <typename T> struct KV { T *x, *r, *r0, *p, *v, *s, *t; const T* J; T *ph, *sh, *dg, *p2; };
struct KP { double* part; int stride; int nb; };
enum { SL_RR = 0, SL_R0R, SL_SP, SL_R0V, SL_SS, SL_SSUM, SL_TS, SL_TT, SL_E, SL_SB, SL_BB, NSLOTS };

template <typename T>
__device__ __forceinline__ T dinv_v(T d) { return d != T(0) ? T(1) / d : T(1); }
__device__ __forceinline__ bool is_red(const Geo& g, int r) { const int jj = r / g.nx; return ((r - jj * g.nx + jj) & 1) == 0; }
template <typename T>
__device__ __forceinline__ T dinv_of(const T* __restrict__ J, int r) { return dinv_v(J[r]); }
// the inputs of one row's product: entries (self, W, E, S, N), the vector and the diagonal at those columns
template <typename T> struct SpIn { T j[5], z[5], d[5]; };
// grid position of a pass's rows: one division for a thread's first row, then krow's steps of 256 rows
struct RowPos { int i, j; };
__device__ __forceinline__ RowPos rowpos(const Geo& g, int r) { const int j = r / g.nx; return RowPos{r - j * g.nx, j}; }
__device__ __forceinline__ RowPos rowpos_step(const Geo& g, RowPos p) { p.i += 256; while (p.i >= g.nx) { p.i -= g.nx; p.j++; } return p; }
// the position of row r in a pass: the running one in the FP32 red-black passes (measured faster there), one division
// per row in the FP64 passes (measured faster there)
template <typename T>
__device__ __forceinline__ void sp_load(const Geo& g, const T* __restrict__ J, const T* __restrict__ z, int r, RowPos pr, bool diag, SpIn<T>& s) {
	const int n = g.n; const Nb b = neighbours_at(g, r, pr.i, pr.j);
	const int col[5] = {r, b.W, b.E, b.S, b.N};
	const int ne = g.twoD ? 5 : 3;
	#pragma unroll
	for (int k = 0; k < 5; k++) if (k < ne) { s.j[k] = J[k * n + r]; s.z[k] = z[col[k]]; if (diag) s.d[k] = J[col[k]]; }
}
// y_r = sum_q A_rq (dinv_c z_c + cc) over the row's entries in the order self, W, E, S, N
template <typename T>
__device__ __forceinline__ T sp_prec(const SpIn<T>& s, int ne, T cc) {
	T y = T(0);
	#pragma unroll
	for (int k = 0; k < 5; k++) if (k < ne) y += s.j[k] * (dinv_v(s.d[k]) * s.z[k] + cc);
	return y;
}
template <typename T>
__device__ __forceinline__ T sp_plain(const SpIn<T>& s, int ne) {
	T y = T(0);
	#pragma unroll
	for (int k = 0; k < 5; k++) if (k < ne) y += s.j[k] * s.z[k];
	return y;
}
template <typename T>
__device__ __forceinline__ T spmv_plain(const Geo& g, const T* __restrict__ J, const T* __restrict__ z, int r) {
	SpIn<T> s; sp_load(g, J, z, r, rowpos(g, r), false, s); return sp_plain(s, g.twoD ? 5 : 3);
}
// logical block of a Krylov pass (xcd_band): the rows a block reads above and below its own (one grid line per block at
// the common sizes) are those its XCD's other blocks are reading at about the same time. The partials are stored by
// logical block, so every reduction sums in the same order as without the mapping.
__device__ __forceinline__ int kblk() { return xcd_band(blockIdx.x, gridDim.x); }
__device__ __forceinline__ int krow(int IT, int k) { return kblk() * 256 * IT + k * 256 + threadIdx.x; }
// this block's partial of one slot (thread 0 writes)
__device__ __forceinline__ void put_partial(const KP& P, int slot, double v, double* sh) {
	const double tot = block_sum256(v, sh);
	if (threadIdx.x == 0) P.part[slot * P.stride + kblk()] = tot;
}
// grid total of one slot written by the previous pass, in the fixed order (identical in every block): each thread adds
// its slots c = tid, tid + 256, ... in that order, the loads of eight of them issued before their additions
__device__ __forceinline__ double grid_total(const KP& P, int slot, double* sh) {
	const double* q = P.part + slot * P.stride;
	constexpr int U = 8;
	double s = 0.0;
	for (int c0 = threadIdx.x; c0 < P.nb; c0 += 256 * U) {
		double a[U];
		#pragma unroll
		for (int u = 0; u < U; u++) { const int c = c0 + u * 256; a[u] = c < P.nb ? q[c] : 0.0; }
		#pragma unroll
		for (int u = 0; u < U; u++) if (c0 + u * 256 < P.nb) s += a[u];
	}
	return block_sum256(s, sh);
}
// rows per register chunk of the vector passes and of the product passes: FP32 passes gain from several rows' loads in
// flight per thread; the FP64 passes measured best row by row (their per-row loads are already twice as wide)
template <typename T> constexpr int chv() { return sizeof(T) == 4 ? 4 : 1; }
template <typename T> constexpr int chs() { return sizeof(T) == 4 ? 2 : 1; }

// the neighbours of row r at grid position pr, as neighbours() gives them (col[1..4] = W, E, S, N; the row itself for a
// missing one), and whether each is red: an interior neighbour has the other colour of a black row, a periodic
// partner's colour follows from its position, as is_red would give it
__device__ __forceinline__ void rb_nbrs(const Geo& g, int r, RowPos pr, int col[5], bool red[5]) {
	const int nx = g.nx, ny = g.ny, ii = pr.i, jj = pr.j;
	const bool yp = g.yper && ny > 1;
	col[0] = r; red[0] = false;
	col[1] = ii > 0 ? r - 1 : (g.xper ? r + nx - 1 : r);            red[1] = ii > 0 || ((nx - 1 + jj) & 1) == 0;
	col[2] = ii < nx - 1 ? r + 1 : (g.xper ? r - (nx - 1) : r);     red[2] = ii < nx - 1 || (jj & 1) == 0;
	col[3] = jj > 0 ? r - nx : (yp ? r + (ny - 1) * nx : r);        red[3] = jj > 0 || ((ii + ny - 1) & 1) == 0;
	col[4] = jj < ny - 1 ? r + nx : (yp ? r - (ny - 1) * nx : r);   red[4] = jj < ny - 1 || (ii & 1) == 0;
}
// row r of y = M^-1 z for the red-black sweep: z_r / d on a red row; on a black row (z_r - sum over red neighbours c
// of A_rc z_c / d_c) / d, with the neighbours' z_c from zc(c) (the pass recomputes them from their own inputs). The
// row's grid position takes one division; its neighbours are those of neighbours(), an interior one being red (the
// other colour) and a periodic partner's colour following from its position, as is_red would give it.
template <typename T, typename ZC>
__device__ __forceinline__ T rb_row(const Geo& g, const KV<T>& V, int r, RowPos pr, T zr, ZC zc) {
	const int jj = pr.j, ii = pr.i, n = g.n, ne = g.twoD ? 5 : 3;
	int col[5]; bool red[5];
	rb_nbrs(g, r, pr, col, red);
	// every neighbour's inputs are loaded first, for every lane (a missing neighbour is the row itself), so that the
	// loads are in flight together; the sum then takes the red neighbours in the order W, E, S, N
	T zd[5], jv[5];
	#pragma unroll
	for (int k = 1; k < 5; k++) if (k < ne) { zd[k] = zc(col[k]) * V.dg[col[k]]; jv[k] = V.J[k * n + r]; }
	const T d = V.dg[r];
	if (((ii + jj) & 1) == 0) return zr * d;
	T acc = zr;
	#pragma unroll
	for (int k = 1; k < 5; k++) if (k < ne && col[k] != r && red[k]) acc -= jv[k] * zd[k];
	return acc * d;
}
// setup 1: partials of E = sum of all entries, sum(b), b.b (PC:726-738)
template <typename T, int IT>
__global__ void __launch_bounds__(256) kb_setup1(Geo g, const T* __restrict__ J, const T* __restrict__ b, KP P) {
	__shared__ double sh[4];
	const int n = g.n; const int nent = g.twoD ? 5 : 3;
	double sE = 0.0, sb = 0.0, sbb = 0.0;
	#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int r = krow(IT, it); if (r >= n) continue;
		T e = T(0); for (int qq = 0; qq < nent; qq++) e += J[qq * n + r];
		sE += (double)e; sb += (double)b[r]; sbb += (double)(b[r] * b[r]);
	}
	put_partial(P, SL_E, sE, sh); put_partial(P, SL_SB, sb, sh); put_partial(P, SL_BB, sbb, sh);
}
// setup 2: Einv, cc_b, bn; x = M^-1 b, p = v = 0
template <typename T, int IT>
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
	#pragma unroll
	for (int it = 0; it < IT; it++) { const int r = krow(IT, it); if (r < n) { V.x[r] = dinv_of(J, r) * b[r] + cc; V.p[r] = T(0); V.v[r] = T(0); } }
}
// setup 3: r = b - A x, r0 = r; partials of r.r, r0.r
template <typename T, int IT>
__global__ void __launch_bounds__(256) kb_setup3(Geo g, const T* __restrict__ b, KV<T> V, KP P) {
	__shared__ double sh[4];
	const int n = g.n;
	double srr = 0.0, sr0r = 0.0;
	#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int r = krow(IT, it); if (r >= n) continue;
		const T ax = spmv_plain(g, V.J, (const T*)V.x, r);
		const T rv = b[r] - ax; V.r[r] = rv; V.r0[r] = rv;
		srr += (double)(rv * rv); sr0r += (double)(rv * rv);
	}
	put_partial(P, SL_RR, srr, sh); put_partial(P, SL_R0R, sr0r, sh);
}
// I1 (iteration it): loop top (PC:741-744) from r.r and r0.r, then p = r + beta (p - omega v); partials of sum(p)
template <typename T, int IT, bool RB>
__global__ void __launch_bounds__(256) kb_p(Geo g, KV<T> V, KS* ks, KP P, int it, double tol, int maxit) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n, cur = it & 1, nxt = (it + 1) & 1;
	constexpr int CHV = chv<T>();
	T rv[CHV], pv[CHV], vv[CHV];
	#pragma unroll
	for (int k = 0; k < CHV; k++) { const int r = krow(IT, k); if (k < IT && r < n) { rv[k] = V.r[r]; pv[k] = V.p[r]; vv[k] = V.v[r]; } }
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
	RowPos pos = RB ? rowpos(g, krow(IT, 0)) : RowPos{0, 0};
	#pragma unroll
	for (int c0 = 0; c0 < IT; c0 += CHV) {
		if (c0 > 0) {
			#pragma unroll
			for (int k = 0; k < CHV; k++) { const int r = krow(IT, c0 + k); if (c0 + k < IT && r < n) { rv[k] = V.r[r]; pv[k] = V.p[r]; vv[k] = V.v[r]; } }
		}
		#pragma unroll
		for (int k = 0; k < CHV; k++) {
			const RowPos pk = pos; if (RB) pos = rowpos_step(g, pos);
			const int r = krow(IT, c0 + k); if (!(c0 + k < IT && r < n)) continue;
			const T pn = rv[k] + beta * (pv[k] - omega * vv[k]);
			if constexpr (RB) {   // new p into the other buffer: black rows read their red neighbours' old p
				V.p2[r] = pn;
				V.ph[r] = rb_row(g, V, r, pk, pn, [&](int c) { return V.r[c] + beta * (V.p[c] - omega * V.v[c]); });
			}
			else { V.p[r] = pn; sp += (double)pn; }
		}
	}
	if constexpr (!RB) put_partial(P, SL_SP, sp, sh);
}
// I2: cc_p = sum(p) Einv; v = A M^-1 p; partials of r0.v
template <typename T, int IT, bool RB>
__global__ void __launch_bounds__(256) kb_v(Geo g, KV<T> V, KS* ks, KP P, int it) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n, ne = g.twoD ? 5 : 3;
	constexpr int CHS = chs<T>();
	const T* z = RB ? (const T*)V.ph : (const T*)V.p;
	SpIn<T> q[CHS]; T r0v[CHS];
	RowPos pos = rowpos(g, krow(IT, 0));
	#pragma unroll
	for (int k = 0; k < CHS; k++) {
		const RowPos pk = pos; pos = rowpos_step(g, pos);
		const int r = krow(IT, k); if (k < IT && r < n) { sp_load(g, V.J, z, r, RB ? pk : rowpos(g, r), !RB, q[k]); r0v[k] = V.r0[r]; }
	}
	T cc = T(0);
	if constexpr (!RB) {
		cc = (T)grid_total(P, SL_SP, sh) * (T)ks->Einv;
		if (blockIdx.x == 0 && threadIdx.x == 0) ks->cc_p = (double)cc;
	}
	double s = 0.0;
	#pragma unroll
	for (int c0 = 0; c0 < IT; c0 += CHS) {
		if (c0 > 0) {
			#pragma unroll
			for (int k = 0; k < CHS; k++) {
				const RowPos pk = pos; pos = rowpos_step(g, pos);
				const int r = krow(IT, c0 + k); if (c0 + k < IT && r < n) { sp_load(g, V.J, z, r, RB ? pk : rowpos(g, r), !RB, q[k]); r0v[k] = V.r0[r]; }
			}
		}
		#pragma unroll
		for (int k = 0; k < CHS; k++) {
			const int r = krow(IT, c0 + k); if (!(c0 + k < IT && r < n)) continue;
			const T vv = RB ? sp_plain(q[k], ne) : sp_prec(q[k], ne, cc); V.v[r] = vv; s += (double)(r0v[k] * vv);
		}
	}
	put_partial(P, SL_R0V, s, sh);
}
// I3: alpha = rho / r0.v; s = r - alpha v; partials of s.s and sum(s)
template <typename T, int IT, bool RB>
__global__ void __launch_bounds__(256) kb_s(Geo g, KV<T> V, KS* ks, KP P, int it) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n, nxt = (it + 1) & 1;
	constexpr int CHV = chv<T>();
	T rv[CHV], vv[CHV];
	#pragma unroll
	for (int k = 0; k < CHV; k++) { const int r = krow(IT, k); if (k < IT && r < n) { rv[k] = V.r[r]; vv[k] = V.v[r]; } }
	const T alpha = (T)ks->rho[nxt] / (T)grid_total(P, SL_R0V, sh);
	if (blockIdx.x == 0 && threadIdx.x == 0) ks->alpha[nxt] = (double)alpha;
	double sss = 0.0, ssum = 0.0;
	RowPos pos = RB ? rowpos(g, krow(IT, 0)) : RowPos{0, 0};
	#pragma unroll
	for (int c0 = 0; c0 < IT; c0 += CHV) {
		if (c0 > 0) {
			#pragma unroll
			for (int k = 0; k < CHV; k++) { const int r = krow(IT, c0 + k); if (c0 + k < IT && r < n) { rv[k] = V.r[r]; vv[k] = V.v[r]; } }
		}
		#pragma unroll
		for (int k = 0; k < CHV; k++) {
			const RowPos pk = pos; if (RB) pos = rowpos_step(g, pos);
			const int r = krow(IT, c0 + k); if (!(c0 + k < IT && r < n)) continue;
			const T sv = rv[k] - alpha * vv[k]; V.s[r] = sv; sss += (double)(sv * sv); ssum += (double)sv;
			if constexpr (RB) V.sh[r] = rb_row(g, V, r, pk, sv, [&](int c) { return V.r[c] - alpha * V.v[c]; });
		}
	}
	put_partial(P, SL_SS, sss, sh);
	if constexpr (!RB) put_partial(P, SL_SSUM, ssum, sh);   // sum(s) feeds the coarse-space constant only
}
// I4: early exit if |s| <= tol |b| (x += alpha M^-1 p, PC:751); otherwise cc_s = sum(s) Einv, t = A M^-1 s; partials of
// t.s and t.t
template <typename T, int IT, bool RB>
__global__ void __launch_bounds__(256) kb_t(Geo g, KV<T> V, KS* ks, KP P, int it, double tol) {
	if (ks->done) return;
	__shared__ double sh[4];
	const int n = g.n, nxt = (it + 1) & 1, ne = g.twoD ? 5 : 3;
	constexpr int CHS = chs<T>();
	const T* z = RB ? (const T*)V.sh : (const T*)V.s;
	SpIn<T> q[CHS]; T sv[CHS];
	RowPos pos = rowpos(g, krow(IT, 0));
	#pragma unroll
	for (int k = 0; k < CHS; k++) {
		const RowPos pk = pos; pos = rowpos_step(g, pos);
		const int r = krow(IT, k); if (k < IT && r < n) { sp_load(g, V.J, z, r, RB ? pk : rowpos(g, r), !RB, q[k]); if constexpr (RB) sv[k] = V.s[r]; }
	}
	const double s2 = grid_total(P, SL_SS, sh), Ss = RB ? 0.0 : grid_total(P, SL_SSUM, sh);
	const T alpha = (T)ks->alpha[nxt];
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if ((T)sqrt((double)(T)s2) <= (T)tol * (T)ks->bn) {
		const T ccp = (T)ks->cc_p;
		#pragma unroll
		for (int k = 0; k < IT; k++) {
			const int r = krow(IT, k);
			if (r < n) V.x[r] += alpha * (RB ? V.ph[r] : dinv_of(V.J, r) * V.p[r] + ccp);
		}
		if (lead) { ks->early = 1; ks->iters = it + 1; ks->status = 0; }
		return;
	}
	const T cc = RB ? T(0) : (T)Ss * (T)ks->Einv;
	if (lead) ks->cc_s = (double)cc;
	double sts = 0.0, stt = 0.0;
	#pragma unroll
	for (int c0 = 0; c0 < IT; c0 += CHS) {
		if (c0 > 0) {
			#pragma unroll
			for (int k = 0; k < CHS; k++) {
				const RowPos pk = pos; pos = rowpos_step(g, pos);
				const int r = krow(IT, c0 + k); if (c0 + k < IT && r < n) { sp_load(g, V.J, z, r, RB ? pk : rowpos(g, r), !RB, q[k]); if constexpr (RB) sv[k] = V.s[r]; }
			}
		}
		#pragma unroll
		for (int k = 0; k < CHS; k++) {
			const int r = krow(IT, c0 + k); if (!(c0 + k < IT && r < n)) continue;
			const T tv = RB ? sp_plain(q[k], ne) : sp_prec(q[k], ne, cc); V.t[r] = tv;
			sts += (double)(tv * (RB ? sv[k] : q[k].z[0])); stt += (double)(tv * tv);
		}
	}
	put_partial(P, SL_TS, sts, sh); put_partial(P, SL_TT, stt, sh);
}
// I5: omega = t.s / t.t; x += alpha M^-1 p + omega M^-1 s; r = s - omega t; partials of r.r and r0.r
template <typename T, int IT, bool RB>
__global__ void __launch_bounds__(256) kb_xr(Geo g, KV<T> V, KS* ks, KP P, int it) {
	if (ks->done) return;
	const int nxt = (it + 1) & 1;
	if (ks->early) { if (blockIdx.x == 0 && threadIdx.x == 0) ks->done = 1; return; }
	__shared__ double sh[4];
	const int n = g.n;
	constexpr int CHV = chv<T>();
	T dv[CHV], pv[CHV], sv[CHV], tv[CHV], xv[CHV], r0v[CHV];
	auto load = [&](int c0) {
		#pragma unroll
		for (int k = 0; k < CHV; k++) {
			const int r = krow(IT, c0 + k);
			if (c0 + k < IT && r < n) {
				if constexpr (RB) { dv[k] = V.sh[r]; pv[k] = V.ph[r]; } else { dv[k] = V.J[r]; pv[k] = V.p[r]; }
				sv[k] = V.s[r]; tv[k] = V.t[r]; xv[k] = V.x[r]; r0v[k] = V.r0[r];
			}
		}
	};
	load(0);
	const double ts = grid_total(P, SL_TS, sh), tt = grid_total(P, SL_TT, sh);
	const T omega = (T)ts / (T)tt;
	const T alpha = (T)ks->alpha[nxt], ccp = (T)ks->cc_p, ccs = (T)ks->cc_s;
	if (blockIdx.x == 0 && threadIdx.x == 0) { ks->omega[nxt] = (double)omega; ks->iters = it + 1; }
	double srr = 0.0, sr0r = 0.0;
	#pragma unroll
	for (int c0 = 0; c0 < IT; c0 += CHV) {
		if (c0 > 0) load(c0);
		#pragma unroll
		for (int k = 0; k < CHV; k++) {
			const int r = krow(IT, c0 + k); if (!(c0 + k < IT && r < n)) continue;
			T ph, shv;
			if constexpr (RB) { ph = pv[k]; shv = dv[k]; }
			else { const T di = dinv_v(dv[k]); ph = di * pv[k] + ccp; shv = di * sv[k] + ccs; }
			V.x[r] = xv[k] + (alpha * ph + omega * shv);
			const T rv = sv[k] - omega * tv[k]; V.r[r] = rv;
			srr += (double)(rv * rv); sr0r += (double)(r0v[k] * rv);
		}
	}
	put_partial(P, SL_RR, srr, sh); put_partial(P, SL_R0R, sr0r, sh);
}

// Red-black variant (the FP32 form): M^-1 is one forward red-black Gauss-Seidel sweep from zero -- red rows
// ((i + j) even) y = z / d, black rows y = (z - sum over red neighbours of A_rq y_q) / d, d being the row's diagonal
// with its boundary self-columns (a neighbour index equal to the row) added. Couplings between rows of one colour (an
// odd periodic wrap) are left out, so M^-1 is a fixed linear operator. kb_p and kb_s write M^-1 p and M^-1 s (rb_row: a
// black row recomputes its red neighbours' new p or s), and the products read them (no constant coarse space: with the
// sweep it slowed convergence on every workload measured, e.g. 3 -> 6 iterations on the 2048x1024 shock-bubble system).
// setup: dg = 1 / d; x = M^-1 b on the red rows (kbr_black completes it); p = v = 0; partials of b.b
template <typename T, int IT>
__global__ void __launch_bounds__(256) kbr_setup1(Geo g, const T* __restrict__ b, KV<T> V, KP P) {
	__shared__ double sh[4];
	const int n = g.n, ne = g.twoD ? 5 : 3;
	double sbb = 0.0;
	#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int r = krow(IT, it); if (r >= n) continue;
		const Nb nb = neighbours(g, r);
		const int col[5] = {r, nb.W, nb.E, nb.S, nb.N};
		T jv[5];
		#pragma unroll
		for (int k = 0; k < 5; k++) if (k < ne) jv[k] = V.J[k * n + r];   // loaded together, added below as before
		T d = jv[0];
		#pragma unroll
		for (int k = 1; k < 5; k++) if (k < ne && col[k] == r) d += jv[k];
		const T di = dinv_v(d);
		V.dg[r] = di; V.x[r] = is_red(g, r) ? b[r] * di : T(0); V.p[r] = T(0); V.v[r] = T(0);
		sbb += (double)(b[r] * b[r]);
	}
	put_partial(P, SL_BB, sbb, sh);
}
// kbr_setup1 on the FP64 system of a Newton step's first sweep: b32 = (float) b64 and dP = 0, and with J64 (a new
// Jacobian) J32 = (float) J64 on the row's ne diagonals, the values the separate conversions would write
template <int IT>
__global__ void __launch_bounds__(256) kbr_setup1_cvt(Geo g, const double* __restrict__ J64, float* __restrict__ J32, const double* __restrict__ b64,
                                                      float* __restrict__ b32, double* __restrict__ dP, KV<float> V, KP P) {
	__shared__ double sh[4];
	const int n = g.n, ne = g.twoD ? 5 : 3;
	double sbb = 0.0;
	#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int r = krow(IT, it); if (r >= n) continue;
		const Nb nb = neighbours(g, r);
		const int col[5] = {r, nb.W, nb.E, nb.S, nb.N};
		float jv[5];
		#pragma unroll
		for (int k = 0; k < 5; k++) if (k < ne) jv[k] = J64 ? (float)J64[k * n + r] : J32[k * n + r];
		if (J64) {
			#pragma unroll
			for (int k = 0; k < 5; k++) if (k < ne) J32[k * n + r] = jv[k];
		}
		float d = jv[0];
		#pragma unroll
		for (int k = 1; k < 5; k++) if (k < ne && col[k] == r) d += jv[k];
		const float di = dinv_v(d), bf = (float)b64[r];
		b32[r] = bf; dP[r] = 0.0;
		V.dg[r] = di; V.x[r] = is_red(g, r) ? bf * di : 0.f; V.p[r] = 0.f; V.v[r] = 0.f;
		sbb += (double)(bf * bf);
	}
	put_partial(P, SL_BB, sbb, sh);
}
// black rows of y = M^-1 z from the red rows already in y; skipped once the solve is done (unless in the setup)
template <typename T, int IT>
__global__ void __launch_bounds__(256) kbr_black(Geo g, const T* __restrict__ z, T* __restrict__ y, const T* __restrict__ J,
                                                 const T* __restrict__ dg, KS* ks, int setup) {
	if (!setup && ks->done) return;
	const int n = g.n, ne = g.twoD ? 5 : 3;
	#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int r = krow(IT, it); if (r >= n) continue;
		const RowPos pr = rowpos(g, r);
		if (((pr.i + pr.j) & 1) == 0) continue;   // red row
		int col[5]; bool red[5];
		rb_nbrs(g, r, pr, col, red);
		T jv[5], yv[5];   // loaded together (a missing neighbour is the row itself), summed in the order W, E, S, N
		#pragma unroll
		for (int k = 1; k < 5; k++) if (k < ne) { jv[k] = J[k * n + r]; yv[k] = y[col[k]]; }
		T acc = z[r];
		#pragma unroll
		for (int k = 1; k < 5; k++) if (k < ne && col[k] != r && red[k]) acc -= jv[k] * yv[k];
		y[r] = acc * dg[r];
	}
}
// setup 2: loop scalars from b.b; r = b - A x, r0 = r; partials of r.r, r0.r
template <typename T, int IT>
__global__ void __launch_bounds__(256) kbr_setup2(Geo g, const T* __restrict__ b, KV<T> V, KS* ks, KP P) {
	__shared__ double sh[4];
	const int n = g.n;
	const double bb = grid_total(P, SL_BB, sh);
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		ks->Einv = 0.0; ks->cc_p = 0.0; ks->cc_s = 0.0; ks->bn = (double)((T)sqrt((double)(T)bb) + (T)1e-300);
		ks->rho[0] = 1; ks->alpha[0] = 1; ks->omega[0] = 1;
		ks->done = 0; ks->iters = 0; ks->early = 0; ks->status = 0;
	}
	double srr = 0.0, sr0r = 0.0;
	#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int r = krow(IT, it); if (r >= n) continue;
		const T rv = b[r] - spmv_plain(g, V.J, (const T*)V.x, r); V.r[r] = rv; V.r0[r] = rv;
		srr += (double)(rv * rv); sr0r += (double)(rv * rv);
	}
	put_partial(P, SL_RR, srr, sh); put_partial(P, SL_R0R, sr0r, sh);
}

// FP64 defect r = b - J (dP + x) of the accumulated step plus the running FP32 solution x, without writing dP (the FP32
// solve may continue), and the Newton step criterion (PC:819-823) of dP + x: block partials of r.r and b.b (slots 0, 1
// of stride) and block maxima of |dP + x| / (|P| + s) (slot 2); the values equal those of r = b - J dP and k_stepcrit
// after k_add_f32.
// IT rows per thread in the Krylov passes' layout (krow: XCD bands, partials by logical block).
template <int IT>
__global__ void __launch_bounds__(256) k_linres_x(Geo g, const double* __restrict__ J, const double* __restrict__ b, const double* __restrict__ dP,
                                                 const float* __restrict__ x, const double* __restrict__ P, const double* __restrict__ sc,
                                                 double* __restrict__ r, double* __restrict__ part, int stride) {
	__shared__ double sh[4];
	const int n = g.n;
	double rr = 0.0, bb = 0.0, crit = 0.0;
	#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int q = krow(IT, it); if (q >= n) continue;
		const Nb nb = neighbours(g, q);
		const int col[5] = {q, nb.W, nb.E, nb.S, nb.N}, ne = g.twoD ? 5 : 3;
		double y = 0.0;
		#pragma unroll
		for (int k = 0; k < 5; k++) if (k < ne) y += J[k * n + q] * (dP[col[k]] + (double)x[col[k]]);
		const double bq = b[q], v = bq - y; r[q] = v; rr += v * v; bb += bq * bq;
		crit = smax(crit, fabs(dP[q] + (double)x[q]) / (fabs(P[q]) + sc[q]));
	}
	const double t1 = block_sum256(rr, sh), t2 = block_sum256(bb, sh);
	crit = wave_max(crit);
	__syncthreads();
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = crit;
	__syncthreads();
	if (threadIdx.x == 0) {
		const int lb = kblk();
		part[lb] = t1; part[stride + lb] = t2;
		part[2 * stride + lb] = smax(smax(sh[0], sh[1]), smax(sh[2], sh[3]));
	}
}
// the partial slots reduced in one block in a fixed order: out = (r.r, b.b, max criterion); each thread issues the
// loads of eight slots before it adds them, so the block waits on memory once per 2048 slots instead of once per 256
__global__ void __launch_bounds__(256) k_sum_part2(const double* __restrict__ part, int nb, int stride, double* __restrict__ out) {
	__shared__ double sh[4];
	constexpr int U = 8;
	double s0 = 0.0, s1 = 0.0, m = 0.0;
	for (int c0 = threadIdx.x; c0 < nb; c0 += 256 * U) {
		double a[U], b[U], e[U];
		#pragma unroll
		for (int u = 0; u < U; u++) {
			const int c = c0 + u * 256;
			const bool in = c < nb;
			a[u] = in ? part[c] : 0.0; b[u] = in ? part[stride + c] : 0.0; e[u] = in ? part[2 * stride + c] : 0.0;
		}
		#pragma unroll
		for (int u = 0; u < U; u++) { s0 += a[u]; s1 += b[u]; m = smax(m, e[u]); }
	}
	s0 = block_sum256(s0, sh); s1 = block_sum256(s1, sh);
	m = wave_max(m);
	__syncthreads();
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
	__syncthreads();
	if (threadIdx.x == 0) { out[0] = s0; out[1] = s1; out[2] = smax(smax(sh[0], sh[1]), smax(sh[2], sh[3])); }
}
// x += (double) d
// the close of an FP32 Krylov sweep, launched behind each chunk so that the host reads the sweep's end and the Newton
// step's first trial residual in one synchronisation: once the solve is done, dP += x (the host then adds nothing) and
// the sweep's defect test |r| <= eta |b| (or b = 0), the host's test, is decided here into *spec; the trial residual
// launched next is gated on it. Block 0 clears that residual's norm slots and flag (the other flags stay for the host).
__global__ void __launch_bounds__(256) k_spec_commit(long n, const float* __restrict__ x, double* __restrict__ dP, const KS* __restrict__ ks,
                                                     const double* __restrict__ dsum, double eta, int* __restrict__ spec,
                                                     unsigned long long* __restrict__ nrm, int* __restrict__ flags) {
	const bool done = ks->done != 0;
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (done && i < n) dP[i] += (double)x[i];
	if (blockIdx.x == 0) {
		if (threadIdx.x < 64) nrm[threadIdx.x] = 0ull;
		if (threadIdx.x == 0) {
			const double rr = dsum[0], bb = dsum[1];
			*spec = done && (sqrt(rr) <= eta * sqrt(bb) || !(bb > 0)) ? 1 : 0;
			flags[FLAG_RES] = 0;
		}
	}
}
__global__ void __launch_bounds__(256) k_add_f32(long n, const float* __restrict__ d, double* __restrict__ x) {
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) x[i] += (double)d[i];
}
void launch_defect_x(const Geo& g, const double* J, const double* b, const double* dP, const float* x, double* r, double* part, int stride,
                     double* sums, const double* P, const double* sc, hipStream_t st) {
	constexpr int IT = 8;
	const unsigned nb = (unsigned)((g.n + 256 * IT - 1) / (256 * IT));
	hipLaunchKernelGGL(k_linres_x<IT>, dim3(nb), dim3(256), 0, st, g, J, b, dP, x, P, sc, r, part, stride);
	hipLaunchKernelGGL(k_sum_part2, dim3(1), dim3(256), 0, st, (const double*)part, (int)nb, stride, sums);
}
void launch_add_f32(long n, const float* d, double* x, hipStream_t st) { hipLaunchKernelGGL(k_add_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, d, x); }

// precision conversions around the FP32 Krylov solve (the Newton system stays FP64). Two elements per thread (16-byte
// FP64 accesses) where the host finds the arrays aligned for it, the odd last element by the thread after the pairs.
__global__ void __launch_bounds__(256) k_to_f32(long n, const double* __restrict__ a, float* __restrict__ b) {
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) b[i] = (float)a[i];
}
__global__ void __launch_bounds__(256) k_to_f32x2(long n, const double* __restrict__ a, float* __restrict__ b) {
	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x, h = n >> 1;
	if (i < h) { const double2 v = reinterpret_cast<const double2*>(a)[i]; reinterpret_cast<float2*>(b)[i] = make_float2((float)v.x, (float)v.y); }
	else if (i == h && (n & 1)) b[n - 1] = (float)a[n - 1];
}
static bool host_aligned(const void* p, unsigned long long al) { return (reinterpret_cast<unsigned long long>(p) & (al - 1)) == 0; }
void launch_to_f32(long n, const double* a, float* b, hipStream_t st) {
	if (host_aligned(a, 16) && host_aligned(b, 8)) {
		const long t = (n >> 1) + (n & 1);
		hipLaunchKernelGGL(k_to_f32x2, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, st, n, a, b);
	}
	else hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b);
}

// rows per thread of the Krylov passes: about a thousand blocks whatever n (4 .. 32 rows per thread)
int krylov_items(int n) {
	int it = 4;
	while (it < 32 && (long)n > (long)256 * it * 2048) it *= 2;
	return it;
}
// FP32 form: no summation order to match; about 1k blocks measured best at 2M rows
static int krylov_items32(int n) {
	int it = 4;
	while (it < 64 && (long)n > (long)256 * it * 1024) it *= 2;
	return it;
}
template <typename T> static int items_of(int n) { return sizeof(T) == 8 ? krylov_items(n) : krylov_items32(n); }
// calls f with the rows-per-thread count as a compile-time constant
template <typename F> static void with_items(int items, F f) {
	switch (items) {
	case 4: f(std::integral_constant<int, 4>{}); break;
	case 8: f(std::integral_constant<int, 8>{}); break;
	case 16: f(std::integral_constant<int, 16>{}); break;
	case 32: f(std::integral_constant<int, 32>{}); break;
	default: f(std::integral_constant<int, 64>{}); break;
	}
}
template <typename T>
static void krylov_setup_t(const Geo& g, const KrylovBufsT<T>& B, const T* b, hipStream_t st) {
	const int items = items_of<T>(g.n), nb = (g.n + 256 * items - 1) / (256 * items);
	KP P{B.part, B.stride, nb};
	KV<T> V{B.x, B.r, B.r0, B.p, B.v, B.s, B.t, B.J, B.ph, B.sh, B.dg, B.p2};
	with_items(items, [&](auto I) {
		constexpr int IT = decltype(I)::value;
		if (B.dg) {
			hipLaunchKernelGGL((kbr_setup1<T, IT>), dim3(nb), dim3(256), 0, st, g, b, V, P);
			hipLaunchKernelGGL((kbr_black<T, IT>), dim3(nb), dim3(256), 0, st, g, b, B.x, B.J, (const T*)B.dg, B.ks, 1);
			hipLaunchKernelGGL((kbr_setup2<T, IT>), dim3(nb), dim3(256), 0, st, g, b, V, B.ks, P);
			return;
		}
		hipLaunchKernelGGL((kb_setup1<T, IT>), dim3(nb), dim3(256), 0, st, g, B.J, b, P);
		hipLaunchKernelGGL((kb_setup2<T, IT>), dim3(nb), dim3(256), 0, st, g, B.J, b, V, B.ks, P);
		hipLaunchKernelGGL((kb_setup3<T, IT>), dim3(nb), dim3(256), 0, st, g, b, V, P);
	});
}
template <typename T>
static void krylov_iter_t(const Geo& g, const KrylovBufsT<T>& B, hipStream_t st, int it, double tol, int maxit) {
	const int items = items_of<T>(g.n), nb = (g.n + 256 * items - 1) / (256 * items);
	KP P{B.part, B.stride, nb};
	KV<T> V{B.x, B.r, B.r0, B.p, B.v, B.s, B.t, B.J, B.ph, B.sh, B.dg, B.p2};
	with_items(items, [&](auto I) {
		constexpr int IT = decltype(I)::value;
		if (B.dg) {   // red-black sweep preconditioner (the FP32 form); p alternates between two buffers
			if (it & 1) { V.p = B.p2; V.p2 = B.p; }
			hipLaunchKernelGGL((kb_p<T, IT, true>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it, tol, maxit);
			hipLaunchKernelGGL((kb_v<T, IT, true>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
			hipLaunchKernelGGL((kb_s<T, IT, true>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
			hipLaunchKernelGGL((kb_t<T, IT, true>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it, tol);
			hipLaunchKernelGGL((kb_xr<T, IT, true>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
			return;
		}
		hipLaunchKernelGGL((kb_p<T, IT, false>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it, tol, maxit);
		hipLaunchKernelGGL((kb_v<T, IT, false>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
		hipLaunchKernelGGL((kb_s<T, IT, false>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
		hipLaunchKernelGGL((kb_t<T, IT, false>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it, tol);
		hipLaunchKernelGGL((kb_xr<T, IT, false>), dim3(nb), dim3(256), 0, st, g, V, B.ks, P, it);
	});
}
void launch_krylov_setup(const Geo& g, const KrylovBufs& B, const double* b, hipStream_t st) { krylov_setup_t<double>(g, B, b, st); }
void launch_krylov_iter(const Geo& g, const KrylovBufs& B, hipStream_t st, int it, double tol, int maxit) { krylov_iter_t<double>(g, B, st, it, tol, maxit); }
void launch_krylov_setup32(const Geo& g, const KrylovBufs32& B, const float* b, hipStream_t st) { krylov_setup_t<float>(g, B, b, st); }
void launch_krylov_setup32_cvt(const Geo& g, const KrylovBufs32& B, const double* J64, float* J32, const double* b64, float* b32, double* dP,
                               hipStream_t st) {
	const int items = items_of<float>(g.n), nb = (g.n + 256 * items - 1) / (256 * items);
	KP P{B.part, B.stride, nb};
	KV<float> V{B.x, B.r, B.r0, B.p, B.v, B.s, B.t, B.J, B.ph, B.sh, B.dg, B.p2};
	with_items(items, [&](auto I) {
		constexpr int IT = decltype(I)::value;
		hipLaunchKernelGGL((kbr_setup1_cvt<IT>), dim3(nb), dim3(256), 0, st, g, J64, J32, b64, b32, dP, V, P);
		hipLaunchKernelGGL((kbr_black<float, IT>), dim3(nb), dim3(256), 0, st, g, (const float*)b32, B.x, B.J, (const float*)B.dg, B.ks, 1);
		hipLaunchKernelGGL((kbr_setup2<float, IT>), dim3(nb), dim3(256), 0, st, g, (const float*)b32, V, B.ks, P);
	});
}
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
	hipLaunchKernelGGL(k_scales, dim3(nblk(g.NC)), dim3(256), 0, st, g, D.ph, Qe, D.vsc, D.esc, D.vu, D.vv, D.flags);
	if (D.variant != VAR_PC) { launch_explicit_var(D, limiter, order, Qe, Qs, adt, st); return; }
	if (g.twoD) {
		constexpr int TX = 32, TY = 8;
		const int tx = (g.NXT + TX - 1) / TX, ty = (g.NYT + TY - 1) / TY;
		hipLaunchKernelGGL((k_explicit_2d<TX, TY>), dim3(tx * ty), dim3(TX * TY), 0, st, g, D.ph, limiter, order, adt / g.dx, adt / g.dy, Qe, Qs,
		                   (const double*)D.vsc, (const double*)D.esc, (const double*)D.vu, (const double*)D.vv, D.donx, D.dony, D.phx, D.phy,
		                   D.Ust, D.a2s, D.ubx, D.uby, ty);
	} else {
		constexpr int TX = 256;
		const int tx = (g.NXT + TX - 1) / TX, ty = g.NYT;
		hipLaunchKernelGGL((k_explicit_1d<TX>), dim3(tx * ty), dim3(TX), 0, st, g, D.ph, limiter, order, adt / g.dx, Qe, Qs,
		                   (const double*)D.vsc, (const double*)D.esc, (const double*)D.vu, (const double*)D.vv, D.donx, D.phx,
		                   D.Ust, D.a2s, D.ubx, ty);
	}
	launch_bc_grid(g, D.Ust, st);
	launch_bc_array(g, D.a2s, st);
}
void launch_recon_full(const Dev& D, int limiter, int order, const double* Qe, double* wxL, double* wxR, double* wyL, double* wyR, hipStream_t st) {
	const Geo& g = D.g;
	hipLaunchKernelGGL(k_recon, dim3(nblk(g.NC)), dim3(256), 0, st, g, D.ph, limiter, order, Qe, (const double*)D.vsc, (const double*)D.esc,
	                   wxL, wxR, wyL, wyR, D.phx, D.phy);
}
// finite mu: the frozen acoustic impedances Z_k = rho_k c_k of each CL_MU cell's stage content per unit volume
// (PC:956-958), 1 elsewhere
__global__ void __launch_bounds__(256) k_mu_impedance(Geo g, Phys P, const double* __restrict__ Xs, const double* __restrict__ a2s,
                                                      const int* __restrict__ ctype, double* __restrict__ Z) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.NC) return;
	const int i = c % g.NXT, j = c / g.NXT, NC = g.NC;
	double z0 = 1.0, z1 = 1.0;
	if (i >= GH && i < g.nx + GH && j >= GH && j < g.ny + GH && ctype[(i - GH) + (j - GH) * g.nx] == CL_MU) {
		const double a1 = Xs[IA1 * NC + c], a2 = a2s[c], sa = a1 + a2;
		if (sa > 0) {
			const double cc[6] = {a1 / sa, a2 / sa, Xs[IM1 * NC + c] / sa, Xs[IM2 * NC + c] / sa, Xs[IE1 * NC + c] / sa, Xs[IE2 * NC + c] / sa};
			double zz[2];
			for (int k = 0; k < 2; k++) {
				const double a = cc[k], m = cc[2 + k], pk = p_ame(P, k, a, m, cc[4 + k]), r = m / a;
				zz[k] = r * sqrt(smax(c2_rp(P, k, r, pk), 1e-30));
			}
			z0 = zz[0]; z1 = zz[1];
		}
	}
	Z[c] = z0; Z[NC + c] = z1;
}
void launch_prep(const Dev& D, int limiter, int order, double tau, const double* Pst, hipStream_t st) {
	const Geo& g = D.g;
	if (D.Z) hipLaunchKernelGGL(k_mu_impedance, dim3(nblk(g.NC)), dim3(256), 0, st, g, D.ph, D.Ust, D.a2s, D.ctype, D.Z);
	hipLaunchKernelGGL(k_prep_cell, dim3(nblk(g.NC)), dim3(256), 0, st, g, D.ph, D.Ust, D.a2s, Pst, D.ust, D.vst, D.cz, D.Kw, D.Pold, D.flags);
	hipLaunchKernelGGL(k_prep_face, dim3(nblk(g.NC)), dim3(256), 0, st, g, order, limiter, tau, D.Ust, D.ust, D.vst, D.cz, D.Pold, D.phx, D.phy,
	                   D.ubsx, D.kpx, D.wLx, D.pvx, D.ubsy, D.kpy, D.wLy, D.pvy);
}
int shock_strong_blocks_per_cu() {
	int b = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void*>(k_shock_strong<false>), 256, 0) != hipSuccess || b < 1) b = 2;
	return b;
}
// zero n ints (a kernel, so that it can be captured in a graph, where memset nodes are avoided)
__global__ void __launch_bounds__(256) k_zero_ints(int* __restrict__ p, long n) {
	for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) p[k] = 0;
}
void launch_zero_ints(int* p, long n, hipStream_t st) {
	hipLaunchKernelGGL(k_zero_ints, dim3(std::max(1, std::min(nblk(n), 64))), dim3(256), 0, st, p, n);
}
void launch_shock(const Dev& D, int* list, int* count, double* rec, unsigned long long* rk_steps, int strong_blocks, int fused, hipStream_t st) {
	launch_zero_ints(count, 2, st);   // the count and the strong pass's cursor
	hipLaunchKernelGGL(k_shock_classify, dim3(nblk(D.g.n)), dim3(256), 0, st, D.g, D.ph, D.Ust, D.Qph, list, count, rec, (long)D.g.n);
	const dim3 grid(std::max(1, std::min(nblk(D.g.n), strong_blocks)));
	if (fused) hipLaunchKernelGGL(k_shock_strong<true>, grid, dim3(256), 0, st, D.g, D.ph, (const double*)rec, (long)D.g.n, D.Qph,
	                              (const int*)list, (const int*)count, count + 1, rk_steps);
	else hipLaunchKernelGGL(k_shock_strong<false>, grid, dim3(256), 0, st, D.g, D.ph, (const double*)rec, (long)D.g.n, D.Qph,
	                        (const int*)list, (const int*)count, count + 1, rk_steps);
}
void launch_residual(const Dev& D, const double* P, const double* dP, double lam, double* Pout, double* G, unsigned long long* nrm, int flag_slot, hipStream_t st,
                     const int* gate) {
	if (general_path(D)) { launch_residual_var(D, P, dP, lam, Pout, G, nrm, flag_slot, st, gate); return; }
	hipLaunchKernelGGL(k_residual, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, dP, lam, Pout, G, nrm, D.flags, flag_slot, gate);
}
void launch_residual_list(const Dev& D, const double* P, const int* list, const int* count, double* G, unsigned long long* nrm, int flag_slot,
                          int blocks, hipStream_t st) {
	hipLaunchKernelGGL(k_residual_list, dim3(std::max(1, std::min(blocks, nblk(D.g.n)))), dim3(256), 0, st, D, P, list, count, G, nrm, D.flags, flag_slot);
}
void launch_spec_commit(long n, const float* x, double* dP, const KS* ks, const double* dsum, double eta, int* spec, unsigned long long* nrm, int* flags,
                        hipStream_t st) {
	hipLaunchKernelGGL(k_spec_commit, dim3(nblk(n)), dim3(256), 0, st, n, x, dP, ks, dsum, eta, spec, nrm, flags);
}
void launch_jacobian(const Dev& D, const double* P, const double* G, double* sc, double* J, double* rhs, bool analytic, hipStream_t st, int* clear,
                     int nclear) {
	hipLaunchKernelGGL(k_fdscale, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, sc, clear, nclear);
	if (general_path(D)) launch_jacobian_var(D, P, G, sc, J, rhs, st);
	else if (analytic) hipLaunchKernelGGL(k_jacobian_an, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, G, J, rhs, D.flags);
	else hipLaunchKernelGGL(k_jacobian, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, G, (const double*)sc, J, rhs, D.flags);
}
void launch_chord_rhs(const Dev& D, const double* P, const double* G, double* sc, double* rhs, hipStream_t st, int* clear, int nclear) {
	hipLaunchKernelGGL(k_chord_rhs, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, G, sc, rhs, clear, nclear);
}
void launch_stepcrit(int n, const double* dP, const double* P, const double* sc, unsigned long long* nrm, hipStream_t st) {
	hipLaunchKernelGGL(k_stepcrit, dim3(nblk(n)), dim3(256), 0, st, n, dP, P, sc, nrm);
}
void launch_floor(const Dev& D, const double* P, double tol, hipStream_t st) {
	if (general_path(D)) { launch_floor_var(D, P, tol, st); return; }
	hipLaunchKernelGGL(k_floor, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, tol, D.flags);
}
void launch_final(const Dev& D, const double* P, const double* Qs, double adt, double* H, unsigned long long* tmax, hipStream_t st, int flag_slot) {
	if (general_path(D)) { launch_final_var(D, P, Qs, adt, H, tmax, st, flag_slot); return; }
	hipLaunchKernelGGL(k_final, dim3(nblk(D.g.n)), dim3(256), 0, st, D, P, Qs, adt, H, D.flags, flag_slot, tmax);
}
void launch_cell_diag(const Geo& g, const Phys& P, const double* U, double* out, hipStream_t st) {
	hipLaunchKernelGGL(k_cell_diag, dim3(nblk(g.n)), dim3(256), 0, st, g, P, U, out);
}
void launch_caldt(const Geo& g, const Phys& P, const double* U, int acoustic, unsigned long long* nrm, hipStream_t st) {
	hipLaunchKernelGGL(k_caldt, dim3(nblk(g.n)), dim3(256), 0, st, g, P, U, acoustic, nrm);
}
void launch_thomas(int n, const double* J, const double* rhs, double* x, double* cp, double* dp, hipStream_t st) {
	hipLaunchKernelGGL(k_thomas, dim3(1), dim3(64), 0, st, n, J, rhs, x, cp, dp);
}
}   // namespace spb
