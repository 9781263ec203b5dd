// spb6_mlp.hip -- learned Newton initial guess: features, FP32 training kernels, BF16 MFMA inference (gfx950).
//
// Network: 175 -> 128 (GELU) -> 128 (GELU) -> 1. Input = the 5x5 stencil of 7 per-cell fields of the stage (explicit
// predictor X*, plain guess P, frozen sound speed), expressed relative to the centre cell and standardised; output =
// the standardised pressure change (P_converged - P_plain) / K_w of the centre cell. Training runs in FP32 with
// hand-written kernels (tiled SGEMM, GELU forward/backward, bias sums, Adam); inference runs the two hidden layers on
// MFMA (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation), persistent blocks with the weights in LDS.
#include <hip/hip_runtime.h>
#include "spb6_device.h"
#include "spb6_kernels.h"
#include "spb6_cell.h"
#include "spb6_mlp.h"
#include <algorithm>

namespace spb {

// ------------------------------------------------------------------------------------------------ features
__device__ __forceinline__ int mapP(const Geo& g, int i, int j) {   // M_P (PC:688-697)
	int I, J;
	if (i < GH) I = g.xper ? g.nx + i : GH;
	else if (i > g.nx + GH - 1) I = g.xper ? i - g.nx : g.nx + GH - 1;
	else I = i;
	if (j < GH) J = (g.yper && g.ny > 1) ? g.ny + j : GH;
	else if (j > g.ny + GH - 1) J = (g.yper && g.ny > 1) ? j - g.ny : g.ny + GH - 1;
	else J = j;
	return (I - GH) + g.nx * (J - GH);
}
// per-cell raw fields over the ghosted grid: F[f*NC + c], f = alpha1, log rho, u*, v*, pbar(X*) - P_REF, plain P,
// log(1 + tau c_frozen / dx)
__device__ __forceinline__ void feat_at(const Geo& g, const Phys& P, double P_REF, const double* __restrict__ Xs, const double* __restrict__ Pg,
                                        const double* __restrict__ ust, const double* __restrict__ vst, const double* __restrict__ cz,
                                        float tau_dx, float* __restrict__ F, int c) {
	const int NC = g.NC, i = c % g.NXT, j = c / g.NXT;
	const double a1 = Xs[IA1 * NC + c], m1 = Xs[IM1 * NC + c], m2 = Xs[IM2 * NC + c], rho = m1 + m2;
	const double pg = Pg[mapP(g, i, j)];
	double pb = pg;
	if (a1 > 0 && a1 < 1 && m1 > 0 && m2 > 0 && rho > 0) {
		const double mu = Xs[IMU * NC + c], mv = Xs[IMV * NC + c];
		const double pm = p_mix(P, a1, m1, m2, Xs[IEN * NC + c] - 0.5 * (mu * mu + mv * mv) / rho) - P_REF;
		if (isfin(pm)) pb = pm;
	}
	const double cf = isfin(cz[c]) ? cz[c] : 0.0;
	F[0 * NC + c] = (float)a1;
	F[1 * NC + c] = (float)(rho > 0 ? log(rho) : -50.0);
	F[2 * NC + c] = (float)ust[c];
	F[3 * NC + c] = (float)vst[c];
	F[4 * NC + c] = (float)pb;
	F[5 * NC + c] = (float)pg;
	F[6 * NC + c] = log1pf(tau_dx * (float)cf);   // acoustic CFL of the cell, log-compressed
}
__global__ void __launch_bounds__(256) k_feat_cell(Geo g, Phys P, double P_REF, const double* __restrict__ Xs, const double* __restrict__ Pg,
                                                   const double* __restrict__ ust, const double* __restrict__ vst, const double* __restrict__ cz,
                                                   float tau_dx, float* __restrict__ F) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < g.NC) feat_at(g, P, P_REF, Xs, Pg, ust, vst, cz, tau_dx, F, c);
}
// the same fields on the cells the inference stages for the listed tiles only: a tile's 16 cells lie in one grid row
// (nx a multiple of 16), and their 5x5 stencils span the 20 x 5 ghosted cells from (i0, j0), the tile's first cell in
// interior coordinates; overlapping tiles write the same values
__global__ void __launch_bounds__(256) k_feat_tiles(Geo g, Phys P, double P_REF, const double* __restrict__ Xs, const double* __restrict__ Pg,
                                                    const double* __restrict__ ust, const double* __restrict__ vst, const double* __restrict__ cz,
                                                    float tau_dx, float* __restrict__ F, const int* __restrict__ tiles, const int* __restrict__ ntl) {
	const long m = 100L * *ntl;
	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (long)gridDim.x * blockDim.x) {
		const int q0 = tiles[e / 100] * 16, r = (int)(e % 100), dy = r / 20, dx = r - 20 * dy;
		feat_at(g, P, P_REF, Xs, Pg, ust, vst, cz, tau_dx, F, (q0 % g.nx + dx) + (q0 / g.nx + dy) * g.NXT);
	}
}
struct FeatCtx {
	Geo g; const float* F; const double* Kw; const double* cz; double tau_dx;
	// activity mask: a cell takes part (is dumped, or takes the network's guess) iff |G_q| / (|P_REF + P_q| + K_q + 1)
	// >= thr for the plain guess's residual G and pressure P; G == nullptr marks every cell active
	const double* G; const double* Pplain; double P_REF; double thr;
	const double* thr_dev;   // when given, thr from device memory (k_guess_thr) instead
};
__device__ __forceinline__ bool cell_active(const FeatCtx& X, int q) {
	if (!X.G) return true;
	const double K = X.Kw[(q % X.g.nx + GH) + (q / X.g.nx + GH) * X.g.NXT];
	const double thr = X.thr_dev ? *X.thr_dev : X.thr;
	return fabs(X.G[q]) >= thr * (fabs(X.P_REF + X.Pplain[q]) + K + 1.0);
}
// the activity threshold from the plain guess's residual on the device: GUESS_ACTIVE x its max-norm (the maximum of
// the 64 slots the residual kernel fills with non-negative bit patterns), 0 if that residual failed (flag set); also
// zeroes the list counters of the passes that follow (c0, c1 when given)
__global__ void __launch_bounds__(64) k_guess_thr(const unsigned long long* __restrict__ nrm, const int* __restrict__ flag, double* __restrict__ thr,
                                                  int* __restrict__ c0, int* __restrict__ c1) {
	unsigned long long m = nrm[threadIdx.x];
	#pragma unroll
	for (int off = 32; off >= 1; off >>= 1) { const unsigned long long o = __shfl_xor(m, off, 64); m = m > o ? m : o; }
	if (threadIdx.x == 0) {
		*thr = *flag ? 0.0 : GUESS_ACTIVE * __longlong_as_double((long long)m);
		if (c0) *c0 = 0;
		if (c1) *c1 = 0;
	}
}
// centre-cell values a row's features are taken relative to
struct Centre { float iK, icz, F1c, F2c, F3c, F5c; };
__device__ __forceinline__ Centre centre_of(const FeatCtx& X, int q) {
	const Geo& g = X.g; const int c = (q % g.nx + GH) + (q / g.nx + GH) * g.NXT, NC = g.NC;
	Centre C;
	C.iK = 1.f / (float)X.Kw[c]; const double cz = X.cz[c]; C.icz = 1.f / (float)(isfin(cz) && cz > 0 ? cz : 1.0);
	C.F1c = X.F[NC + c]; C.F2c = X.F[2 * NC + c]; C.F3c = X.F[3 * NC + c]; C.F5c = X.F[5 * NC + c];
	return C;
}
// feature of field f from the field value v at a stencil cell: alpha1 and log(1 + tau c / dx) as they are, log rho
// relative to the centre, velocities relative to the centre in units of the centre's frozen sound speed, pressures
// relative to the centre's plain guess in units of the centre's K_w
__device__ __forceinline__ float feat_tr(int f, float v, const Centre& C) {
	switch (f) {
	case 1: return v - C.F1c;
	case 2: return (v - C.F2c) * C.icz;
	case 3: return (v - C.F3c) * C.icz;
	case 4: case 5: return (v - C.F5c) * C.iK;
	default: return v;
	}
}
// raw feature k (0..174) of interior row q: stencil offset s = k / 7 (5x5, row-major from (-2,-2)), field f = k % 7
__device__ __forceinline__ float feature(const FeatCtx& X, int q, int k, const Centre& C) {
	const Geo& g = X.g;
	const int s = k / 7, f = k - 7 * s, dxo = s % 5 - 2, dyo = s / 5 - 2;
	const int c = (q % g.nx + GH + dxo) + (q / g.nx + GH + dyo) * g.NXT;
	return feat_tr(f, X.F[f * g.NC + c], C);
}
// training dump: raw features of interior rows q = off, off + stride, ... and the target (P_conv - P_plain) / K_w
__global__ void __launch_bounds__(256) k_dump(FeatCtx X, const double* __restrict__ Pconv, const double* __restrict__ Pplain, int off, int stride,
                                              long cap, float* __restrict__ Xout, float* __restrict__ Yout, unsigned long long* __restrict__ count) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int q = off + t * stride;
	__shared__ unsigned long long base;
	const bool act = q < X.g.n && cell_active(X, q);
	// one counter update per block reserves the block's rows
	__shared__ int nact;
	if (threadIdx.x == 0) nact = 0;
	__syncthreads();
	int slot = -1;
	if (act) slot = atomicAdd(&nact, 1);
	__syncthreads();
	if (threadIdx.x == 0) base = atomicAdd(count, (unsigned long long)nact);
	__syncthreads();
	if (!act) return;
	const long row = (long)base + slot;
	if (row >= cap) return;
	const Centre C = centre_of(X, q);
	for (int k = 0; k < MLP_IN; k++) Xout[row * MLP_IN + k] = feature(X, q, k, C);
	Yout[row] = (float)((Pconv[q] - Pplain[q]) / X.Kw[(q % X.g.nx + GH) + (q / X.g.nx + GH) * X.g.NXT]);
}

// ------------------------------------------------------------------------------------------------ FP32 training kernels
// C[M x N] = op(A)[M x K] op(B)[K x N] (+ beta C): op(A)(m, k) = ta ? A[k*lda + m] : A[m*lda + k], likewise B.
// 64x64 output tiles, K chunks of 16 staged in LDS, 256 threads with a 4x4 register tile each.
__global__ void __launch_bounds__(256) k_sgemm(int M, int N, int K, const float* __restrict__ A, int lda, int ta,
                                              const float* __restrict__ B, int ldb, int tb, float* __restrict__ C, int ldc, float beta) {
	__shared__ float As[16][64 + 4], Bs[16][64 + 4];
	const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
	const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
	float acc[4][4] = {};
	for (int k0 = 0; k0 < K; k0 += 16) {
		for (int e = threadIdx.x; e < 16 * 64; e += 256) {
			const int kk = e / 64, mm = e % 64;
			const int m = m0 + mm, k = k0 + kk;
			As[kk][mm] = (m < M && k < K) ? (ta ? A[(long)k * lda + m] : A[(long)m * lda + k]) : 0.f;
			const int n = n0 + mm;
			Bs[kk][mm] = (n < N && k < K) ? (tb ? B[(long)n * ldb + k] : B[(long)k * ldb + n]) : 0.f;
		}
		__syncthreads();
		#pragma unroll
		for (int kk = 0; kk < 16; kk++) {
			float a[4], b[4];
			#pragma unroll
			for (int r = 0; r < 4; r++) { a[r] = As[kk][ty * 4 + r]; b[r] = Bs[kk][tx * 4 + r]; }
			#pragma unroll
			for (int r = 0; r < 4; r++)
				#pragma unroll
				for (int s = 0; s < 4; s++) acc[r][s] += a[r] * b[s];
		}
		__syncthreads();
	}
	#pragma unroll
	for (int r = 0; r < 4; r++)
		#pragma unroll
		for (int s = 0; s < 4; s++) {
			const int m = m0 + ty * 4 + r, n = n0 + tx * 4 + s;
			if (m < M && n < N) { const long o = (long)m * ldc + n; C[o] = acc[r][s] + (beta != 0.f ? beta * C[o] : 0.f); }
		}
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
// GELU in its tanh form x sigma(2u), u = sqrt(2/pi) (x + 0.044715 x^3) (Hendrycks and Gimpel), for the BF16 inference:
// within 4.8e-4 of the erf form (below the BF16 step of the activations) with one native exp2 and one reciprocal;
// the constants are -2 log2(e) sqrt(2/pi) and 0.044715 times that
__device__ __forceinline__ float gelu_fast(float x) {
	return x * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * (-2.3022081981f - 0.10294323958f * x * x)));
}
__device__ __forceinline__ float dgelu(float x) { return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x); }
// Z += bias (per column), H = GELU(Z)
__global__ void __launch_bounds__(256) k_bias_gelu(long rows, int N, float* __restrict__ Z, const float* __restrict__ b, float* __restrict__ H) {
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= rows * N) return;
	const float z = Z[e] + b[e % N];
	Z[e] = z; H[e] = gelu(z);
}
// y = H w + b (one output), loss partial and dy = 2 (y - t) / B
__global__ void __launch_bounds__(256) k_head(int Bn, const float* __restrict__ H, const float* __restrict__ w, const float* __restrict__ bp,
                                              const float* __restrict__ t, float* __restrict__ y, float* __restrict__ dy, float* __restrict__ loss) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	float l = 0.f;
	if (r < Bn) {
		float s = *bp;
		for (int k = 0; k < MLP_H; k++) s += H[(long)r * MLP_H + k] * w[k];
		y[r] = s;
		if (t) { const float d = s - t[r]; l = d * d; if (dy) dy[r] = 2.f * d / (float)Bn; }
	}
	if (loss) {
		for (int off = 32; off >= 1; off >>= 1) l += __shfl_xor(l, off, 64);
		if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
	}
}
// dZ2 = (dy w^T) * GELU'(Z2), and the head gradients dw = H^T dy, db = sum dy
__global__ void __launch_bounds__(256) k_head_back(int Bn, const float* __restrict__ dy, const float* __restrict__ w, const float* __restrict__ Z,
                                                   float* __restrict__ dZ) {
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)Bn * MLP_H) return;
	const int r = (int)(e / MLP_H), k = (int)(e % MLP_H);
	dZ[e] = dy[r] * w[k] * dgelu(Z[e]);
}
__global__ void __launch_bounds__(256) k_mul_dgelu(long n, float* __restrict__ dH, const float* __restrict__ Z) {
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e < n) dH[e] *= dgelu(Z[e]);
}
// column sums of a rows x N matrix (bias gradients), one block per column, fixed order
__global__ void __launch_bounds__(256) k_colsum(int rows, int N, const float* __restrict__ A, float* __restrict__ out) {
	__shared__ float sh[256];
	const int col = blockIdx.x;
	float s = 0.f;
	for (int r = threadIdx.x; r < rows; r += 256) s += A[(long)r * N + col];
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int h = 128; h > 0; h >>= 1) { if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h]; __syncthreads(); }
	if (threadIdx.x == 0) out[col] = sh[0];
}
// Adam (Kingma & Ba 2015) with bias correction
__global__ void __launch_bounds__(256) k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                                              float lr, float b1, float b2, float eps, float c1, float c2) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
	m[i] = mi; v[i] = vi;
	p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
__global__ void __launch_bounds__(256) k_gather_rows(int Bn, int K, const float* __restrict__ X, const float* __restrict__ Y, const int* __restrict__ idx,
                                                     float* __restrict__ Xb, float* __restrict__ Yb) {
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)Bn * K) return;
	const int r = (int)(e / K), k = (int)(e % K);
	const long src = idx[r];
	Xb[e] = X[src * K + k];
	if (k == 0) Yb[r] = Y[src];
}

// ------------------------------------------------------------------------------------------------ BF16 MFMA inference
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
union BF8 { bf16x8 v; __bf16 e[8]; uint4 u; };

// MFMA k order of layer 1: slot t = 32 ks + 8 g + e (k step ks, lane group g = lane / 16, element e of the lane's 8)
// holds field f = (8 ks + e) / 7 at stencil point s = 4 ((8 ks + e) % 7) + g (s < 25). The field is then a compile-time
// constant of the unrolled (ks, e) loops and only the stencil offset varies by lane. f = 7 and s >= 25 are padding;
// slot MLP_BIAS_SLOT (ks 6, g 0, e 1: a field-7 slot) takes the constant input 1 and carries layer 1's bias.
constexpr int MLP_BIAS_SLOT = 6 * 32 + 1;
__host__ __device__ constexpr int slot_feature(int t) {
	const int ks = t / 32, lk = (t / 8) % 4, e = t % 8, p = 8 * ks + e, f = p / 7, s = 4 * (p % 7) + lk;
	return (f < 7 && s < 25) ? 7 * s + f : -1;
}
// layer-2 k order: element s of lane group g in k step j is hidden unit h2k(j, g, s), i.e. the rows 4g .. 4g + 3 of
// layer 1's output tiles 2j and 2j + 1 that the lane already holds, so layer 1's results feed layer 2 from registers
__host__ __device__ constexpr int h2k(int j, int g, int s) { return 32 * j + (s < 4 ? 4 * g + s : 16 + 4 * g + (s - 4)); }
// Folded standardisation for the stage's features: W1 (x - mean) istd + b1 = (W1 istd) (x - mu_f) + b1f with the
// per-field shift mu_f (the mean of field f's slot means) and b1f[n] = b1[n] + sum_k W1[k][n] (mu_f(k) - mean_k) istd_k.
// The BF16 input is then x - mu_f (the field centred, as precise as the standardised value); mub1 = (mu[7], b1f[128]).
__global__ void __launch_bounds__(MLP_H) k_pack_fold(const float* __restrict__ params, const float* __restrict__ mean,
                                                     const float* __restrict__ istd, float* __restrict__ mub1) {
	__shared__ float mu[7];
	const int n = threadIdx.x;
	if (n < 7) {
		double s = 0.0;
		for (int sp = 0; sp < 25; sp++) s += mean[7 * sp + n];
		mu[n] = (float)(s / 25.0);
		mub1[n] = mu[n];
	}
	__syncthreads();
	double b = params[MLP_OFF_B1 + n];
	for (int k = 0; k < MLP_IN; k++) b += (double)params[MLP_OFF_W1 + k * MLP_H + n] * ((double)mu[k % 7] - (double)mean[k]) * (double)istd[k];
	mub1[7 + n] = (float)b;
}
// transposed BF16 weights, one lane's 8 consecutive k = 16 bytes: W1s[n][t] for standardised inputs and W1f[n][t] for
// inputs x - mu_f (istd folded in), each with its bias in MLP_BIAS_SLOT; W2p[m][32 j + 8 g + s] = W2[h2k(j, g, s)][m]
__global__ void __launch_bounds__(256) k_pack_bf16(const float* __restrict__ params, const float* __restrict__ istd, const float* __restrict__ mub1,
                                                   __bf16* __restrict__ W1s, __bf16* __restrict__ W1f, __bf16* __restrict__ W2p) {
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < MLP_H * MLP_KPAD) {
		const int n = e / MLP_KPAD, t = e % MLP_KPAD, k = slot_feature(t);
		float ws = 0.f, wf = 0.f;
		if (k >= 0) { ws = params[MLP_OFF_W1 + k * MLP_H + n]; wf = ws * istd[k]; }
		else if (t == MLP_BIAS_SLOT) { ws = params[MLP_OFF_B1 + n]; wf = mub1[7 + n]; }
		W1s[e] = (__bf16)ws; W1f[e] = (__bf16)wf;
	}
	if (e < MLP_H * MLP_H) {
		const int m = e / MLP_H, kk = e % MLP_H;
		W2p[e] = (__bf16)params[MLP_OFF_W2 + h2k(kk / 32, (kk / 8) % 4, kk % 8) * MLP_H + m];
	}
}
// MFMA 16x16x32 operand layout: lane l holds A rows i = l % 16 with k = 8 (l / 16) .. +7 of a 32-wide k step, B columns
// j = l % 16 with the same k; the 16x16 FP32 result gives lane l column j = l % 16, rows 4 (l / 16) .. +3.
// Both layers run transposed, Z^T = W^T X^T: the weights are the A operand (rows = hidden units) and a 16-cell tile
// is the B operand (columns = cells), so lane l works on cell l % 16 throughout and layer 1's FP32 results, after GELU
// and BF16 rounding, are layer 2's B operand in place. Persistent blocks of 16 waves share the BF16 weights in LDS;
// each wave takes 16-cell tiles. mode 0: features from the stage (guess); mode 1: standardised features given in
// Xin[rows][175] (tests).
constexpr int INF_WAVES = 16;
// first pass of the learned guess: every cell gets the plain guess, and the 16-cell tiles with an active cell are
// listed (in any order: a tile's result depends only on its own cells) for the network pass, which then takes them
// in turn from the list. Active cells are a few percent of a grid and cluster (the shock, the bubble), so taking the
// tiles in grid order would leave most wavefronts idle while a few run many active tiles.
// With rlist (the learned guess's residual evaluated only where it can differ from the plain guess's): a cell is
// touched when it or one of its four neighbours (those its residual reads P at) is active, i.e. changes its pressure.
// An untouched cell's residual inputs are the plain guess's, so its residual is the plain one, bit for bit: it is
// copied into Gout, its norm term |G| / (|P_REF + P| + K + 1) taken into nrm2 (and a failed closure, G = NaN, into
// *flag2) as the residual kernel would; touched cells are listed in rlist for k_residual_list.
__global__ void __launch_bounds__(256) k_guess_tiles(FeatCtx X, int nrows, const double* __restrict__ Pplain, double* __restrict__ Pout,
                                                     int* __restrict__ list, int* __restrict__ count, double* __restrict__ Gout,
                                                     int* __restrict__ rlist, int* __restrict__ rcount, unsigned long long* __restrict__ nrm2,
                                                     int* __restrict__ flag2) {
	const int q = blockIdx.x * 256 + threadIdx.x, lane = threadIdx.x & 63;
	const bool valid = q < nrows;
	if (valid) Pout[q] = Pplain[q];
	const bool act = valid && cell_active(X, q);
	if (rlist) {
		bool touched = false;
		double nv = 0.0;
		if (valid) {
			const Nb b = neighbours(X.g, q);
			touched = act || cell_active(X, b.W) || cell_active(X, b.E) || (X.g.twoD && (cell_active(X, b.S) || cell_active(X, b.N)));
			if (!touched) {
				const double gq = X.G[q];
				Gout[q] = gq;
				if (isfin(gq)) nv = fabs(gq) / (fabs(X.P_REF + X.Pplain[q]) + X.Kw[(q % X.g.nx + GH) + (q / X.g.nx + GH) * X.g.NXT] + 1.0);
				else *flag2 = 1;
			}
		}
		block_append(touched, q, rlist, rcount);
		atomic_max_nonneg(nrm2, nv);
	}
	const unsigned long long m = __ballot(act);
	__shared__ int nt, base, tl[16];
	if (threadIdx.x == 0) nt = 0;
	__syncthreads();
	if ((lane & 15) == 0 && ((m >> lane) & 0xFFFFull)) tl[atomicAdd(&nt, 1)] = q >> 4;
	__syncthreads();
	if (threadIdx.x == 0) base = nt ? atomicAdd(count, nt) : 0;
	__syncthreads();
	if (threadIdx.x < nt) list[base + threadIdx.x] = tl[threadIdx.x];
}
__global__ void __launch_bounds__(64 * INF_WAVES) k_mlp_infer(FeatCtx X, int nrows, const float* __restrict__ Xin, int mode,
                                                   const float* __restrict__ mu_f, const __bf16* __restrict__ W1, const __bf16* __restrict__ W2p,
                                                   const float* __restrict__ b2, const float* __restrict__ w3, float b3, float mu_t, float sd_t,
                                                   const double* __restrict__ Pplain, double* __restrict__ Pout, float* __restrict__ yout,
                                                   const int* __restrict__ tiles, const int* __restrict__ ntl) {
	__shared__ uint4 sW1[MLP_H * MLP_KPAD / 8];     // 56 KB
	__shared__ uint4 sW2[MLP_H * MLP_H / 8];        // 32 KB
	__shared__ float4 sB2[MLP_H / 4], sW3[MLP_H / 4];
	__shared__ float sF[INF_WAVES][7][5][20];       // per-wave 5x5 stencils of a 16-cell tile (20 x 5 cells x 7 fields), 44 KB
	for (int e = threadIdx.x; e < MLP_H * MLP_KPAD / 8; e += 64 * INF_WAVES) sW1[e] = ((const uint4*)W1)[e];
	for (int e = threadIdx.x; e < MLP_H * MLP_H / 8; e += 64 * INF_WAVES) sW2[e] = ((const uint4*)W2p)[e];
	for (int e = threadIdx.x; e < MLP_H / 4; e += 64 * INF_WAVES) { sB2[e] = ((const float4*)b2)[e]; sW3[e] = ((const float4*)w3)[e]; }
	__syncthreads();
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, g = lane >> 4;
	float mu[7];
	#pragma unroll
	for (int f = 0; f < 7; f++) mu[f] = mode == 0 ? mu_f[f] : 0.f;
	// this lane group's stencil points s = 4 m + g (m = 0..6) as offsets into a staged tile, and whether they exist
	int soff[7]; bool sin[7];
	#pragma unroll
	for (int m = 0; m < 7; m++) { const int sp = 4 * m + g; sin[m] = sp < 25; soff[m] = sin[m] ? (sp / 5) * 20 + sp % 5 : 0; }
	const int ntiles = tiles ? *ntl : (nrows + 15) / 16;
	const bool staged = mode == 0 && X.g.nx % 16 == 0 && nrows % 16 == 0;
	const float* tileF = &sF[wave][0][0][0] + lr;
	for (int it = blockIdx.x * INF_WAVES + wave; it < ntiles; it += gridDim.x * INF_WAVES) {
		const int tile = tiles ? tiles[it] : it;
		const int q = tile * 16 + lr;           // this lane's cell
		const bool valid = q < nrows;
		// a tile none of whose cells is active keeps the plain guess and skips the network (wave-uniform branch; listed
		// tiles have an active cell)
		if (!tiles && mode == 0 && __ballot(valid && cell_active(X, q)) == 0ull) {
			if (g == 0 && valid && Pout) Pout[q] = Pplain[q];
			continue;
		}
		Centre C{};
		int cq = 0;                              // ghosted index of the lane's cell
		if (staged) {
			// the tile's 16 cells lie in one grid row: stage the 20 x 5 cells x 7 fields of their stencils in LDS with
			// coalesced loads, one round trip, and take every feature from there
			const int q0 = tile * 16, i0 = q0 % X.g.nx, j0 = q0 / X.g.nx;
			for (int e = lane; e < 700; e += 64) {
				const int f = e / 100, rem = e - 100 * f, dy = rem / 20, dx = rem - 20 * dy;
				sF[wave][f][dy][dx] = X.F[f * X.g.NC + (i0 + dx) + (j0 + dy) * X.g.NXT];   // ghosted (i0 - 2 + dx + GH, j0 - 2 + dy + GH)
			}
			cq = (q % X.g.nx + GH) + (q / X.g.nx + GH) * X.g.NXT;
			C.iK = 1.f / (float)X.Kw[cq]; const double cz = X.cz[cq]; C.icz = 1.f / (float)(isfin(cz) && cz > 0 ? cz : 1.0);
			__builtin_amdgcn_s_waitcnt(0);
			__builtin_amdgcn_wave_barrier();
			C.F1c = tileF[100 + 42]; C.F2c = tileF[200 + 42]; C.F3c = tileF[300 + 42]; C.F5c = tileF[500 + 42];   // (dy, dx) = (2, 2)
		}
		else if (valid && mode == 0) { C = centre_of(X, q); cq = (q % X.g.nx + GH) + (q / X.g.nx + GH) * X.g.NXT; }
		// layer 1: Z1^T[128 x 16] = W1^T[128 x 224] X^T[224 x 16]; acc1[nt][r] = hidden unit 16 nt + 4 g + r of cell lr
		f32x4 acc1[8];
		#pragma unroll
		for (int nt = 0; nt < 8; nt++) acc1[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
		#pragma unroll
		for (int ks = 0; ks < MLP_KPAD / 32; ks++) {
			BF8 b;
			#pragma unroll
			for (int e = 0; e < 8; e++) {
				const int p = 8 * ks + e, f = p / 7, m = p % 7;
				float x = 0.f;
				if (f < 7) {
					if (valid && sin[m]) {
						if (staged) x = feat_tr(f, tileF[100 * f + soff[m]], C) - mu[f];
						else if (mode == 0) {
							const int sp = 4 * m + g;
							x = feat_tr(f, X.F[f * X.g.NC + cq + (sp % 5 - 2) + (sp / 5 - 2) * X.g.NXT], C) - mu[f];
						} else x = Xin[(long)q * MLP_IN + 7 * (4 * m + g) + f];
					}
				} else if (ks * 32 + e == MLP_BIAS_SLOT && g == 0) x = 1.f;
				b.e[e] = (__bf16)x;
			}
			#pragma unroll
			for (int nt = 0; nt < 8; nt++) {
				BF8 a; a.u = sW1[((nt * 16 + lr) * MLP_KPAD + ks * 32 + g * 8) / 8];
				acc1[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc1[nt], 0, 0, 0);
			}
		}
		// GELU (the bias came in through MLP_BIAS_SLOT) and BF16: layer 2's B fragments in the h2k order
		BF8 h[4];
		#pragma unroll
		for (int j = 0; j < 4; j++)
			#pragma unroll
			for (int r = 0; r < 4; r++) { h[j].e[r] = (__bf16)gelu_fast(acc1[2 * j][r]); h[j].e[4 + r] = (__bf16)gelu_fast(acc1[2 * j + 1][r]); }
		// layer 2: Z2^T[128 x 16] = W2^T[128 x 128] H1^T[128 x 16]; acc2[mt][r] = hidden unit 16 mt + 4 g + r of cell lr
		f32x4 acc2[8];
		#pragma unroll
		for (int mt = 0; mt < 8; mt++) acc2[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
		#pragma unroll
		for (int j = 0; j < 4; j++)
			#pragma unroll
			for (int mt = 0; mt < 8; mt++) {
				BF8 a; a.u = sW2[((mt * 16 + lr) * MLP_H + j * 32 + g * 8) / 8];
				acc2[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, h[j].v, acc2[mt], 0, 0, 0);
			}
		// head: y = sum_n GELU(Z2[n] + b2[n]) w3[n] + b3 over the lane's 32 units, then over the four lane groups
		float part = 0.f;
		#pragma unroll
		for (int mt = 0; mt < 8; mt++) {
			const float4 bb = sB2[mt * 4 + g], ww = sW3[mt * 4 + g];
			part += gelu_fast(acc2[mt][0] + bb.x) * ww.x + gelu_fast(acc2[mt][1] + bb.y) * ww.y
			      + gelu_fast(acc2[mt][2] + bb.z) * ww.z + gelu_fast(acc2[mt][3] + bb.w) * ww.w;
		}
		part += __shfl_xor(part, 16, 64);
		part += __shfl_xor(part, 32, 64);
		if (g == 0 && valid) {
			const float y = part + b3;
			if (yout) yout[q] = y;
			if (Pout) {
				const Geo& gg = X.g;
				const double Kc = X.Kw[(q % gg.nx + GH) + (q / gg.nx + GH) * gg.NXT];
				Pout[q] = cell_active(X, q) ? Pplain[q] + Kc * (double)(mu_t + sd_t * y) : Pplain[q];
			}
		}
	}
}

// ------------------------------------------------------------------------------------------------ launchers
static inline unsigned nb(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }
void mlp_feat_cell(const Dev& D, const double* Pplain, float* F, hipStream_t st) {
	hipLaunchKernelGGL(k_feat_cell, dim3(nb(D.g.NC)), dim3(256), 0, st, D.g, D.ph, D.P_REF, (const double*)D.Ust, Pplain, (const double*)D.ust,
	                   (const double*)D.vst, (const double*)D.cz, (float)(D.tau / D.g.dx), F);
}
void mlp_dump(const Dev& D, const float* F, double tau, const double* Pconv, const double* Pplain, const double* Gplain, double thr,
              int off, int stride, long cap, float* X, float* Y, unsigned long long* count, hipStream_t st) {
	FeatCtx fx{D.g, F, D.Kw, D.cz, tau / D.g.dx, Gplain, Pplain, D.P_REF, thr, nullptr};
	const long rows = (D.g.n - off + stride - 1) / stride;
	if (rows <= 0) return;
	hipLaunchKernelGGL(k_dump, dim3(nb(rows)), dim3(256), 0, st, fx, Pconv, Pplain, off, stride, cap, X, Y, count);
}
void mlp_pack(const float* params, const float* mean, const float* istd, __bf16* W1s, __bf16* W1f, __bf16* W2p, float* mub1, hipStream_t st) {
	hipLaunchKernelGGL(k_pack_fold, dim3(1), dim3(MLP_H), 0, st, params, mean, istd, mub1);
	hipLaunchKernelGGL(k_pack_bf16, dim3(nb(MLP_H * MLP_KPAD)), dim3(256), 0, st, params, istd, (const float*)mub1, W1s, W1f, W2p);
}
void mlp_infer(const Dev* D, float* F, double tau, int nrows, const float* Xin, const MlpDev& W, const double* Pplain,
               const double* Gplain, double thr, double* Pout, float* yout, hipStream_t st, int* list, int* count,
               const unsigned long long* nrm, const int* flag, double* thr_dev, double* Gout, int* rlist, int* rcount, unsigned long long* nrm2,
               int* flag2, bool feats) {
	FeatCtx fx{};
	if (D) { fx.g = D->g; fx.F = F; fx.Kw = D->Kw; fx.cz = D->cz; fx.tau_dx = tau / D->g.dx; fx.G = Gplain; fx.Pplain = Pplain; fx.P_REF = D->P_REF; fx.thr = thr; }
	if (D && thr_dev) {   // the threshold from the plain residual's norm on the device
		const bool lists = Gplain && Pout && list && count;
		hipLaunchKernelGGL(k_guess_thr, dim3(1), dim3(64), 0, st, nrm, flag, thr_dev, lists ? count : nullptr, lists ? rcount : nullptr);
		fx.thr_dev = thr_dev;
	}
	int dev = 0; (void)hipGetDevice(&dev);
	int cus = 256; (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	const int tiles = (nrows + 15) / 16;
	const int blocks = std::max(1, std::min(cus, (tiles + INF_WAVES - 1) / INF_WAVES));
	const bool listed = D && Gplain && Pout && list && count;   // the guess with an activity mask: listed active tiles
	// feats: the features are computed here, on the listed tiles' staged stencils only where the inference stages the
	// tiles (tiles in one grid row, as in k_mlp_infer), else on every cell first
	const bool staged = D && !Xin && D->g.nx % 16 == 0 && nrows % 16 == 0;
	if (feats && D && !(listed && staged)) mlp_feat_cell(*D, Pplain, F, st);
	if (listed) {
		if (!fx.thr_dev) (void)hipMemsetAsync(count, 0, sizeof(int), st);   // else k_guess_thr has cleared the counters
		const bool sparse = fx.thr_dev && Gout && rlist && rcount && nrm2 && flag2;
		hipLaunchKernelGGL(k_guess_tiles, dim3(nb(nrows)), dim3(256), 0, st, fx, nrows, Pplain, Pout, list, count, Gout, sparse ? rlist : nullptr,
		                   rcount, nrm2, flag2);
		if (feats && staged)
			hipLaunchKernelGGL(k_feat_tiles, dim3(4 * cus), dim3(256), 0, st, D->g, D->ph, D->P_REF, (const double*)D->Ust, Pplain,
			                   (const double*)D->ust, (const double*)D->vst, (const double*)D->cz, (float)(D->tau / D->g.dx), F,
			                   (const int*)list, (const int*)count);
	}
	hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * INF_WAVES), 0, st, fx, nrows, Xin, Xin ? 1 : 0, (const float*)W.mub1,
	                   Xin ? W.W1s : W.W1f, W.W2p, W.params + MLP_OFF_B2, W.params + MLP_OFF_W3, W.b3, W.mu_t, W.sd_t, Pplain, Pout, yout,
	                   listed ? (const int*)list : nullptr, listed ? (const int*)count : nullptr);
}

// forward (and optionally backward) of the FP32 network on a batch Xb[Bn x 175]; grads into G (same layout as params)
void mlp_fwd_bwd(const float* P, const float* Xb, const float* Yb, int Bn, float* ws, float* G, float* yout, float* loss, hipStream_t st) {
	float* Z1 = ws; float* H1 = Z1 + (long)Bn * MLP_H; float* Z2 = H1 + (long)Bn * MLP_H; float* H2 = Z2 + (long)Bn * MLP_H;
	float* dy = H2 + (long)Bn * MLP_H; float* dZ2 = dy + Bn; float* dZ1 = dZ2 + (long)Bn * MLP_H; float* y = dZ1 + (long)Bn * MLP_H;
	const dim3 g1(nb(MLP_H, 64), nb(Bn, 64));
	hipLaunchKernelGGL(k_sgemm, g1, dim3(256), 0, st, Bn, MLP_H, MLP_IN, Xb, MLP_IN, 0, P + MLP_OFF_W1, MLP_H, 0, Z1, MLP_H, 0.f);
	hipLaunchKernelGGL(k_bias_gelu, dim3(nb((long)Bn * MLP_H)), dim3(256), 0, st, (long)Bn, MLP_H, Z1, P + MLP_OFF_B1, H1);
	hipLaunchKernelGGL(k_sgemm, g1, dim3(256), 0, st, Bn, MLP_H, MLP_H, (const float*)H1, MLP_H, 0, P + MLP_OFF_W2, MLP_H, 0, Z2, MLP_H, 0.f);
	hipLaunchKernelGGL(k_bias_gelu, dim3(nb((long)Bn * MLP_H)), dim3(256), 0, st, (long)Bn, MLP_H, Z2, P + MLP_OFF_B2, H2);
	hipLaunchKernelGGL(k_head, dim3(nb(Bn)), dim3(256), 0, st, Bn, (const float*)H2, P + MLP_OFF_W3, P + MLP_OFF_B3, Yb, yout ? yout : y, G ? dy : nullptr, loss);
	if (!G) return;
	// head gradients
	hipLaunchKernelGGL(k_sgemm, dim3(1, nb(MLP_H, 64)), dim3(256), 0, st, MLP_H, 1, Bn, (const float*)H2, MLP_H, 1, (const float*)dy, 1, 0, G + MLP_OFF_W3, 1, 0.f);
	hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, st, Bn, 1, (const float*)dy, G + MLP_OFF_B3);
	hipLaunchKernelGGL(k_head_back, dim3(nb((long)Bn * MLP_H)), dim3(256), 0, st, Bn, (const float*)dy, P + MLP_OFF_W3, (const float*)Z2, dZ2);
	// layer 2 gradients: dW2 = H1^T dZ2, db2 = colsum dZ2, dH1 = dZ2 W2^T
	hipLaunchKernelGGL(k_sgemm, dim3(nb(MLP_H, 64), nb(MLP_H, 64)), dim3(256), 0, st, MLP_H, MLP_H, Bn, (const float*)H1, MLP_H, 1, (const float*)dZ2, MLP_H, 0, G + MLP_OFF_W2, MLP_H, 0.f);
	hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, st, Bn, MLP_H, (const float*)dZ2, G + MLP_OFF_B2);
	hipLaunchKernelGGL(k_sgemm, g1, dim3(256), 0, st, Bn, MLP_H, MLP_H, (const float*)dZ2, MLP_H, 0, P + MLP_OFF_W2, MLP_H, 1, dZ1, MLP_H, 0.f);
	hipLaunchKernelGGL(k_mul_dgelu, dim3(nb((long)Bn * MLP_H)), dim3(256), 0, st, (long)Bn * MLP_H, dZ1, (const float*)Z1);
	// layer 1 gradients: dW1 = X^T dZ1, db1 = colsum dZ1
	hipLaunchKernelGGL(k_sgemm, dim3(nb(MLP_H, 64), nb(MLP_IN, 64)), dim3(256), 0, st, MLP_IN, MLP_H, Bn, Xb, MLP_IN, 1, (const float*)dZ1, MLP_H, 0, G + MLP_OFF_W1, MLP_H, 0.f);
	hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, st, Bn, MLP_H, (const float*)dZ1, G + MLP_OFF_B1);
}
void mlp_adam(float* P, const float* G, float* m, float* v, float lr, int t, hipStream_t st) {
	const float b1 = 0.9f, b2 = 0.999f;
	const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
	hipLaunchKernelGGL(k_adam, dim3(nb(MLP_NPARAM)), dim3(256), 0, st, MLP_NPARAM, P, G, m, v, lr, b1, b2, 1e-8f, c1, c2);
}
void mlp_gather(int Bn, const float* X, const float* Y, const int* idx, float* Xb, float* Yb, hipStream_t st) {
	hipLaunchKernelGGL(k_gather_rows, dim3(nb((long)Bn * MLP_IN)), dim3(256), 0, st, Bn, MLP_IN, X, Y, idx, Xb, Yb);
}

}   // namespace spb
