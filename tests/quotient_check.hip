// GPU self-check of the strong-shock kernel's shared-reciprocal quotients (spb6_device.h: rcp_nr, quot, prof_rhs_rcp,
// prof_step<true>). Built with the library into tests/bin/quotient_check; run by tests/test_gpu_quotients.py.
//
//   quot   quot(a, b, rcp_nr(b)) against the compiler's a / b on random operand pairs whose exponents lie in
//          [-384, 383] (random mantissas, and all-ones, zero and one-ulp mantissas), and on zero numerators (a zero of
//          either sign is accepted there);
//   rhs    prof_rhs_rcp against prof_rhs for four phase pairs on random states at the profile problems' scales and on
//          states built to fail the band test (p = P, p + pi_k or v2 near zero, tiny or huge Y_k, lam_k / Y_k, p, v and
//          random bit patterns over exponents -400..400): wherever the band test passes, the flag and the three
//          derivatives must agree bit for bit;
//   step   prof_step<true> against prof_step<false> over whole integrations, step by step: a 2 must leave the run
//          unchanged, any other return code and the run state must agree bit for bit.
//
// Prints one line per part, "<part> cases N fast F fallback B differ D"; exit status 0 when it ran.
#include "../diffusionimex_amd/csrc/spb6_device.h"

#include <cstdio>
#include <cstring>

using namespace spb;

#define CK(x)                                                                                        \
	do {                                                                                             \
		hipError_t e_ = (x);                                                                         \
		if (e_ != hipSuccess) { std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); return 1; } \
	} while (0)

namespace {

enum { C_CASES, C_FAST, C_FALLBACK, C_DIFFER, C_N };

__device__ __forceinline__ uint64_t mix(uint64_t z) {
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
struct Rng {
	uint64_t s;
	__device__ uint64_t next() { s = mix(s); return s; }
	__device__ double unit() { return (double)(next() >> 11) * 0x1p-53; }                      // [0, 1)
	__device__ double logu(double lo, double hi) { return lo * pow(hi / lo, unit()); }         // log-uniform
	__device__ double bits(int emin, int emax) {                                               // random sign, exponent, mantissa
		const uint64_t r = next();
		const int e = emin + (int)((r >> 52) % (uint64_t)(emax - emin + 1));
		uint64_t m = next() & 0xFFFFFFFFFFFFFull;
		const int kind = (int)((r >> 40) & 15);
		if (kind == 0) m = 0xFFFFFFFFFFFFFull; else if (kind == 1) m = 0; else if (kind == 2) m = 1;
		return __longlong_as_double((long long)(((r >> 63) << 63) | ((uint64_t)(e + 1023) << 52) | m));
	}
};
__device__ __forceinline__ bool same(double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); }
__device__ void count(unsigned long long* c, int k, unsigned long long v) { if (v) atomicAdd(&c[k], v); }

__global__ void k_quot(uint64_t seed, int iters, unsigned long long* c) {
	Rng g{seed ^ (0x100000001B3ull * (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x))};
	unsigned long long n = 0, d = 0;
	for (int k = 0; k < iters; k++) {
		const double b = g.bits(-384, 383);
		const double a = (k & 63) == 0 ? ((k & 64) ? -0.0 : 0.0) : g.bits(-384, 383);
		const double q = quot(a, b, rcp_nr(b)), e = a / b;
		n++;
		if (a == 0.0 ? q != 0.0 : !same(q, e)) d++;
	}
	count(c, C_CASES, n); count(c, C_DIFFER, d);
}

__device__ Phys phases(int k) {
	Phys P{};
	P.cv[0] = P.cv[1] = 1000.0;
	if (k == 0) { P.g[0] = 4.4; P.pi[0] = 6e8; P.g[1] = 1.4; }                          // water | air
	else if (k == 1) { P.g[0] = 1.19; P.pi[0] = 6.2e8; P.b[0] = 6.7e-4; P.g[1] = 1.4; }  // covolume water | air
	else if (k == 2) { P.g[0] = 1.4; P.g[1] = 1.648; }                                   // air | helium
	else { P.g[0] = 1.0 + 0x1p-12; P.g[1] = 1.4; }                                       // g0 - 1 out of band 8
	return P;
}
// one profile problem at the recorded problems' scales
__device__ void problem(Rng& g, Prof& S, double& lr0, double& lr1) {
	S.Y0 = g.logu(1e-12, 1.0); S.Y1 = 1.0 - S.Y0 > 0 ? 1.0 - S.Y0 : 1e-12;
	S.p0 = g.logu(1e3, 1e9); S.j2 = g.logu(1e3, 1e13);
	S.v0 = g.logu(1e-3, 1.0); S.vend = S.v0 * g.logu(0.05, 0.999); S.vk00 = S.v0 * g.logu(0.5, 2.0);
	const double c = g.unit();
	lr0 = c / S.Y0; lr1 = (1.0 - c) / S.Y1;
}

__global__ void k_rhs(uint64_t seed, int iters, unsigned long long* c) {
	Rng g{seed ^ (0x100000001B3ull * (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x))};
	unsigned long long n = 0, f = 0, b = 0, d = 0;
	for (int k = 0; k < iters; k++) {
		const Phys P = phases((int)(g.next() & 3));
		const ShockRcp rc = shock_rcp(P);
		Prof S; double lr0, lr1;
		problem(g, S, lr0, lr1);
		double v = S.v0 + (S.vend - S.v0) * g.unit(), y[3] = {S.p0 * g.logu(0.5, 4.0), S.vk00 * g.logu(0.3, 1.5), 0.0};
		switch ((int)(g.next() % 10)) {
			case 0: v = S.v0; y[0] = S.p0; break;                                  // p = P: zero lam numerators
			case 1: y[0] = -P.pi[0] + P.pi[0] * 1e-15 * g.unit(); break;           // p + pi0 -> 0
			case 2: y[0] = -P.pi[1] + g.bits(-300, 10); break;                     // p + pi1 tiny or zero
			case 3: v = S.Y0 * y[1] * (1.0 + 1e-15 * (g.unit() - 0.5)); break;     // v2 -> 0
			case 4: S.Y0 = fabs(g.bits(-200, -100)); lr0 = g.bits(-5, 200); break;
			case 5: lr1 = g.bits(-400, 400); break;
			case 6: y[0] = g.bits(-400, 400); y[1] = g.bits(-400, 400); v = g.bits(-400, 400); break;
			default: break;                                                        // at scale
		}
		const bool fast0 = rc.ok && in_band(S.Y0, BAND_LEAF) && in_band(S.Y1, BAND_LEAF) && in_band(lr0, BAND_LEAF) &&
		                   in_band(lr1, BAND_LEAF);
		bool fast = fast0;
		double d1[3], d2[3];
		const bool ok1 = prof_rhs_rcp(P, S, rc, rcp_nr(S.Y1), lr0, lr1, v, y, d1, fast);
		const bool ok2 = prof_rhs(P, S, lr0, lr1, v, y, d2);
		n++;
		if (fast) {
			f++;
			if (ok1 != ok2 || !same(d1[0], d2[0]) || !same(d1[1], d2[1]) || !same(d1[2], d2[2])) d++;
		} else b++;
	}
	count(c, C_CASES, n); count(c, C_FAST, f); count(c, C_FALLBACK, b); count(c, C_DIFFER, d);
}

__device__ bool same_run(const ProfRun& a, const ProfRun& b) {
	return same(a.y0, b.y0) && same(a.y1, b.y1) && same(a.y2, b.y2) && same(a.v, b.v) && same(a.h, b.h) &&
	       same(a.lr0, b.lr0) && same(a.lr1, b.lr1) && a.it == b.it;
}
__global__ void k_step(uint64_t seed, unsigned long long* c) {
	Rng g{seed ^ (0x100000001B3ull * (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x))};
	const Phys P = phases((int)(g.next() % 3));
	const ShockRcp rc = shock_rcp(P);
	Prof S; double lr0, lr1;
	problem(g, S, lr0, lr1);
	if ((g.next() & 7) == 0) S.Y0 = g.logu(1e-45, 1e-35);   // Y0 out of band 120: the whole integration falls back
	ProfRun R;
	prof_begin(S, g.unit(), R);
	unsigned long long n = 0, f = 0, b = 0, d = 0;
	for (int k = 0; k < 4100; k++) {   // prof_step returns -1 once 4000 steps are taken
		ProfRun R1 = R, R2 = R;
		int r1 = prof_step<true>(P, S, R1, &rc);
		n++;
		if (r1 == 2) {
			b++;
			if (!same_run(R1, R)) d++;
			r1 = prof_step(P, S, R1);
		} else f++;
		const int r2 = prof_step(P, S, R2);
		if (r1 != r2 || !same_run(R1, R2)) { d++; break; }
		R = R2;
		if (r2 != 0) break;
	}
	count(c, C_CASES, n); count(c, C_FAST, f); count(c, C_FALLBACK, b); count(c, C_DIFFER, d);
}

int report(const char* name, unsigned long long* dc) {
	unsigned long long h[C_N];
	CK(hipDeviceSynchronize());
	CK(hipMemcpy(h, dc, sizeof h, hipMemcpyDeviceToHost));
	std::printf("%s cases %llu fast %llu fallback %llu differ %llu\n", name, h[C_CASES], h[C_FAST], h[C_FALLBACK], h[C_DIFFER]);
	std::fflush(stdout);
	return 0;
}

}   // namespace

int main() {
	unsigned long long* dc = nullptr;
	CK(hipMalloc(&dc, 3 * C_N * sizeof(unsigned long long)));
	CK(hipMemset(dc, 0, 3 * C_N * sizeof(unsigned long long)));
	hipLaunchKernelGGL(k_quot, dim3(2048), dim3(256), 0, 0, 0x51ull, 128, dc);
	CK(hipGetLastError());
	if (report("quot", dc)) return 1;
	hipLaunchKernelGGL(k_rhs, dim3(1024), dim3(256), 0, 0, 0x52ull, 16, dc + C_N);
	CK(hipGetLastError());
	if (report("rhs", dc + C_N)) return 1;
	hipLaunchKernelGGL(k_step, dim3(256), dim3(256), 0, 0, 0x53ull, dc + 2 * C_N);
	CK(hipGetLastError());
	if (report("step", dc + 2 * C_N)) return 1;
	CK(hipFree(dc));
	return 0;
}
