// Self-check of the strong-shock kernel's select form of the shooting state machine (spb6_device.h: shoot_next_sel
// against shoot_next). Built with the library into tests/bin/shoot_check; `shoot_check host` runs on the CPU
// (tests/test_shoot_select.py), `shoot_check device` runs the same cases in a kernel on the GPU
// (tests/test_gpu_shoot_select.py).
//
//   single  one transition from random states: every phase (the finished ones included), brackets at and near the
//           -40 / 40 limits and narrower than 1e-10, F of either sign, zeros of either sign, values at the 1e-10
//           convergence bound, NaN and infinities, the Illinois side and its iteration count at the 100 cap;
//   chain   whole shootings on F(t) = s (t* - t) (1 + 0.3 tanh(t)) and a few t -> NaN holes, the state of both forms
//           compared after every transition.
//
// The returned t and every field of the state (doubles as bit patterns) must agree. Prints one line per part,
// "<part> cases N differ D"; exit status 0 when it ran.
#include "../diffusionimex_amd/csrc/spb6_device.h"

#include <cstdio>
#include <cstring>

using namespace spb;

namespace {

struct Rng {
	unsigned long long s;
	SPB_HD unsigned long long next() {
		s += 0x9E3779B97F4A7C15ull;
		unsigned long long z = s;
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	SPB_HD double uni() { return (double)(next() >> 11) * 0x1.0p-53; }
	SPB_HD int pick(int n) { return (int)(next() % (unsigned long long)n); }
};

SPB_HD double special(Rng& r, double scale) {
	switch (r.pick(12)) {
	case 0: return 0.0;
	case 1: return -0.0;
	case 2: return dnan();
	case 3: return dinf();
	case 4: return -dinf();
	case 5: return 1e-10;
	case 6: return -1e-10;
	case 7: return (r.uni() - 0.5) * 2e-10;
	default: return (r.uni() - 0.5) * scale;
	}
}

SPB_HD void random_state(Rng& r, Shoot& H, double& f, double& q) {
	H.phase = r.pick(8) == 0 ? SH_DONE + r.pick(2) : r.pick(SH_DONE);
	const int b = r.pick(6);
	const double ta = b == 0 ? -40.0 : b == 1 ? -40.0 + r.uni() : -40.0 + 80.0 * r.uni();
	double tb = b == 2 ? 40.0 : b == 3 ? 40.0 - r.uni() : ta + 4.0 * r.uni();
	if (r.pick(8) == 0) tb = ta + 1e-10 * r.uni() * 2.0;
	if (tb > 40.0) tb = 40.0;
	H.ta = ta; H.tb = tb;
	H.t = r.pick(4) == 0 ? special(r, 80.0) : ta + (tb - ta) * r.uni();
	H.fa = r.pick(3) == 0 ? special(r, 2.0) : -r.uni();
	H.fb = r.pick(3) == 0 ? special(r, 2.0) : r.uni();
	H.qa = special(r, 1e3); H.qb = special(r, 1e3); H.qt = special(r, 1e3);
	H.st = (double)(2 << r.pick(6));
	H.side = r.pick(3) - 1;
	H.it = r.pick(2) ? 95 + r.pick(6) : r.pick(100);
	f = r.pick(3) == 0 ? special(r, 2.0) : (r.uni() - 0.5) * 2.0;
	q = special(r, 1e3);
}

SPB_HD unsigned long long bits(double x) { return __builtin_bit_cast(unsigned long long, x); }

SPB_HD bool same(const Shoot& a, const Shoot& b, double ra, double rb) {
	return bits(ra) == bits(rb) && bits(a.ta) == bits(b.ta) && bits(a.tb) == bits(b.tb) && bits(a.fa) == bits(b.fa) &&
	       bits(a.fb) == bits(b.fb) && bits(a.qa) == bits(b.qa) && bits(a.qb) == bits(b.qb) && bits(a.st) == bits(b.st) &&
	       bits(a.t) == bits(b.t) && bits(a.qt) == bits(b.qt) && a.phase == b.phase && a.side == b.side && a.it == b.it;
}

// one case of part `single`: 1 when the two forms differ
SPB_HD int single_case(unsigned long long seed) {
	Rng r{seed};
	Shoot H; double f, q;
	random_state(r, H, f, q);
	Shoot A = H, B = H;
	const double ra = shoot_next(A, f, q), rb = shoot_next_sel(B, f, q);
	return same(A, B, ra, rb) ? 0 : 1;
}

// one case of part `chain`: a whole shooting on a synthetic F; the number of transitions where the forms differ
SPB_HD int chain_case(unsigned long long seed, int& steps) {
	Rng r{seed};
	const double ts = -45.0 + 90.0 * r.uni(), s = r.pick(2) ? 0.05 + r.uni() : 1e-3 * r.uni();
	const double hole = r.pick(4) == 0 ? -40.0 + 80.0 * r.uni() : 1e9;
	Shoot A, B;
	double t0 = -45.0 + 90.0 * r.uni();
	double ta = shoot_begin(A, t0), tb = shoot_begin(B, t0);
	int bad = bits(ta) != bits(tb);
	for (int k = 0; k < 200 && A.phase < SH_DONE; k++) {
		const double t = ta;
		double f = s * (ts - t) * (1.0 + 0.3 * tanh(t));
		if (fabs(t - hole) < 0.5) f = dnan();
		const double q = 1e-3 * t * t;
		ta = shoot_next(A, f, q);
		tb = shoot_next_sel(B, f, q);
		bad += same(A, B, ta, tb) ? 0 : 1;
		steps++;
	}
	return bad;
}

__global__ void k_single(unsigned long long base, int n, int* differ) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && single_case(base + (unsigned long long)i * 7919ull)) atomicAdd(differ, 1);
}
__global__ void k_chain(unsigned long long base, int n, int* differ, int* steps) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int st = 0;
	const int d = chain_case(base + (unsigned long long)i * 104729ull, st);
	if (d) atomicAdd(differ, d);
	atomicAdd(steps, st);
}

#define CK(x)                                                                                        \
	do {                                                                                             \
		hipError_t e_ = (x);                                                                         \
		if (e_ != hipSuccess) { std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); return 1; } \
	} while (0)

int run_device(int n_single, int n_chain) {
	int* d = nullptr;
	CK(hipMalloc(&d, 3 * sizeof(int)));
	CK(hipMemset(d, 0, 3 * sizeof(int)));
	hipLaunchKernelGGL(k_single, dim3((n_single + 255) / 256), dim3(256), 0, 0, 1ull, n_single, d);
	CK(hipGetLastError());
	hipLaunchKernelGGL(k_chain, dim3((n_chain + 255) / 256), dim3(256), 0, 0, 2ull, n_chain, d + 1, d + 2);
	CK(hipGetLastError());
	int h[3];
	CK(hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost));
	CK(hipFree(d));
	std::printf("single cases %d differ %d\n", n_single, h[0]);
	std::printf("chain cases %d steps %d differ %d\n", n_chain, h[2], h[1]);
	return 0;
}

int run_host(int n_single, int n_chain) {
	int bad = 0, steps = 0, badc = 0;
	for (int i = 0; i < n_single; i++) bad += single_case(1ull + (unsigned long long)i * 7919ull);
	for (int i = 0; i < n_chain; i++) badc += chain_case(2ull + (unsigned long long)i * 104729ull, steps);
	std::printf("single cases %d differ %d\n", n_single, bad);
	std::printf("chain cases %d steps %d differ %d\n", n_chain, steps, badc);
	return 0;
}

}  // namespace

int main(int argc, char** argv) {
	const bool dev = argc > 1 && std::strcmp(argv[1], "device") == 0;
	return dev ? run_device(1 << 22, 1 << 18) : run_host(1 << 20, 1 << 16);
}
