// Host check of the analytic Newton Jacobian (jacobian_analytic_row, spb6_cell.h) against difference quotients of the
// residual in the reference's operation order (cell_res, the same code the plain form runs on the device).
//
// Random physical cell states and face data over three phase pairs (air / helium, water pi = 6e8 / air, covolume gas /
// gas with q), 1D and 2D rows, wall faces, quiescent faces sitting on the upwind kink u_f = 0, and both defect segments of
// the closure (shock budget Q = 0 and Q > 0). Reference: central difference quotients with h = 1e-6 (|p| + 1) (forward
// ones, Richardson-extrapolated, on the kink rows, whose analytic value is the right derivative); allowed
// difference 10 x (truncation estimated from the quotients at h and 2h + round-off of G, about
// eps (|p| + max pi + max |q| (rho_1 + rho_2)), over h) + 1e-9 (|entry| + 1). Prints "entries N bad B kinks K
// worst W" and exits non-zero when an entry fails.
#include "spb6_cell.h"
#include <cstdio>
#include <random>

using namespace spb;

int main(int argc, char** argv) {
	const int trials = argc > 1 ? atoi(argv[1]) : 20000;
	std::mt19937_64 rng(7);
	std::uniform_real_distribution<double> U(0.0, 1.0);
	int bad = 0, tot = 0, kinks = 0;
	double worst = 0.0;
	for (int trial = 0; trial < trials; trial++) {
		Phys P{};
		const int mat = trial % 3;
		if (mat == 0) { P.g[0] = 1.4; P.g[1] = 1.648; }
		else if (mat == 1) { P.g[0] = 4.4; P.pi[0] = 6e8; P.g[1] = 1.4; }
		else { P.g[0] = 1.4; P.g[1] = 1.4; P.b[0] = 1e-3; P.q[1] = -1e5; }
		const double p0 = 1e5 * (0.5 + U(rng) * 4), P_REF = 1e5;
		const double a1 = 0.05 + 0.9 * U(rng), r1 = (mat == 1 ? 1000 : 1.2) * (0.5 + U(rng)), r2 = 0.2 + 2 * U(rng);
		const double m1 = a1 * r1, m2 = (1 - a1) * r2, u = 50 * (U(rng) - 0.5), v = 50 * (U(rng) - 0.5);
		CellIn C{};
		C.xs[IA1] = a1; C.xs[IM1] = m1; C.xs[IM2] = m2; C.xs[IMU] = (m1 + m2) * u; C.xs[IMV] = (m1 + m2) * v;
		C.xs[IE1] = eps_amp(P, 0, a1, m1, p0); C.xs[IE2] = eps_amp(P, 1, 1 - a1, m2, p0);
		C.xs[IEN] = C.xs[IE1] + C.xs[IE2] + 0.5 * (m1 + m2) * (u * u + v * v);
		C.a2s = 1 - a1;
		C.Q0 = (trial % 4 == 0) ? 0.0 : 1e3 * U(rng);
		C.Q1 = (trial % 4 == 0) ? 0.0 : 1e3 * U(rng);
		FaceIn* F[4] = {&C.fe, &C.fw, &C.fn, &C.fs};
		for (int f = 0; f < 4; f++) {
			FaceIn& x = *F[f];
			x.wl = U(rng); x.pv = 100 * (U(rng) - 0.5); x.ubs = 20 * (U(rng) - 0.5); x.kp = 1e-4 * U(rng); x.ube = 20 * (U(rng) - 0.5);
			for (int k = 0; k < 4; k++) {
				const double base = k == 0 ? m1 : k == 1 ? m2 : k == 2 ? C.xs[IE1] : C.xs[IE2];
				x.cl[k] = base * (0.8 + 0.4 * U(rng));
				x.cr[k] = base * (0.8 + 0.4 * U(rng));
			}
			x.utl = 30 * (U(rng) - 0.5); x.utr = 30 * (U(rng) - 0.5); x.wall = -1; x.bw = 0;
			if (trial % 7 == 0 && f == 0) { x.wall = 0; x.bw = 0.0; }          // resting wall face
			if (trial % 13 == 0 && f == 2) { x.wall = 2; x.bw = 3.0; }         // moving wall face
			if (trial % 11 == 0 && f == 1) { x.ubs = 0; x.ube = 0; }           // quiescent face: u_f = kp (PL - PR) = 0
		}
		const double rx = 0.05 * U(rng), ry = 0.05 * U(rng);
		const bool twoD = trial % 5 != 0;
		double Pv[5];
		for (int k = 0; k < 5; k++) Pv[k] = p0 - P_REF + (trial % 11 == 0 ? 0.0 : 1e3 * (U(rng) - 0.5));
		if (trial % 11 == 0) { Pv[1] = Pv[0]; kinks++; }
		double G;
		if (!cell_res(P, P_REF, rx, ry, twoD, C, Pv[0], Pv[1], Pv[2], Pv[3], Pv[4], G, nullptr)) continue;
		const double p = (G + Pv[0]) + P_REF, plo = smax(-P.pi[0], -P.pi[1]);
		double Jr[5];
		auto face_at = [&](int w, FaceIn& f) { f = w == 0 ? C.fe : w == 1 ? C.fw : w == 2 ? C.fn : C.fs; };
		if (!jacobian_analytic_row(P, P_REF, rx, ry, twoD, C.xs, C.Q0, C.Q1, Pv, G, face_at, Jr)) {
			bad++;
			printf("trial %d: analytic row failed\n", trial);
			continue;
		}
		// difference quotients with h = 1e-6 (|p| + 1): central ones, and forward ones on the kink rows (the analytic value
		// there is the right derivative); truncation from the quotients at h and 2h (Richardson), round-off of G about
		// eps (|p| + max pi + max |q| (rho_1 + rho_2)) divided by h
		const double h = 1e-6 * (fabs(p) + 1.0), noise = 4.0 * 2.2e-16 * (fabs(p) + smax(P.pi[0], P.pi[1]) + smax(fabs(P.q[0]), fabs(P.q[1])) * (r1 + r2)) / h;
		const bool kink_row = trial % 11 == 0;
		auto G_at = [&](int j, double d) {
			double Q[5];
			for (int k = 0; k < 5; k++) Q[k] = Pv[k];
			Q[j] += d;
			double G2 = NAN;
			return cell_res(P, P_REF, rx, ry, twoD, C, Q[0], Q[1], Q[2], Q[3], Q[4], G2, nullptr) ? G2 : NAN;
		};
		for (int j = 0; j < (twoD ? 5 : 3); j++) {
			double d1, d2;
			if (kink_row) { d1 = (G_at(j, h) - G) / h; d2 = (G_at(j, 2 * h) - G) / (2 * h); }
			else { d1 = (G_at(j, h) - G_at(j, -h)) / (2 * h); d2 = (G_at(j, 2 * h) - G_at(j, -2 * h)) / (4 * h); }
			const double ref = kink_row ? 2 * d1 - d2 : d1;     // forward: first-order error removed
			const double trunc = kink_row ? fabs(d2 - d1) : fabs(d2 - d1) / 3.0;
			const double tol = 10.0 * (trunc + noise) + 1e-9 * (fabs(ref) + 1.0);
			const double err = fabs(Jr[j] - ref);
			tot++;
			if (!(err <= tol)) {
				bad++;
				if (bad <= 10) printf("trial %d mat %d entry %d: analytic %.10e difference quotient %.10e (allowed %.2e)\n", trial, mat, j, Jr[j], ref, tol);
			}
			if (err / tol > worst) worst = err / tol;
		}
	}
	printf("entries %d bad %d kinks %d worst %.3f\n", tot, bad, kinks, worst);
	return bad != 0;
}
