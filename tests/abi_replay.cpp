// The reference's time loop (src/SPB6_PC_IMEX.cpp, PC:1230-1296) written against the C ABI of include/spb6_imex.h, as a
// maintainer would bind it: caldt and the start-up ramp, the attempt loop with reject-and-halve, and
// each attempt replayed stage by stage through spb6_attempt_begin / spb6_stage_combine / spb6_stage_update /
// spb6_stage_pressure / spb6_attempt_accept with the caller's own Butcher tables (set_tab, PC:1195-1218). A second
// context on the same initial state runs spb6_run; the two must agree bit for bit (state, pressure, Newton and Krylov
// counts).
//
// usage: abi_replay <case> <rk> <steps>   case 0: 200-cell water-air tube (1 GPa | 1e5 Pa, transmissive),
//                                          case 1: 48 x 24 shock-bubble (walls at the bottom and top)
// prints "identical steps S newton N krylov K" and exits 0, or the first difference and exits 1.
#include "spb6_imex.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// conserved state of a stiffened-gas cell (prim2cell, PHY:849-854, with q = b = 0)
void prim2cell(const spb6_config_t& c, double a1, double r1, double r2, double u, double v, double p, double* Q) {
	const double a2 = 1.0 - a1, m1 = a1 * r1, m2 = a2 * r2, rho = m1 + m2;
	const double g0 = c.ph[0].gamma, pi0 = c.ph[0].pi_inf, g1 = c.ph[1].gamma, pi1 = c.ph[1].pi_inf;
	Q[0] = a1; Q[1] = m1; Q[2] = m2; Q[3] = rho * u; Q[4] = rho * v;
	Q[5] = (p + g0 * pi0) * a1 / (g0 - 1.0);
	Q[6] = (p + g1 * pi1) * a2 / (g1 - 1.0);
	Q[7] = Q[5] + Q[6] + 0.5 * rho * (u * u + v * v);
}

void defaults(spb6_config_t& c) {   // the reference defaults (PC:57-80)
	std::memset(&c, 0, sizeof c);
	c.x0 = 0.0; c.x1 = 1.0; c.y0 = 0.0; c.y1 = 1.0;
	for (int k = 0; k < 2; k++) c.ph[k].cv = 1000.0;
	c.rk = 2; c.cfl = 0.5; c.cfl_ac = 0.4; c.ramp = 1; c.dtfix = 0.0; c.acdt = 0;
	c.newton_tol = 1e-13; c.newton_iter = 60; c.krylov_tol = 1e-13; c.krylov_iter = 4000;
	c.limiter = 1; c.order = 2; c.rus_local = 1;
}

void make_case(int which, spb6_config_t& c, std::vector<double>& U) {
	defaults(c);
	if (which == 0) {
		c.nx = 200; c.ny = 1;
		c.ph[0].gamma = 4.4; c.ph[0].pi_inf = 6e8; c.ph[1].gamma = 1.4;
		U.assign((size_t)c.nx * 8, 0.0);
		for (int i = 0; i < c.nx; i++) {
			const double x = (i + 0.5) / c.nx;
			const bool L = x < 0.7;
			prim2cell(c, L ? 1 - 1e-8 : 1e-8, 1000.0, 50.0, 0.0, 0.0, L ? 1e9 : 1e5, &U[(size_t)i * 8]);
		}
	} else {
		c.nx = 48; c.ny = 24; c.x1 = 0.2; c.y1 = 0.1;
		c.bc[2] = c.bc[3] = SPB6_BC_WALL;
		c.ph[0].gamma = 1.4; c.ph[1].gamma = 1.648;
		const double g = 1.4, Ms = 1.22, p0 = 1e5, r0 = 1.2, c0 = std::sqrt(g * p0 / r0);
		const double p1 = p0 * (1 + 2 * g / (g + 1) * (Ms * Ms - 1)), r1 = r0 * (g + 1) * Ms * Ms / ((g - 1) * Ms * Ms + 2);
		const double u1 = Ms * c0 * (1 - r0 / r1);
		U.assign((size_t)c.nx * c.ny * 8, 0.0);
		for (int j = 0; j < c.ny; j++)
			for (int i = 0; i < c.nx; i++) {
				const double x = (i + 0.5) * 0.2 / c.nx, y = (j + 0.5) * 0.1 / c.ny;
				const bool in = (x - 0.14) * (x - 0.14) + (y - 0.05) * (y - 0.05) < 0.025 * 0.025, post = x < 0.1;
				prim2cell(c, in ? 1e-6 : 1 - 1e-6, post ? r1 : r0, 0.166, post ? u1 : 0.0, 0.0, post ? p1 : p0,
				          &U[((size_t)i + (size_t)j * c.nx) * 8]);
			}
	}
}

// set_tab (PC:1195-1218)
void set_tab(int rk, double Ae[3][3], double Ai[3][3]) {
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++) Ae[a][b] = Ai[a][b] = 0.0;
	if (rk == 1) Ai[0][0] = 1.0;
	else if (rk == 2) {
		const double bb = 1.0 - 1.0 / std::sqrt(2.0), cc = 1.0 / (2.0 * bb);
		Ae[1][0] = cc; Ai[0][0] = bb; Ai[1][0] = 1 - bb; Ai[1][1] = bb;
	} else {
		const double r12 = 0.5, r13 = 1.0 / 3.0, r14 = 1.0 / 4.0;
		Ae[1][0] = r12; Ae[2][0] = r12; Ae[2][1] = r12;
		Ai[0][0] = r14; Ai[1][1] = r14; Ai[2][0] = r13; Ai[2][1] = r13; Ai[2][2] = r13;
	}
}

// the time loop (PC:1230-1296) with each attempt replayed stage by stage; returns 0 or the failure status
int replay(spb6_ctx* g, const spb6_config_t& cfg, double t_end, long max_steps) {
	const double ramp_factor = 1.2;
	double t = 0.0, dtprev = -1.0, ir = 0.0;
	bool startup = true;
	for (long done = 0; t < t_end * (1 - 1e-12) && done < max_steps; done++) {
		bool ramping = false;
		double dt;
		if (cfg.dtfix > 0) dt = cfg.dtfix;
		else {
			if (spb6_caldt(g, cfg.acdt, &ir)) return SPB6_HIP_ERROR;
			dt = cfg.cfl * ir;
			if (cfg.ramp) {
				if (dtprev < 0) { if (spb6_caldt(g, 1, &ir)) return SPB6_HIP_ERROR; dtprev = cfg.cfl_ac * ir / ramp_factor; }
				if (dt > ramp_factor * dtprev) { dt = ramp_factor * dtprev; ramping = true; }
			}
			if (!std::isfinite(dt)) { if (spb6_caldt(g, 1, &ir)) return SPB6_HIP_ERROR; dt = cfg.cfl_ac * ir; }
		}
		dt = (t_end - t) < dt ? (t_end - t) : dt;
		if (!(dt > 1e-12 * t_end)) return SPB6_DT_UNDERFLOW;
		bool ok = false;
		for (int att = 0; att < 30 && !ok; att++) {
			const int rk = (cfg.ramp && ramping && startup) ? 1 : cfg.rk;
			double Ae[3][3], Ai[3][3];
			set_tab(rk, Ae, Ai);
			if (spb6_attempt_begin(g)) return SPB6_HIP_ERROR;
			ok = true;
			for (int rks = 0; rks < rk && ok; rks++) {
				if (spb6_stage_combine(g, rks, dt, Ae[rks], Ai[rks])) return SPB6_HIP_ERROR;
				const double adt = Ai[rks][rks] * dt;
				int it = 0;
				double res = 0.0;
				ok = spb6_stage_update(g, adt) == SPB6_OK;
				ok = ok && spb6_stage_pressure(g, adt, rks, &it, &res) == SPB6_OK;
			}
			if (ok) spb6_attempt_accept(g, dt);
			else dt *= 0.5;
		}
		if (!ok) return SPB6_STEP_FAILED;
		t = t + dt;
		dtprev = dt;
		if (!ramping) startup = false;
	}
	return SPB6_OK;
}

}   // namespace

int main(int argc, char** argv) {
	if (argc < 4) { std::fprintf(stderr, "usage: abi_replay <case> <rk> <steps>\n"); return 2; }
	const int which = std::atoi(argv[1]), rk = std::atoi(argv[2]);
	const long steps = std::atol(argv[3]);
	spb6_config_t cfg;
	std::vector<double> U;
	make_case(which, cfg, U);
	cfg.rk = rk;
	spb6_ctx* a = spb6_create(&cfg, 0);
	spb6_ctx* b = spb6_create(&cfg, 0);
	if (!a || !b) { std::fprintf(stderr, "spb6_create: %s\n", spb6_last_error()); return 1; }
	if (spb6_set_state(a, U.data()) || spb6_set_state(b, U.data())) { std::fprintf(stderr, "set_state failed\n"); return 1; }
	const int ra = spb6_run(a, 1.0, steps);
	const int rb = replay(b, cfg, 1.0, steps);
	if (ra != SPB6_OK || rb != SPB6_OK) { std::fprintf(stderr, "run %d replay %d\n", ra, rb); return 1; }
	const size_t n = (size_t)cfg.nx * cfg.ny;
	std::vector<double> Ua(n * 8), Ub(n * 8), Pa(n), Pb(n);
	double pra, prb;
	spb6_get_state(a, Ua.data()); spb6_get_state(b, Ub.data());
	spb6_get_pressure(a, Pa.data(), &pra); spb6_get_pressure(b, Pb.data(), &prb);
	spb6_stats_t sa, sb;
	spb6_get_stats(a, &sa); spb6_get_stats(b, &sb);
	int rc = 0;
	for (size_t q = 0; q < n * 8; q++)
		if (std::memcmp(&Ua[q], &Ub[q], sizeof(double))) { std::printf("state differs at cell %zu field %zu\n", q / 8, q % 8); rc = 1; break; }
	for (size_t q = 0; q < n && !rc; q++)
		if (std::memcmp(&Pa[q], &Pb[q], sizeof(double))) { std::printf("pressure differs at cell %zu\n", q); rc = 1; }
	if (!rc && (sa.steps != sb.steps || sa.newton_it != sb.newton_it || sa.krylov_it != sb.krylov_it || sa.stages != sb.stages)) {
		std::printf("counters differ: steps %ld/%ld newton %ld/%ld krylov %ld/%ld stages %ld/%ld\n", sa.steps, sb.steps, sa.newton_it,
		            sb.newton_it, sa.krylov_it, sb.krylov_it, sa.stages, sb.stages);
		rc = 1;
	}
	if (!rc) std::printf("identical steps %ld newton %ld krylov %ld\n", sb.steps, sb.newton_it, sb.krylov_it);
	spb6_destroy(a);
	spb6_destroy(b);
	return rc;
}
