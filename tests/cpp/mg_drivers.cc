// Drives a multigrid hierarchy through the drop-in C++ API as the
// reference's drivers do:
//  - cycles: v_cycle() then l2_error_all() until sqrt(l2/rhs) < tol (the
//    stop rule the fixtures use), printing each l2_error_all and the sweep
//    count region::gs_num, and writing the final x to argv[6];
//  - guide: guide.set_tol(scale) then solve_v_cycles() three times
//    (guide_test.cc:26-50), printing the return code, residual and cycle
//    prediction;
//  - fmg: one fmg_cycle() from x = 0 (multigrid.cc:109-134), then l2, and
//    the iterate written to argv[7];
//  - slices: two V-cycles, then output_x / output_r / output_residual of
//    z-slice k (argv[8]) at level 0 and k/2 at level 1, files <x_out>_x0.bin
//    etc. (as the reference's output routines name nothing, the names are
//    the caller's);
//  - pcg: conjugate gradients preconditioned by one V-cycle from x = 0
//    (the way PGMG preconditions CG, README.md:55-61), with Ax and
//    source_soln_dot, printing |r|^2 per iteration; then the other
//    outside-array helpers (copy_rhs_to_soln, soln_diff, overwrite_r_field,
//    update_r_field) on the result.
//  - diag: diagnostic_S / diagnostic_A / diagnostic_x of level argv[7]
//    (argv[8] = S, A or x; x after two V-cycles), nothing else printed;
//  - lu: the bottom LU made live by enable_exact() + lu_decomposition() on
//    the bottom region, then cycles to 1e-10, x written to argv[7];
//  - ctor_lu: lu_decomposition() before the operators exist.
//   mg_drivers <cycles|guide|fmg|slices|pcg|diag|lu|ctor_lu> <family> <seed> <m> <n> <o> <x_out|level> [smoother|k|S/A/x] [bottom]
// <family> may carry the periodic axes as ":<mask>" (1 = x, 2 = y, 4 = z);
// "simple" is problem_simple.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/pgmg/multigrid.hh"
#include "../../include/pgmg/problems.hh"

template <class P>
int drive(int argc, char **argv, geometry &gm, P &p);

// overwrite_r_field / update_r_field need the problem class's reset_r_field
// and update_r_field, which only synthetic_problem has.
void reset_source(multigrid &mg, synthetic_problem &p) {mg.overwrite_r_field(p);}
void add_source(multigrid &mg, synthetic_problem &p) {mg.update_r_field(p);}
void reset_source(multigrid &, problem_simple &) {p_fatal_error("problem_simple has no reset_r_field", 1);}
void add_source(multigrid &, problem_simple &) {p_fatal_error("problem_simple has no update_r_field", 1);}

int main(int argc, char **argv) {
	if (argc < 8) {fprintf(stderr, "usage: mg_drivers cycles|guide|fmg family seed m n o x_out [smoother] [bottom]\n"); return 2;}
	const char *colon = strchr(argv[2], ':');
	const int prd = colon ? atoi(colon + 1) : 0;
	geometry gm(atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), (prd & 1) != 0, (prd & 2) != 0, (prd & 4) != 0);
	if (strncmp(argv[2], "simple", 6) == 0) {
		problem_simple p(gm);
		return drive(argc, argv, gm, p);
	}
	synthetic_problem p(gm, atoi(argv[2]), strtoull(argv[3], NULL, 10));
	return drive(argc, argv, gm, p);
}

template <class P>
int drive(int argc, char **argv, geometry &gm, P &p) {
	const char *mode = argv[1];
	comm_buffer com;
	pgmg_opts opts;
	pgmg_default_opts(&opts);
	const bool numeric_opts = strcmp(mode, "slices") != 0 && strcmp(mode, "diag") != 0;
	opts.smoother = argc > 8 && numeric_opts ? atoi(argv[8]) : PGMG_SMOOTH_EXACT;
	opts.bottom = argc > 9 && numeric_opts ? atoi(argv[9]) : PGMG_BOTTOM_GS24;
	multigrid mg(p, gm, com, &opts);
	mg.setup_matrices(p);
	mg.setup_fields(p);
	if (strcmp(mode, "diag") == 0) {
		// region diagnostics of level argv[7] (S, A, or x after two V-cycles),
		// alone on standard output
		const int level = atoi(argv[7]);
		const char what = argc > 8 ? argv[8][0] : 'S';
		if (what == 'S') mg.diagnostic_S(level);
		else if (what == 'A') mg.diagnostic_A(level);
		else {mg.v_cycle(); mg.v_cycle(); mg.diagnostic_x(level);}
		fflush(stdout);
		mg.free();
		return 0;
	}
	if (strcmp(mode, "lu") == 0) {
		// the bottom LU made live as the reference's users do
		// (reg[nl-1]->enable_exact(); reg[nl-1]->lu_decomposition();), then
		// V-cycles to 1e-10 as "cycles"; prints whether Aexact is set
		region *b = mg.reg[mg.nlevels - 1];
		printf("aexact_before %d\n", b->Aexact != NULL);
		b->enable_exact();
		b->lu_decomposition();
		printf("aexact_after %d\n", b->Aexact != NULL);
		if (b->Aexact) printf("lu00 %.17g piv0 %d\n", b->Aexact[0], b->ipiv[0]);
		mg.compute_rats();
		printf("aexact_after_rats %d\n", b->Aexact != NULL);
		double bn = mg.rhs_l2norm(), e = mg.l2_error_all();
		int c = 0;
		const int g0 = region::gs_num;
		while (c < 100 && !(sqrt(e / bn) < 1e-10)) {mg.v_cycle(); c++; e = mg.l2_error_all();}
		printf("cycles %d res %.17g sweeps_per_cycle %d\n", c, e, c ? (region::gs_num - g0) / c : 0);
		mg.reg[0]->sync_from_device();
		FILE *f = fopen(argv[7], "wb");
		if (!f) return 1;
		fwrite(mg.reg[0]->x0, sizeof(double), (size_t)mg.reg[0]->smno, f);
		fclose(f);
		mg.free();
		return 0;
	}
	if (strcmp(mode, "ctor_lu") == 0) {
		// lu_decomposition before any operator exists (what the reference's
		// constructor does, multigrid.cc:81-84): the factor fails, Aexact
		// stays NULL and the bottom keeps 24 sweeps after setup
		multigrid mg2(p, gm, com, &opts);
		region *b = mg2.reg[mg2.nlevels - 1];
		b->enable_exact();
		b->lu_decomposition();
		printf("aexact_ctor %d\n", b->Aexact != NULL);
		mg2.setup_matrices(p);
		mg2.setup_fields(p);
		printf("aexact_setup %d\n", b->Aexact != NULL);
		const int g0 = region::gs_num;
		mg2.v_cycle();
		printf("sweeps_per_cycle %d\n", region::gs_num - g0);
		return 0;
	}
	printf("levels %d", mg.nlevels);
	for (int l = 0; l < mg.nlevels; l++) printf(" %dx%dx%d", mg.reg[l]->m, mg.reg[l]->n, mg.reg[l]->o);
	printf("\n");
	if (strcmp(mode, "cycles") == 0) {
		double bn = mg.rhs_l2norm(), e = mg.l2_error_all();
		printf("res 0 %.17g\n", e);
		int c = 0;
		while (c < 100 && !(sqrt(e / bn) < 1e-10)) {
			int g0 = region::gs_num;
			mg.v_cycle();
			c++;
			e = mg.l2_error_all();
			printf("res %d %.17g sweeps %d\n", c, e, region::gs_num - g0);
		}
		mg.reg[0]->sync_from_device();
		FILE *f = fopen(argv[7], "wb");
		if (!f) return 1;
		fwrite(mg.reg[0]->x0, sizeof(double), (size_t)mg.reg[0]->smno, f);
		fclose(f);
	} else if (strcmp(mode, "guide") == 0) {
		mg.guide.set_tol(1.0);
		for (int t = 0; t < 3; t++) {
			mg.setup_fields(p);
			double res = -1;
			int pred = mg.guide.pred_iters();
			int rc = mg.solve_v_cycles(res);
			printf("guide %d rc %d res %.6g pred %d next %d\n", t, rc, res, pred, mg.guide.pred_iters());
		}
	} else if (strcmp(mode, "slices") == 0) {
		const int k = argc > 8 ? atoi(argv[8]) : 0;
		mg.v_cycle();
		mg.v_cycle();
		for (int l = 0; l < 2 && l < mg.nlevels; l++) {
			const int kl = l == 0 ? k : k / 2;
			char nm[512];
			snprintf(nm, sizeof nm, "%s_x%d.bin", argv[7], l); mg.output_x(nm, kl, l);
			snprintf(nm, sizeof nm, "%s_r%d.bin", argv[7], l); mg.output_r(nm, kl, l);
			snprintf(nm, sizeof nm, "%s_res%d.bin", argv[7], l); mg.output_residual(nm, kl, l);
		}
		printf("slices done\n");
	} else if (strcmp(mode, "pcg") == 0) {
		region &R = *mg.reg[0];
		const int N = R.smno;
		R.sync_from_device();
		double *b = new double[N], *u = new double[N], *r = new double[N], *z = new double[N];
		double *pv = new double[N], *Ap = new double[N];
		double bb = 0;
		for (int q = 0; q < N; q++) {b[q] = R.r0[q]; u[q] = 0; r[q] = b[q]; bb += b[q] * b[q];}
		double rz_old = 0, rr = bb;
		printf("pcg 0 %.17g\n", rr);
		int it = 0;
		while (it < 40 && rr > 1e-20 * bb) {
			// z = M^-1 r: one V-cycle on A z = r from z = 0
			memcpy(R.r0, r, sizeof(double) * N);
			R.sync_r_to_device();
			R.allclear_x_field();
			mg.v_cycle();
			const double rz = mg.source_soln_dot();
			R.sync_from_device();
			memcpy(z, R.x0, sizeof(double) * N);
			const double beta = it == 0 ? 0 : rz / rz_old;
			for (int q = 0; q < N; q++) pv[q] = z[q] + beta * pv[q];
			mg.Ax(pv, Ap);
			double pAp = 0;
			for (int q = 0; q < N; q++) pAp += pv[q] * Ap[q];
			const double alpha = rz / pAp;
			rr = 0;
			for (int q = 0; q < N; q++) {u[q] += alpha * pv[q]; r[q] -= alpha * Ap[q]; rr += r[q] * r[q];}
			rz_old = rz;
			it++;
			printf("pcg %d %.17g\n", it, rr);
		}
		// check the solution against the operator: |b - A u|^2
		mg.Ax(u, Ap);
		double tr = 0;
		for (int q = 0; q < N; q++) tr += (b[q] - Ap[q]) * (b[q] - Ap[q]);
		printf("true %.17g bb %.17g\n", tr, bb);
		// x = u on the device: soln_diff against u is 0
		memcpy(R.x0, u, sizeof(double) * N);
		R.sync_x_to_device();
		printf("diff_self %.17g\n", mg.soln_diff(u));
		// source back to b, then x = r and soln_diff against b is 0
		reset_source(mg, p);
		mg.copy_rhs_to_soln();
		printf("diff_rhs %.17g\n", mg.soln_diff(b));
		// update_r_field adds 0.5 b: rhs_l2norm grows by 2.25
		add_source(mg, p);
		printf("rhs_ratio %.17g\n", mg.rhs_l2norm() / bb);
		delete[] b; delete[] u; delete[] r; delete[] z; delete[] pv; delete[] Ap;
	} else {
		mg.fmg_cycle();
		printf("fmg %.17g rhs %.17g\n", mg.l2_error_all(), mg.rhs_l2norm());
		mg.reg[0]->sync_from_device();
		FILE *f = fopen(argv[7], "wb");
		if (!f) return 1;
		fwrite(mg.reg[0]->x0, sizeof(double), (size_t)mg.reg[0]->smno, f);
		fclose(f);
	}
	mg.free();
	return 0;
}
