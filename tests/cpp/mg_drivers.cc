// Drives a multigrid hierarchy through the drop-in C++ API as the
// reference's drivers do:
//  - cycles: v_cycle() then l2_error_all() until sqrt(l2/rhs) < tol (the
//    stop rule the fixtures use), printing each l2_error_all and the sweep
//    count region::gs_num, and writing the final x to argv[6];
//  - guide: guide.set_tol(scale) then solve_v_cycles() three times
//    (guide_test.cc:26-50), printing the return code, residual and cycle
//    prediction;
//  - fmg: one fmg_cycle() from x = 0 (multigrid.cc:109-134), then l2.
//   mg_drivers <cycles|guide|fmg> <family> <seed> <m> <n> <o> <x_out> [smoother] [bottom]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/pgmg/multigrid.hh"
#include "../../include/pgmg/problems.hh"

int main(int argc, char **argv) {
	if (argc < 8) {fprintf(stderr, "usage: mg_drivers cycles|guide|fmg family seed m n o x_out [smoother] [bottom]\n"); return 2;}
	const char *mode = argv[1];
	geometry gm(atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), false, false, false);
	synthetic_problem p(gm, atoi(argv[2]), strtoull(argv[3], NULL, 10));
	comm_buffer com;
	pgmg_opts opts;
	pgmg_default_opts(&opts);
	opts.smoother = argc > 8 ? atoi(argv[8]) : PGMG_SMOOTH_EXACT;
	opts.bottom = argc > 9 ? atoi(argv[9]) : PGMG_BOTTOM_GS24;
	multigrid mg(p, gm, com, &opts);
	mg.setup_matrices(p);
	mg.setup_fields(p);
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
	} else {
		mg.fmg_cycle();
		printf("fmg %.17g rhs %.17g\n", mg.l2_error_all(), mg.rhs_l2norm());
	}
	mg.free();
	return 0;
}
