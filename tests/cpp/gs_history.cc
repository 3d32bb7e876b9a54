// Drives a lone region through the drop-in C++ API as gs_parallel.cc:25-40
// does: setup_matrices, setup_fields, then Gauss-Seidel sweeps each followed
// by l2_error, printed one per line with full precision.  Also writes the
// final x (point-major doubles) to the file named by argv[6].
//   gs_history <simple|synthetic> <m> <n> <o> <sweeps> <x_out>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/pgmg/problems.hh"
#include "../../include/pgmg/region.hh"

template <class P>
static int run(P &p, geometry &gm, int sweeps, const char *xout) {
	comm_buffer com;
	region r(p, gm, com);
	r.setup_matrices(p);
	r.setup_fields(p);
	printf("0 %.17g\n", r.l2_error());
	for (int i = 1; i <= sweeps; i++) {
		r.gauss_seidel();
		printf("%d %.17g\n", i, r.l2_error());
	}
	r.sync_from_device();
	FILE *f = fopen(xout, "wb");
	if (!f) return 1;
	fwrite(r.x0, sizeof(double), (size_t)r.smno, f);
	fclose(f);
	printf("gs_num %d\n", region::gs_num);
	return 0;
}

int main(int argc, char **argv) {
	if (argc < 7) {fprintf(stderr, "usage: gs_history simple|synthetic m n o sweeps x_out\n"); return 2;}
	geometry gm(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), false, false, false);
	int sweeps = atoi(argv[5]);
	if (strcmp(argv[1], "simple") == 0) {
		problem_simple p(gm);
		return run(p, gm, sweeps, argv[6]);
	}
	synthetic_problem p(gm, PGMG_FAM_FACE7_UNIFORM, 12345);
	return run(p, gm, sweeps, argv[6]);
}
