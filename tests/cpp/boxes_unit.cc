// Host-only check of the drop-in API's box bookkeeping: the top level's
// boxes as a problem class gives them (ingest_operator) and every coarse
// level's boxes by the Galerkin bound rule (pgmg_detail::rat_bounds), printed
// in region::diagnostic_S's format, level after level.  No GPU call is made.
//   boxes_unit <family|simple>[:prd] <seed> <m> <n> <o>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pgmg/problems.hh"
#include "../../include/pgmg/region.hh"

template <class P>
void run(P &p, int m, int n, int o, int prd) {
	pgmg_dims d = {m, n, o, prd & 1, (prd >> 1) & 1, (prd >> 2) & 1};
	int nl = 0, dims[48];
	pgmg_check(pgmg_hierarchy_shape(&d, &nl, dims), "shape");
	std::vector<double> A27;
	std::vector<int> boxes, coarse;
	pgmg_detail::ingest_operator(p, m, n, o, prd, A27, &boxes);
	for (int l = 0; l < nl; l++) {
		if (l) {
			pgmg_detail::rat_bounds(&boxes[0], dims[3 * l - 3], dims[3 * l - 2], dims[3 * l - 1], prd, dims[3 * l],
						dims[3 * l + 1], dims[3 * l + 2], coarse);
			boxes.swap(coarse);
		}
		printf("level %d\n", l);
		const int sm = dims[3 * l], sn = dims[3 * l + 1], so = dims[3 * l + 2];
		for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++) {
			const int *Sp = &boxes[6 * (i + sm * (j + sn * k))];
			printf("0: (%d,%d,%d) [%d:%d] [%d:%d] [%d:%d]\n", i, j, k, Sp[0], Sp[1], Sp[2], Sp[3], Sp[4], Sp[5]);
		}
	}
}

int main(int argc, char **argv) {
	if (argc < 6) {fprintf(stderr, "usage: boxes_unit family[:prd] seed m n o\n"); return 2;}
	const char *colon = strchr(argv[1], ':');
	const int prd = colon ? atoi(colon + 1) : 0;
	const int m = atoi(argv[3]), n = atoi(argv[4]), o = atoi(argv[5]);
	geometry gm(m, n, o, (prd & 1) != 0, (prd & 2) != 0, (prd & 4) != 0);
	if (strncmp(argv[1], "simple", 6) == 0) {
		problem_simple p(gm);
		run(p, m, n, o, prd);
	} else {
		synthetic_problem p(gm, atoi(argv[1]), strtoull(argv[2], NULL, 10));
		run(p, m, n, o, prd);
	}
	return 0;
}
