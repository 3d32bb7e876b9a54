// Drives the drop-in multigrid::mg_guide through the sequence of update()
// calls that tests/golden/guide_sequence.txt records for the reference, and
// prints the state in the same format.  Needs no GPU.
#include <cstdio>

#include "../../include/pgmg/multigrid.hh"

static const int guide_extras[] = {0, 0, 1, 0, 3, 0, 0, 2, 0, 0, 0, 0, 1, 6, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 0};

int main() {
	for (int avg = 0; avg < 2; avg++) {
		multigrid::mg_guide g;
		g.set_tol(3);
		if (avg) g.average_on();
		printf("tol %.17g\n", g.tol);
		for (size_t q = 0; q < sizeof guide_extras / sizeof guide_extras[0]; q++) {
			g.update(guide_extras[q]);
			printf("%d %d %d %.17g %.17g %d\n", avg, guide_extras[q], g.iter, g.iter_avg, g.iter_var, g.pred_iters());
		}
	}
	return 0;
}
