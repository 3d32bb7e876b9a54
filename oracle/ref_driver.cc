/* ref_driver.cc - drives the real PGMG reference (compiled from its own
 * sources by oracle/Makefile.ref into oracle/_ref/) to write golden vectors
 * and to time the reference CPU path.  Test infrastructure only: nothing in
 * the product links or runs it.
 *
 * The problem class below is a p_class in the sense of problem_simple.hh:14-73:
 * every grid point is an unknown ("internal"), its stencil box is the
 * [-1,1]^3 box, clipped to the grid on non-periodic axes and whole on
 * periodic ones (range_* below), and fill_entries writes the box entries in
 * k,j,i order from a 27-entry row.  Rows come from include/pgmg_problems.h.
 *
 * <family> is a family number, optionally followed by ":<prd>", the periodic
 * axes as a bit mask (1 = x, 2 = y, 4 = z), e.g. "1:5".
 *
 * Usage:
 *   pgmg_ref fixture <family> <seed> <m> <n> <o> <out.bin>
 *   pgmg_ref solve   <family> <seed> <m> <n> <o> <bottom:gs24|lu> <tol> <maxcycles> <out.bin|->
 *   pgmg_ref time    <family> <seed> <m> <n> <o> <bottom> <tol> <maxcycles>
 *   pgmg_ref fmg     <family> <seed> <m> <n> <o> <out.bin>
 *   pgmg_ref diag    <family|simple> <seed> <m> <n> <o> <level> <S|A|x>
 *     prints the reference's diagnostic_S / diagnostic_A / diagnostic_x of
 *     one level (x after two V-cycles from x = 0); "simple" is the
 *     reference's own problem_simple class (problem_simple.hh/.cc), whose
 *     boundary rows have 1x1x1 boxes.
 *
 * Output is a flat little-endian record stream, each record
 *   char name[32] | int32 dtype (0=f64, 1=i32) | int64 count | data
 * read by tests/golden/make_golden.py. */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>
#include <string>
#include "mpi.h"
#include "multigrid.hh"
#include "problem_simple.hh"
#include "../include/pgmg_problems.h"

class array_problem {
	public:
		const bool x_prd, y_prd, z_prd;
		const int m, n, o;
		int family, prd;
		uint64_t seed;
		static double dummy[1];
		array_problem(geometry &gm, int fam, uint64_t sd)
			: x_prd(gm.x_prd), y_prd(gm.y_prd), z_prd(gm.z_prd), m(gm.m), n(gm.n), o(gm.o),
			  family(fam), prd((gm.x_prd ? 1 : 0) | (gm.y_prd ? 2 : 0) | (gm.z_prd ? 4 : 0)), seed(sd) {}
		inline int range_xd(int i, int j, int k) {return x_prd || i > 0 ? -1 : 0;}
		inline int range_xu(int i, int j, int k) {return x_prd || i < m - 1 ? 2 : 1;}
		inline int range_yd(int i, int j, int k) {return y_prd || j > 0 ? -1 : 0;}
		inline int range_yu(int i, int j, int k) {return y_prd || j < n - 1 ? 2 : 1;}
		inline int range_zd(int i, int j, int k) {return z_prd || k > 0 ? -1 : 0;}
		inline int range_zu(int i, int j, int k) {return z_prd || k < o - 1 ? 2 : 1;}
		inline bool internal(int i, int j, int k) {return true;}
		inline int mem_size(int i, int j, int k) {
			return (range_xu(i, j, k) - range_xd(i, j, k)) * (range_yu(i, j, k) - range_yd(i, j, k))
			     * (range_zu(i, j, k) - range_zd(i, j, k));
		}
		void fill_entries(int i, int j, int k, double *&en) {
			double row[27];
			pgmg_row27_prd(family, seed, m, n, o, prd, i, j, k, row);
			for (int dk = range_zd(i, j, k); dk < range_zu(i, j, k); dk++)
				for (int dj = range_yd(i, j, k); dj < range_yu(i, j, k); dj++)
					for (int di = range_xd(i, j, k); di < range_xu(i, j, k); di++)
						*(en++) = row[(dk + 1) * 9 + (dj + 1) * 3 + (di + 1)];
		}
		inline double *external_ptr(int i, int j, int k) {return dummy;}
		inline double x_field(int i, int j, int k) {return 0;}
		inline double r_field(int i, int j, int k) {return pgmg_rhs(seed, m, n, o, i, j, k);}
};
double array_problem::dummy[1] = {0};

static FILE *outf = NULL;

static void put(const char *name, int dtype, long long count, const void *data) {
	if (outf == NULL) return;
	char nm[32];
	memset(nm, 0, sizeof nm);
	strncpy(nm, name, 31);
	fwrite(nm, 1, 32, outf);
	fwrite(&dtype, 4, 1, outf);
	fwrite(&count, 8, 1, outf);
	fwrite(data, dtype == 0 ? 8 : 4, count, outf);
}
static void put_d(const char *name, const std::vector<double> &v) {put(name, 0, (long long)v.size(), v.data());}
static void put_i(const char *name, const std::vector<int> &v) {put(name, 1, (long long)v.size(), v.data());}

/* Level operator expanded to 27-entry rows (zero outside the stored box). */
static std::vector<double> level_A27(region *r) {
	std::vector<double> out((size_t)r->smno * 27, 0.0);
	for (int k = 0; k < r->so; k++) for (int j = 0; j < r->sn; j++) for (int i = 0; i < r->sm; i++) {
		int l = i + r->sm * (j + r->sn * k), *Sp = r->S + 6 * l;
		double *Ap = r->A[l];
		for (int dk = Sp[4]; dk < Sp[5]; dk++) for (int dj = Sp[2]; dj < Sp[3]; dj++)
			for (int di = Sp[0]; di < Sp[1]; di++) {
				if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) {
					fprintf(stderr, "box outside [-1,1]^3 at level point %d\n", l);
					exit(3);
				}
				out[(size_t)l * 27 + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)] = *(Ap++);
			}
	}
	return out;
}
static std::vector<int> level_S(region *r) {return std::vector<int>(r->S, r->S + 6 * (size_t)r->smno);}
static std::vector<double> level_x(region *r) {
	std::vector<double> out((size_t)r->smno);
	for (int k = 0; k < r->so; k++) for (int j = 0; j < r->sn; j++) for (int i = 0; i < r->sm; i++)
		out[i + r->sm * (j + (size_t)r->sn * k)] = r->x0[i + r->mg * (j + (size_t)r->ng * k)];
	return out;
}
static std::vector<double> level_r(region *r) {return std::vector<double>(r->r0, r->r0 + r->smno);}
static void set_x(region *r, const std::vector<double> &v) {
	for (int k = 0; k < r->so; k++) for (int j = 0; j < r->sn; j++) for (int i = 0; i < r->sm; i++)
		r->x0[i + r->mg * (j + (size_t)r->ng * k)] = v[i + r->sm * (j + (size_t)r->sn * k)];
}
/* Deterministic test field for a level (same formula in the tests). */
static std::vector<double> test_field(int m, int n, int o, uint64_t seed) {
	std::vector<double> v((size_t)m * n * o);
	for (size_t p = 0; p < v.size(); p++) v[p] = 2.0 * pgmg_u01(seed, 99, p) - 1.0;
	return v;
}

static void enable_lu(multigrid &mg) {
	region *b = mg.reg[mg.nlevels - 1];
	if (b->Aexact == NULL) {
		b->enable_exact();
		if (b->Aexact != NULL) b->lu_decomposition();
	}
}

/* The reference's mg_guide (multigrid.hh:34-118, multigrid.cc:268-292)
 * driven through a fixed sequence of update(extras) calls, with and without
 * the moving average; prints the state after each call. */
static const int guide_extras[] = {0, 0, 1, 0, 3, 0, 0, 2, 0, 0, 0, 0, 1, 6, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 0};

static void run_guide() {
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
	fflush(stdout);
}

int main(int argc, char **argv) {
	MPI_Init(&argc, &argv);
	if (argc >= 2 && std::string(argv[1]) == "guide") {
		run_guide();
		MPI_Finalize();
		return 0;
	}
	if (argc < 7) {
		fprintf(stderr, "usage: pgmg_ref fixture|solve|time family seed m n o ...\n");
		MPI_Finalize();
		return 2;
	}
	std::string cmd = argv[1];
	if (cmd == "diag") {
		if (argc < 9) {fprintf(stderr, "usage: pgmg_ref diag family seed m n o level S|A|x\n"); MPI_Finalize(); return 2;}
		const char *colon = strchr(argv[2], ':');
		const int dprd = colon ? atoi(colon + 1) : 0;
		const bool simple = strncmp(argv[2], "simple", 6) == 0;
		int m = atoi(argv[4]), n = atoi(argv[5]), o = atoi(argv[6]), level = atoi(argv[7]);
		comm_buffer com;
		geometry gm(m, n, o, (dprd & 1) != 0, (dprd & 2) != 0, (dprd & 4) != 0);
		problem_simple ps(gm);
		array_problem pa(gm, simple ? 1 : atoi(argv[2]), strtoull(argv[3], NULL, 10));
		multigrid *mg = simple ? new multigrid(ps, gm, com) : new multigrid(pa, gm, com);
		if (simple) {mg->setup_matrices(ps); mg->setup_fields(ps);}
		else {mg->setup_matrices(pa); mg->setup_fields(pa);}
		char what = argv[8][0];
		if (what == 'S') mg->diagnostic_S(level);
		else if (what == 'A') mg->diagnostic_A(level);
		else {mg->v_cycle(); mg->v_cycle(); mg->diagnostic_x(level);}
		fflush(stdout);
		mg->free();
		delete mg;
		gm.free();
		MPI_Finalize();
		return 0;
	}
	int family = atoi(argv[2]), prd = 0;
	const char *colon = strchr(argv[2], ':');
	if (colon) prd = atoi(colon + 1);
	uint64_t seed = strtoull(argv[3], NULL, 10);
	int m = atoi(argv[4]), n = atoi(argv[5]), o = atoi(argv[6]);

	comm_buffer com;
	geometry gm(m, n, o, (prd & 1) != 0, (prd & 2) != 0, (prd & 4) != 0);
	array_problem pr(gm, family, seed);
	multigrid mg(pr, gm, com);
	int aexact_after_ctor = mg.reg[mg.nlevels - 1]->Aexact != NULL;
	double t0 = MPI_Wtime();
	mg.setup_matrices(pr);
	double t_setup = MPI_Wtime() - t0;
	mg.setup_fields(pr);

	if (cmd == "fixture") {
		outf = fopen(argv[7], "wb");
		if (!outf) {perror("open"); return 1;}
		std::vector<int> dims;
		for (int l = 0; l < mg.nlevels; l++) {
			dims.push_back(mg.reg[l]->m); dims.push_back(mg.reg[l]->n); dims.push_back(mg.reg[l]->o);
		}
		put_i("dims", dims);
		put_i("aexact_after_ctor", std::vector<int>(1, aexact_after_ctor));
		for (int l = 0; l < mg.nlevels; l++) {
			char nm[32];
			snprintf(nm, sizeof nm, "A%d", l); put_d(nm, level_A27(mg.reg[l]));
			snprintf(nm, sizeof nm, "S%d", l); put_i(nm, level_S(mg.reg[l]));
		}
		region *r0 = mg.reg[0];
		put_d("b", level_r(r0));
		put_d("rhs_l2norm", std::vector<double>(1, mg.rhs_l2norm()));
		/* One forward sweep, then one symmetric sweep, from x = test field. */
		std::vector<double> xt = test_field(m, n, o, seed + 1);
		set_x(r0, xt);
		put_d("l2_error_all_xt", std::vector<double>(1, mg.l2_error_all()));
		mg.gauss_seidel(0, 0);
		put_d("x_gs0", level_x(r0));
		mg.gauss_seidel(1, 0);
		put_d("x_gs1", level_x(r0));
		if (mg.nlevels > 1) {
			/* Restriction of the residual of x_gs1, then interpolation of a
			 * test coarse field added onto x_gs1. */
			mg.restriction(1);
			put_d("r1_restrict", level_r(mg.reg[1]));
			put_d("x1_after_restrict", level_x(mg.reg[1]));
			region *r1 = mg.reg[1];
			set_x(r1, test_field(r1->m, r1->n, r1->o, seed + 2));
			mg.interpolation(1);
			put_d("x_interp", level_x(r0));
		}
		/* Full V-cycles from x = 0 with the default (gs24) bottom. */
		mg.setup_fields(pr);
		std::vector<double> res;
		res.push_back(mg.l2_error_all());
		for (int c = 0; c < 4; c++) {
			mg.v_cycle();
			res.push_back(mg.l2_error_all());
			if (c == 0) put_d("x_v1", level_x(r0));
		}
		put_d("vcycle_res_gs24", res);
		put_d("x_v4_gs24", level_x(r0));
		put_i("gs_num_4cycles_plus", std::vector<int>(1, region::gs_num));
		/* The same with the bottom LU enabled after the operators exist. */
		if (mg.nlevels > 1) {
			enable_lu(mg);
			put_i("lu_enabled", std::vector<int>(1, mg.reg[mg.nlevels - 1]->Aexact != NULL));
			mg.setup_fields(pr);
			res.clear();
			res.push_back(mg.l2_error_all());
			for (int c = 0; c < 4; c++) {
				mg.v_cycle();
				res.push_back(mg.l2_error_all());
			}
			put_d("vcycle_res_lu", res);
			put_d("x_v4_lu", level_x(r0));
		}
		fclose(outf);
		printf("{\"ok\": true, \"levels\": %d}\n", mg.nlevels);
		fflush(stdout);
	} else if (cmd == "fmg") {
		/* One full multigrid cycle from x = 0 (multigrid.cc:109-134), then
		 * one V-cycle on top of it. */
		outf = fopen(argv[7], "wb");
		if (!outf) {perror("open"); return 1;}
		mg.fmg_cycle();
		put_d("x_fmg", level_x(mg.reg[0]));
		if (mg.nlevels > 1) put_d("x1_fmg", level_x(mg.reg[1]));
		put_d("l2_fmg", std::vector<double>(1, mg.l2_error_all()));
		mg.v_cycle();
		put_d("x_fmg_v", level_x(mg.reg[0]));
		put_d("l2_fmg_v", std::vector<double>(1, mg.l2_error_all()));
		fclose(outf);
		printf("{\"ok\": true, \"levels\": %d}\n", mg.nlevels);
		fflush(stdout);
	} else if (cmd == "slices") {
		/* Two V-cycles from x = 0, then z-slice k of x, r and the residual
		 * at levels 0 and 1 through the reference's own output routines. */
		int k = atoi(argv[7]);
		std::string prefix = argv[8];
		mg.v_cycle();
		mg.v_cycle();
		for (int l = 0; l < 2 && l < mg.nlevels; l++) {
			int kl = l == 0 ? k : k / 2;
			char nm[512];
			snprintf(nm, sizeof nm, "%s_x%d.bin", prefix.c_str(), l); mg.output_x(nm, kl, l);
			snprintf(nm, sizeof nm, "%s_r%d.bin", prefix.c_str(), l); mg.output_r(nm, kl, l);
			snprintf(nm, sizeof nm, "%s_res%d.bin", prefix.c_str(), l); mg.output_residual(nm, kl, l);
		}
		printf("{\"ok\": true}\n");
		fflush(stdout);
	} else if (cmd == "solve" || cmd == "time") {
		std::string bottom = argc > 7 ? argv[7] : "gs24";
		double tol = argc > 8 ? atof(argv[8]) : 1e-10;
		int maxc = argc > 9 ? atoi(argv[9]) : 100;
		if (bottom == "lu") enable_lu(mg);
		if (cmd == "solve" && argc > 10 && strcmp(argv[10], "-") != 0) outf = fopen(argv[10], "wb");
		// rhs_l2norm is the global sum on rank 0 only (region.cc:1549-1558);
		// every rank needs it to take the same stop decision
		double bn = mg.rhs_l2norm();
		MPI_Bcast(&bn, 1, MPI_DOUBLE, 0, MPI_COMM_WORLD);
		std::vector<double> res;
		int gs0 = region::gs_num;
		double ts = MPI_Wtime();
		int c = 0;
		double rel = 1;
		res.push_back(mg.l2_error_all());
		while (c < maxc) {
			mg.v_cycle();
			c++;
			double e = mg.l2_error_all();
			res.push_back(e);
			rel = sqrt(e / bn);
			if (!(rel >= tol)) break;
		}
		double tsolve = MPI_Wtime() - ts;
		if (outf) {
			put_d("res", res);
			put_d("x", level_x(mg.reg[0]));
			put_d("rhs_l2norm", std::vector<double>(1, bn));
			fclose(outf);
		}
		int rank = 0, size = 1;
		MPI_Comm_rank(MPI_COMM_WORLD, &rank);
		MPI_Comm_size(MPI_COMM_WORLD, &size);
		if (rank == 0) {
			printf("{\"cycles\": %d, \"rel\": %.17g, \"solve_s\": %.6f, \"setup_s\": %.6f, "
			       "\"levels\": %d, \"gs_sweeps_per_cycle\": %d, \"lu\": %d, \"ranks\": %d}\n",
			       c, rel, tsolve, t_setup, mg.nlevels, c ? (region::gs_num - gs0) / c : 0,
			       mg.reg[mg.nlevels - 1]->Aexact != NULL, size);
			fflush(stdout);
		}
	}
	mg.free();
	gm.free();
	MPI_Finalize();
	return 0;
}
