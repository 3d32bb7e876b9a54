/** \file multigrid.hh
 * \brief multigrid of the drop-in API (reference: multigrid.hh:20-319,
 * multigrid.cc:7-292).
 *
 * The constructor builds the whole level hierarchy in HBM (one pgmg_hier
 * handle); every level operation runs as gfx950 kernels.  The default
 * options reproduce the reference's arithmetic: FP64, the exact (hyperplane)
 * Gauss-Seidel smoother, the 24-sweep bottom.  An optional fourth constructor
 * argument selects the eight-colour smoother, mixed precision or the bottom
 * LU.  reg[l] are host views of the levels (see region.hh for the mirror
 * semantics); nlevels == mlevels (one rank). */
#ifndef PGMG_AMD_MULTIGRID_HH
#define PGMG_AMD_MULTIGRID_HH

#include <cmath>
#include <limits>
#include <vector>

#include "buffer.hh"
#include "common.hh"
#include "geometry.hh"
#include "region.hh"

class multigrid {
	public:
		/** Cycle-count predictor and tolerance of solve_v_cycles
		 * (multigrid.hh:34-118, multigrid.cc:268-292), same constants and
		 * update rule. */
		struct mg_guide {
			static const int DEF_PERIOD = 16;
			static const int INIT_VAR = 1000000;
			static const int DEF_ALPHA = 3;
			static const int MAX_ITERS = 100;
			static const int FUDGE_FACTOR = 10000;
			static const int INSURANCE = 2;
			int period, iter;
			bool use_avg;
			double iter_avg, iter_var, alpha, tol;
			bool tol_set_manually;
			mg_guide() : use_avg(false), tol_set_manually(false) {
				set_period(DEF_PERIOD);
				alpha = (1. * DEF_ALPHA) / period;
			}
			/** Tolerance from a characteristic scale of one node's value. */
			inline void set_tol(double p_scale) {
				const double eps = std::numeric_limits<double>::epsilon();
				tol = p_scale * p_scale * eps * eps * FUDGE_FACTOR;
				tol_set_manually = true;
			}
			inline void average_on() {use_avg = true;}
			inline void average_on(double alpha_) {use_avg = true; alpha = alpha_;}
			inline void average_off() {use_avg = false;}
			inline void set_period(int period_) {
				period = period_;
				iter_avg = iter = period;
				iter_var = INIT_VAR;
			}
			/** Lower the iteration index by one per solve and raise it by a
			 * period per extra cycle needed; with averaging, also lower it by
			 * each standard deviation the index exceeds its moving mean. */
			void update(int extras) {
				int diter = period * extras - 1;
				if (use_avg) {
					const double di = iter - iter_avg;
					const int zscore = static_cast<int>(di / std::sqrt(iter_var));
					if (zscore > 0) diter -= zscore;
					iter_avg = (1 - alpha) * iter_avg + alpha * iter;
					iter_var = (1 - alpha) * (iter_var + alpha * di * di);
				}
				iter += diter;
				if (iter < period) iter = 2 * period - 1;
			}
			inline int pred_iters() {return iter / period;}
		};

		int v_down, v_bottom, v_up;
		int rank, nlevels, mlevels;
		mg_guide guide;
		geometry *geom;
		region **reg;
		geometry **geo;
		/** The device hierarchy and its options. */
		pgmg_hier *hier;
		pgmg_opts opts;
		/** Whether setup_matrices / compute_rats fill the host mirrors
		 * reg[l]->S, A, Am, Asize (the reference's own storage; about 230
		 * bytes per point of every level).  Default on. */
		bool mirror_operator;

		template <class p_class>
		multigrid(p_class &pr, geometry &gm, comm_buffer &com_, const pgmg_opts *opts_ = NULL)
			: v_down(2), v_bottom(20), v_up(2), rank(0), nlevels(0), mlevels(0), geom(&gm), hier(NULL),
			  mirror_operator(true), com(com_) {
			(void)pr;
			if (opts_) opts = *opts_;
			else {
				pgmg_default_opts(&opts);
				opts.smoother = PGMG_SMOOTH_EXACT;
			}
			v_down = opts.v_down; v_up = opts.v_up; v_bottom = opts.v_bottom;
			pgmg_dims d = {gm.m, gm.n, gm.o, gm.x_prd ? 1 : 0, gm.y_prd ? 1 : 0, gm.z_prd ? 1 : 0};
			pgmg_check(pgmg_hier_create(&d, &opts, &hier), "multigrid");
			pgmg_check(pgmg_num_levels(hier, &nlevels), "multigrid");
			mlevels = nlevels;
			reg = new region *[mg_max_levels];
			geo = new geometry *[mg_max_levels];
			for (int l = 0; l < nlevels; l++) {
				int dl[3];
				pgmg_check(pgmg_level_dims(hier, l, dl), "multigrid");
				reg[l] = new region(l ? reg[l - 1] : NULL, hier, l, dl[0], dl[1], dl[2], gm.x_prd, gm.y_prd, gm.z_prd);
				geo[l] = NULL;
			}
		}
		~multigrid() {
			for (int l = nlevels - 1; l >= 0; l--) delete reg[l];
			delete[] reg;
			delete[] geo;
			pgmg_hier_destroy(hier);
		}

		/** multigrid.hh:301-305: top operator, then the Galerkin levels. */
		template <class p_class>
		void setup_matrices(p_class &pr) {
			reg[0]->setup_matrices(pr);
			compute_rats();
		}
		template <class p_class>
		void setup_fields(p_class &pr) {reg[0]->setup_fields(pr);}
		template <class p_class>
		void setup_r_field(p_class &pr) {reg[0]->setup_r_field(pr);}
		inline void allclear_x_field() {reg[0]->allclear_x_field();}
		/** multigrid::compute_rats (multigrid.cc:138-152): the Galerkin
		 * operators of every coarse level on the device (and the bottom LU
		 * again where one is live), then the host mirrors: the coarse boxes
		 * by the reference's bound rule (fill_rat_bounds) from the top
		 * level's boxes, and the entries inside them. */
		void compute_rats() {
			pgmg_check(pgmg_compute_rats(hier), "compute_rats");
			reg[nlevels - 1]->refresh_exact();
			if (!mirror_operator || reg[0]->top_boxes.empty()) return;
			std::vector<int> boxes(reg[0]->top_boxes);
			reg[0]->mirror_operator(boxes);
			for (int l = 1; l < nlevels; l++) {
				std::vector<int> coarse;
				const region &F = *reg[l - 1], &C = *reg[l];
				pgmg_detail::rat_bounds(&boxes[0], F.m, F.n, F.o, F.prd_mask(), C.m, C.n, C.o, coarse);
				reg[l]->mirror_operator(coarse);
				boxes.swap(coarse);
			}
		}
		/** multigrid::diagnostic_S / _A / _x (multigrid.hh:256-263). */
		inline void diagnostic_S(int level = 0) {if (level < nlevels) reg[level]->diagnostic_S();}
		inline void diagnostic_A(int level = 0) {if (level < nlevels) reg[level]->diagnostic_A();}
		inline void diagnostic_x(int level = 0) {if (level < nlevels) reg[level]->diagnostic_x();}

		/** multigrid::v_cycle (multigrid.cc:88-106). */
		void v_cycle(int mode = 0) {
			sync_sweeps();
			long long a, b;
			pgmg_gs_count(hier, &a);
			pgmg_check(pgmg_v_cycle(hier, mode), "v_cycle");
			pgmg_gs_count(hier, &b);
			region::gs_num += (int)(b - a);
		}
		/** multigrid::fmg_cycle (multigrid.cc:109-134). */
		void fmg_cycle() {
			int i, j, l;
			const int mode = 0;
			for (i = 0; i < mlevels - 1; i++) {
				restriction(i + 1);
				for (l = 0; l < v_down; l++) gauss_seidel(mode, i + 1);
			}
			for (j = mlevels - 1; j > 0; j--) {
				for (i = j; i < mlevels - 1; i++) {
					restriction(i + 1);
					for (l = 0; l < v_down; l++) gauss_seidel(mode, i + 1);
				}
				for (l = 0; l < v_bottom; l++) gauss_seidel(mode, mlevels - 1);
				for (i = mlevels - 1; i >= j; i--) {
					interpolation(i);
					for (l = 0; l < v_up; l++) gauss_seidel(mode, i - 1);
				}
			}
		}
		double l2_error(int level = 0) {return nlevels > level ? reg[level]->l2_error() : 0;}
		double l2_error_all(int level = 0) {return nlevels > level ? reg[level]->l2_error_all() : 0;}
		double l0_error(int level = 0) {return nlevels > level ? reg[level]->l0_error() : 0;}
		double l0_error_all(int level = 0) {return nlevels > level ? reg[level]->l0_error_all() : 0;}
		double rhs_l2norm(int level = 0) {return nlevels > level ? reg[level]->rhs_l2norm() : 0;}
		/** multigrid::gauss_seidel (multigrid.cc:180-182). */
		void gauss_seidel(int mode = 0, int level = 0) {
			if (level < nlevels && reg[level]->gs_enable) reg[level]->gauss_seidel(mode);
		}
		/** multigrid::interpolation / restriction (multigrid.cc:196-224). */
		void interpolation(int level) {
			if (level < nlevels && level > 0) pgmg_check(pgmg_interpolation(hier, level), "interpolation");
		}
		void restriction(int level) {
			if (level < nlevels && level > 0) pgmg_check(pgmg_restriction(hier, level), "restriction");
		}
		/** Halo exchange: nothing to exchange on one GPU. */
		void communicate(int level = 0) {(void)level;}
		inline void solve_exact(int level = 0, bool smooth = 0) {
			if (level < nlevels) {
				if (!smooth && level == nlevels - 1) {
					int lu = 0;
					pgmg_lu_active(hier, &lu);
					if (!lu) region::gs_num += pgmg_gs_exact;
				} else region::gs_num += pgmg_gs_exact;
				reg[level]->solve_exact(smooth);
			}
		}

		/** multigrid::solve_v_cycles (multigrid.hh:182-219): the predicted
		 * number of cycles, then 1, 2, 3... more until sum res^2 <= tol*m*n*o,
		 * then INSURANCE cycles.  0 = converged, 1 = MAX_ITERS exceeded,
		 * 2 = NaN. */
		int solve_v_cycles(double &res) {
			if (!guide.tol_set_manually) p_fatal_error(
				"multigrid::solve_v_cycles cannot be used before the class mg_guide instance uses"
				" the function multigrid::mg_pred::set_tol to set the target tolerance to be reached", 1);
			double tol = guide.tol * geom->m;
			tol *= geom->n;
			tol *= geom->o;
			for (int i = 0; i < guide.pred_iters(); i++) v_cycle();
			int extra_rounds = 0, n = 0;
			while ((res = l2_error_all()) > tol && n < guide.MAX_ITERS) {
				if (std::isnan(res)) return 2;
				extra_rounds++;
				for (int i = 0; i < extra_rounds; i++, n++) v_cycle();
			}
			if (n > guide.MAX_ITERS) return 1;
			for (int i = 0; i < guide.INSURANCE; i++) v_cycle();
			guide.update(extra_rounds * (extra_rounds + 1) / 2);
			return 0;
		}
		int solve_v_cycles() {double res; return solve_v_cycles(res);}

		/** Solve to a relative residual sqrt(l2_error_all/rhs_l2norm) < tol with
		 * V-cycles or, for mixed precision, defect correction (no counterpart
		 * in the reference; its drivers loop v_cycle and l2_error_all). */
		int solve_relative(double tol, int max_cycles, int *cycles = NULL) {
			sync_sweeps();
			int c = 0;
			int rc = pgmg_solve(hier, tol, max_cycles, &c, NULL);
			if (cycles) *cycles = c;
			return rc;
		}

		/** PCG helpers (multigrid.hh:222-251), on the top level. */
		inline void compute_residual(double *ext_r) {reg[0]->compute_residual(ext_r);}
		void Ax(const double *in, double *out) {reg[0]->Ax(in, out);}
		void unpadded_Ax(const double *in, double *out, const int lev = 0) {
			if (lev >= nlevels) p_fatal_error("Unpadded_Ax only available for single MPI process, within multigrid hierarchy.", 1);
			reg[lev]->unpadded_Ax(in, out);
		}
		void copy_rhs_to_soln() {reg[0]->copy_rhs_to_soln();}
		double source_soln_dot() {return reg[0]->source_soln_dot();}
		template <class p_class>
		void update_r_field(p_class &pr) {reg[0]->update_r_field(pr);}
		template <class p_class>
		void overwrite_r_field(p_class &pr) {reg[0]->overwrite_r_field(pr);}
		double soln_diff(double *ext_x) {return reg[0]->soln_diff(ext_x);}
		/** z-slice output (multigrid.cc:158-176). */
		void output_x(const char *filename, int k, int level = 0) {if (level < nlevels) reg[level]->output_x(filename, k);}
		void output_r(const char *filename, int k, int level = 0) {if (level < nlevels) reg[level]->output_r(filename, k);}
		void output_residual(const char *filename, int k, int level = 0) {
			if (level < nlevels) reg[level]->output_residual(filename, k);
		}
		inline void free() {}
		inline void setup_test(int ca, bool xfield, int level = 0) {
			if (level < nlevels) reg[level]->setup_test(ca, xfield);
		}
	protected:
		comm_buffer &com;
	private:
		void sync_sweeps() {pgmg_check(pgmg_set_sweeps(hier, v_down, v_up), "sweeps");}
		multigrid(const multigrid &);
		multigrid &operator=(const multigrid &);
};

#endif
