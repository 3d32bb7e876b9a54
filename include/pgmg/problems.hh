/** \file problems.hh
 * \brief Problem classes for the drop-in API: p_class implementations in the
 * sense of problem_simple.hh:14-73 (range_*, internal, mem_size,
 * fill_entries, external_ptr, x_field, r_field).
 *
 * synthetic_problem: the seeded families of pgmg_problems.h, every point an
 *   unknown with the [-1,1]^3 box clipped to the grid.
 * problem_simple: the reference's example problem (problem_simple.hh,
 *   problem_simple.cc): the 7-point cell-centred Laplacian {1,1,1,-6,1,1,1}
 *   on interior points, identity rows scaled by -0.375 on the boundary
 *   (Dirichlet points), sources +1 at (3,3,3) and -1 at (12,12,12). */
#ifndef PGMG_AMD_PROBLEMS_HH
#define PGMG_AMD_PROBLEMS_HH

#include "../pgmg_problems.h"
#include "geometry.hh"

class synthetic_problem {
	public:
		const bool x_prd, y_prd, z_prd;
		const int m, n, o;
		int family;
		uint64_t seed;
		synthetic_problem(geometry &gm, int family_, uint64_t seed_)
			: x_prd(gm.x_prd), y_prd(gm.y_prd), z_prd(gm.z_prd), m(gm.m), n(gm.n), o(gm.o),
			  family(family_), seed(seed_) {}
		inline int range_xd(int i, int, int) {return x_prd || i > 0 ? -1 : 0;}
		inline int range_xu(int i, int, int) {return x_prd || i < m - 1 ? 2 : 1;}
		inline int range_yd(int, int j, int) {return y_prd || j > 0 ? -1 : 0;}
		inline int range_yu(int, int j, int) {return y_prd || j < n - 1 ? 2 : 1;}
		inline int range_zd(int, int, int k) {return z_prd || k > 0 ? -1 : 0;}
		inline int range_zu(int, int, int k) {return z_prd || k < o - 1 ? 2 : 1;}
		inline bool internal(int, int, int) {return true;}
		inline int mem_size(int i, int j, int k) {
			return (range_xu(i, j, k) - range_xd(i, j, k)) * (range_yu(i, j, k) - range_yd(i, j, k))
			     * (range_zu(i, j, k) - range_zd(i, j, k));
		}
		void fill_entries(int i, int j, int k, double *&en) {
			double row[27];
			pgmg_row27_prd(family, seed, m, n, o, (x_prd ? 1 : 0) | (y_prd ? 2 : 0) | (z_prd ? 4 : 0), i, j, k, row);
			for (int dk = range_zd(i, j, k); dk < range_zu(i, j, k); dk++)
				for (int dj = range_yd(i, j, k); dj < range_yu(i, j, k); dj++)
					for (int di = range_xd(i, j, k); di < range_xu(i, j, k); di++)
						*(en++) = row[(dk + 1) * 9 + (dj + 1) * 3 + (di + 1)];
		}
		inline double *external_ptr(int, int, int) {return NULL;}
		inline double x_field(int, int, int) {return 0;}
		inline double r_field(int i, int j, int k) {return pgmg_rhs(seed, m, n, o, i, j, k);}
		/** For overwrite_r_field / update_r_field (region.hh:207-224). */
		inline double reset_r_field(int i, int j, int k) {return r_field(i, j, k);}
		inline double update_r_field(int i, int j, int k) {return 0.5 * r_field(i, j, k);}
};

class problem_simple {
	public:
		const bool x_prd, y_prd, z_prd;
		const int m, n, o;
		problem_simple(geometry &gm)
			: x_prd(gm.x_prd), y_prd(gm.y_prd), z_prd(gm.z_prd), m(gm.m), n(gm.n), o(gm.o) {
			for (int e = 0; e < 27; e++) stencil[e] = 0;
			stencil[4] = stencil[10] = stencil[12] = stencil[14] = stencil[16] = stencil[22] = 1;
			stencil[13] = -6;
			fix[0] = -0.375;
		}
		inline int range_xd(int i, int j, int k) {return interior(i, j, k) ? -1 : 0;}
		inline int range_xu(int i, int j, int k) {return interior(i, j, k) ? 2 : 1;}
		inline int range_yd(int i, int j, int k) {return interior(i, j, k) ? -1 : 0;}
		inline int range_yu(int i, int j, int k) {return interior(i, j, k) ? 2 : 1;}
		inline int range_zd(int i, int j, int k) {return interior(i, j, k) ? -1 : 0;}
		inline int range_zu(int i, int j, int k) {return interior(i, j, k) ? 2 : 1;}
		inline int mem_size(int, int, int) {return 0;}
		void fill_entries(int, int, int, double *&) {}
		inline double x_field(int, int, int) {return 0;}
		inline double r_field(int i, int j, int k) {
			return i == 3 && j == 3 && k == 3 ? 1 : (i == 12 && j == 12 && k == 12 ? -1 : 0);
		}
		inline bool internal(int, int, int) {return false;}
		inline double *external_ptr(int i, int j, int k) {return interior(i, j, k) ? stencil : fix;}
		inline bool interior(int i, int j, int k) {
			return (x_prd || (i != 0 && i != m - 1)) && (y_prd || (j != 0 && j != n - 1)) &&
			       (z_prd || (k != 0 && k != o - 1));
		}
	private:
		double stencil[27];
		double fix[1];
};

#endif
