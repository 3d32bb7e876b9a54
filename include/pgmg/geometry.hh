/** \file geometry.hh
 * \brief geometry of the drop-in API (reference: geometry.hh:8-75,
 * geometry.cc:12-26, 139-158).
 *
 * The reference splits the grid over a minimum-surface processor grid with a
 * Cartesian MPI communicator.  Here one process drives one GPU that holds the
 * whole grid, which is the reference's one-rank case: procs = 1, the
 * processor grid is 1x1x1, this rank owns [0,m)x[0,n)x[0,o). */
#ifndef PGMG_AMD_GEOMETRY_HH
#define PGMG_AMD_GEOMETRY_HH

#include "common.hh"

class geometry {
	public:
		/** Global grid size. */
		const int m, n, o;
		/** Periodicity per axis. */
		const bool x_prd, y_prd, z_prd;
		/** Rank and processor count (0 and 1). */
		int rank, procs;
		/** Processor grid size and this rank's position in it. */
		int mp, np, op, ip, jp, kp;
		/** Owned global index ranges [ai,bi) etc. */
		int ai, aj, ak, bi, bj, bk;
		/** Owned sizes. */
		int sm, sn, so, smn, smno;
		/** Processor table (unused with one rank). */
		int *ranks;
		geometry(int m_, int n_, int o_, bool x_prd_, bool y_prd_, bool z_prd_)
			: m(m_), n(n_), o(o_), x_prd(x_prd_), y_prd(y_prd_), z_prd(z_prd_),
			  rank(0), procs(1), mp(1), np(1), op(1), ip(0), jp(0), kp(0),
			  ai(0), aj(0), ak(0), bi(m_), bj(n_), bk(o_), sm(m_), sn(n_), so(o_),
			  smn(m_ * n_), smno(m_ * n_ * o_), ranks(NULL) {
			if (m < 1 || n < 1 || o < 1) p_fatal_error("No valid processor decomposition", PGMG_MATH_ERROR);
		}
		/** geometry::free (geometry.cc): releases the communicator; nothing
		 * to release here. */
		void free() {}
		/** 27-neighbour table (geometry.cc:164-172): only the rank itself, and
		 * periodic wrap-arounds to itself. */
		void set_up_neighbors(int *neigh) {
			for (int k = -1; k <= 1; k++) for (int j = -1; j <= 1; j++) for (int i = -1; i <= 1; i++, neigh++)
				*neigh = ((z_prd || k == 0) && (y_prd || j == 0) && (x_prd || i == 0)) ? 0 : -1;
		}
};

#endif
