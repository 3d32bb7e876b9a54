/** \file region.hh
 * \brief region of the drop-in API (reference: region.hh:12-529).
 *
 * A region is one level of a multigrid hierarchy.  In the reference it owns
 * the level's arrays in host memory; here the arrays live in HBM inside a
 * pgmg_hier handle and a region is a view of one level of it.  The host
 * arrays x0 and r0 are mirrors: call sync_from_device() before reading them
 * and sync_to_device() after writing them.  The level layout is the
 * reference's one-rank layout: no ghost points, mg = sm, ng = sn, x0 = x,
 * index i + mg*(j + ng*k).
 *
 * region(p_class&, geometry&, comm_buffer&) builds a lone top-level region,
 * as gs_parallel.cc:25 does; regions inside a multigrid are made by it. */
#ifndef PGMG_AMD_REGION_HH
#define PGMG_AMD_REGION_HH

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "buffer.hh"
#include "common.hh"
#include "geometry.hh"

/** Static sweep counter shared by all regions (region.hh:17), in a template
 * so that the header defines it without a separate translation unit. */
template <int D> struct pgmg_region_counter {static int gs_num;};
template <int D> int pgmg_region_counter<D>::gs_num = 0;

namespace pgmg_detail {

/** Ingests a p_class operator (region.hh:474-529, 95-101) into 27-entry rows.
 * Each point's box [range_xd, range_xu) x ... must lie inside [-1,1]^3, and
 * inside the grid along the non-periodic axes (prd: bit 0 x, bit 1 y, bit 2
 * z); along a periodic axis an offset that leaves the grid couples to the
 * wrapped point.  The entries of internal points come from fill_entries in
 * k,j,i order, those of the others from external_ptr, in the box's k,j,i
 * order. */
template <class p_class>
void ingest_operator(p_class &pr, int m, int n, int o, int prd, std::vector<double> &A27, std::vector<int> *S6 = NULL) {
	A27.assign((size_t)m * n * o * 27, 0.0);
	if (S6) S6->resize((size_t)m * n * o * 6);
	std::vector<double> tmp(27);
	for (int k = 0; k < o; k++) for (int j = 0; j < n; j++) for (int i = 0; i < m; i++) {
		int xd = pr.range_xd(i, j, k), xu = pr.range_xu(i, j, k);
		int yd = pr.range_yd(i, j, k), yu = pr.range_yu(i, j, k);
		int zd = pr.range_zd(i, j, k), zu = pr.range_zu(i, j, k);
		if (xd < -1 || xu > 2 || yd < -1 || yu > 2 || zd < -1 || zu > 2 || xd >= xu || yd >= yu || zd >= zu)
			p_fatal_error("Stencil box outside [-1,1]^3 is not supported", PGMG_SETUP_ERROR);
		if ((!(prd & 1) && (i + xd < 0 || i + xu > m)) || (!(prd & 2) && (j + yd < 0 || j + yu > n)) ||
		    (!(prd & 4) && (k + zd < 0 || k + zu > o)))
			p_fatal_error("Stencil box reaches outside the grid along a non-periodic axis", PGMG_SETUP_ERROR);
		const double *src;
		if (pr.internal(i, j, k)) {
			int sz = pr.mem_size(i, j, k);
			if (sz != (xu - xd) * (yu - yd) * (zu - zd)) p_fatal_error("mem_size differs from the box size", PGMG_SETUP_ERROR);
			double *en = &tmp[0];
			pr.fill_entries(i, j, k, en);
			src = &tmp[0];
		} else {
			src = pr.external_ptr(i, j, k);
		}
		const size_t pt = (size_t)i + (size_t)m * (j + (size_t)n * k);
		if (S6) {
			int *Sp = &(*S6)[6 * pt];
			Sp[0] = xd; Sp[1] = xu; Sp[2] = yd; Sp[3] = yu; Sp[4] = zd; Sp[5] = zu;
		}
		double *row = &A27[27 * pt];
		for (int dk = zd; dk < zu; dk++) for (int dj = yd; dj < yu; dj++) for (int di = xd; di < xu; di++)
			row[(dk + 1) * 9 + (dj + 1) * 3 + (di + 1)] = *(src++);
	}
}

/** The fine points (unwrapped indices) whose interpolation weight onto
 * coarse point I is non-zero, along an axis of fine size M (region.cc:665-746:
 * the 2x2x2 cube scan and its masks at one rank; coarse size (M+2)>>1, or
 * (M+1)>>1 when periodic).  On an even non-periodic axis the last fine point
 * is injected into the last coarse point only; on an even periodic axis the
 * last cube reaches coarse point 0 across the wrap (fine -1 of coarse 0). */
inline int fine_support(int I, int M, bool prd, int *f) {
	int c = 0;
	if (prd && !(M & 1)) {
		for (int d = -1; d <= 1; d++) f[c++] = 2 * I + d;
		return c;
	}
	if (!(M & 1) && 2 * I == M) {  // the last coarse point of an even axis
		f[c++] = M - 1;
		return c;
	}
	const int top = (M & 1) ? M - 1 : M - 2;
	for (int d = -1; d <= 1; d++)
		if (2 * I + d >= 0 && 2 * I + d <= top) f[c++] = 2 * I + d;
	return c;
}

/** region::grid_map (region.cc:1201-1217): a fine index range [lo, hi) on an
 * axis of fine size M to the coarse points whose interpolation reaches it. */
inline void grid_map(int &lo, int &hi, int M, bool prd) {
	if (prd) {
		lo = lo < 0 ? (lo - ((M & 1) ? 2 : 1)) / 2 : lo >> 1;
		hi = (hi + (((M & 1) && hi > M) ? 3 : 2)) >> 1;
	} else {
		lo = (!(M & 1) && lo == M - 1 ? lo + 1 : lo) >> 1;
		hi = (hi + 2) >> 1;
	}
}

/** The coarse stencil boxes of the Galerkin operator (region::fill_rat_bounds,
 * region.cc:598-623, with rat_bound_full/partial, b_red and b_ex,
 * region.cc:916-1003, 1105-1110, 1276-1295): coarse point I starts with the
 * box [0,1)^3 and is widened by the coarse image (grid_map) of the box of
 * every fine point that interpolates from it.  Sf and Sc hold six bounds per
 * point (xd, xu, yd, yu, zd, zu) in point order i + m*(j + n*k). */
inline void rat_bounds(const int *Sf, int mf, int nf, int of, int prd, int mc, int nc, int oc, std::vector<int> &Sc) {
	Sc.assign((size_t)mc * nc * oc * 6, 0);
	const bool px = prd & 1, py = (prd & 2) != 0, pz = (prd & 4) != 0;
	int fi[3], fj[3], fk[3];
	for (int K = 0; K < oc; K++) {
		const int ck = fine_support(K, of, pz, fk);
		for (int J = 0; J < nc; J++) {
			const int cj = fine_support(J, nf, py, fj);
			for (int I = 0; I < mc; I++) {
				const int ci = fine_support(I, mf, px, fi);
				int b[6] = {0, 1, 0, 1, 0, 1};
				for (int c3 = 0; c3 < ck; c3++) for (int c2 = 0; c2 < cj; c2++) for (int c1 = 0; c1 < ci; c1++) {
					const int gi = fi[c1], gj = fj[c2], gk = fk[c3];
					const int wi = (gi % mf + mf) % mf, wj = (gj % nf + nf) % nf, wk = (gk % of + of) % of;
					const int *Sp = Sf + 6 * ((size_t)wi + (size_t)mf * (wj + (size_t)nf * wk));
					int lo[3] = {gi + Sp[0], gj + Sp[2], gk + Sp[4]}, hi[3] = {gi + Sp[1], gj + Sp[3], gk + Sp[5]};
					grid_map(lo[0], hi[0], mf, px);
					grid_map(lo[1], hi[1], nf, py);
					grid_map(lo[2], hi[2], of, pz);
					const int ctr[3] = {I, J, K};
					for (int a = 0; a < 3; a++) {
						if (lo[a] - ctr[a] < b[2 * a]) b[2 * a] = lo[a] - ctr[a];
						if (hi[a] - ctr[a] > b[2 * a + 1]) b[2 * a + 1] = hi[a] - ctr[a];
					}
				}
				int *Sp = &Sc[6 * ((size_t)I + (size_t)mc * (J + (size_t)nc * K))];
				for (int a = 0; a < 6; a++) Sp[a] = b[a];
			}
		}
	}
}

/** Size of a box (region.hh:430-433, rat_size). */
inline int box_size(const int *Sp) {return (Sp[1] - Sp[0]) * (Sp[3] - Sp[2]) * (Sp[5] - Sp[4]);}

}  // namespace pgmg_detail

class region : public pgmg_region_counter<0> {
	public:
		/** Parent (finer) region, or NULL at the top. */
		region *p;
		int rank;
		const bool x_prd, y_prd, z_prd;
		bool gs_enable;
		/** Global level size. */
		const int m, n, o;
		int ip, jp, kp, mp, np, op;
		/** Host mirrors of the solution and source (sync_from_device first). */
		double *x, *x0, *r0;
		int mg, ng, og, mng, indent;
		int ai, aj, ak, bi, bj, bk, sm, sn, so, smn, smno;
		/** Host mirror of the level operator in the reference's storage
		 * (region.hh:57-61, 256): S holds six box bounds per point (xd, xu,
		 * yd, yu, zd, zu, offsets from the point), A[p] points at point p's
		 * entries in Am, in the box's k, j, i order, Asize is the number of
		 * entries.  Filled when the operator is set up (setup_matrices,
		 * compute_rats) unless the hierarchy's mirror_operator is off;
		 * NULL before.  Changing them does not change the device operator. */
		int *S;
		double **A;
		double *Am;
		int Asize;
		/** The bottom LU factor as dgetrf leaves it and its 1-based pivots
		 * (region_tr.cc:472-534), NULL when the level has no live factor. */
		double *Aexact;
		int *ipiv;
		/** The device hierarchy and the level this region views. */
		pgmg_hier *hier;
		int level;

		/** A lone top-level region (gs_parallel.cc:25): one device level. */
		template <class p_class>
		region(p_class &pr, geometry &gm, comm_buffer &com_)
			: p(NULL), rank(0), x_prd(gm.x_prd), y_prd(gm.y_prd), z_prd(gm.z_prd), gs_enable(true),
			  m(gm.m), n(gm.n), o(gm.o), S(NULL), A(NULL), Am(NULL), Asize(0), Aexact(NULL), ipiv(NULL),
			  hier(NULL), level(0), owns(true), exact_enabled(false) {
			(void)com_;
			pgmg_dims d = {m, n, o, x_prd ? 1 : 0, y_prd ? 1 : 0, z_prd ? 1 : 0};
			pgmg_opts opt;
			pgmg_default_opts(&opt);
			opt.smoother = PGMG_SMOOTH_EXACT;
			opt.max_levels = 1;
			pgmg_check(pgmg_hier_create(&d, &opt, &hier), "region");
			init_fields();
		}
		/** A view of level l of an existing hierarchy (used by multigrid). */
		region(region *parent, pgmg_hier *h, int l, int m_, int n_, int o_, bool xp = false, bool yp = false,
		       bool zp = false)
			: p(parent), rank(0), x_prd(xp), y_prd(yp), z_prd(zp), gs_enable(true),
			  m(m_), n(n_), o(o_), S(NULL), A(NULL), Am(NULL), Asize(0), Aexact(NULL), ipiv(NULL),
			  hier(h), level(l), owns(false), exact_enabled(false) {init_fields();}
		~region() {
			delete[] x;
			delete[] r0;
			clear_operator_mirror();
			delete[] Aexact;
			delete[] ipiv;
			if (owns) pgmg_hier_destroy(hier);
		}
		inline int prd_mask() const {return (x_prd ? 1 : 0) | (y_prd ? 2 : 0) | (z_prd ? 4 : 0);}

		/** Fills S, A, Am and Asize from box bounds (six per point) and the
		 * device operator of this level. */
		void mirror_operator(const std::vector<int> &boxes) {
			clear_operator_mirror();
			const size_t N = (size_t)smno;
			S = new int[6 * N];
			memcpy(S, &boxes[0], sizeof(int) * 6 * N);
			long long tot = 0;
			for (size_t q = 0; q < N; q++) tot += pgmg_detail::box_size(S + 6 * q);
			Asize = (int)tot;
			Am = new double[(size_t)tot];
			A = new double *[N];
			std::vector<double> A27(27 * N);
			pgmg_check(pgmg_get_operator27(hier, level, &A27[0]), "operator mirror");
			double *ap = Am;
			for (size_t q = 0; q < N; q++) {
				const int *Sp = S + 6 * q;
				A[q] = ap;
				for (int dk = Sp[4]; dk < Sp[5]; dk++) for (int dj = Sp[2]; dj < Sp[3]; dj++)
					for (int di = Sp[0]; di < Sp[1]; di++)
						*(ap++) = A27[27 * q + (dk + 1) * 9 + (dj + 1) * 3 + (di + 1)];
			}
		}
		void clear_operator_mirror() {
			delete[] S; delete[] A; delete[] Am;
			S = NULL; A = NULL; Am = NULL; Asize = 0;
		}
		/** region::enable_exact (region_tr.cc:472-478): with one rank, a
		 * level of fewer than pgmg_max_exact points is eligible; here only the
		 * bottom level of a hierarchy can hold the factor. */
		void enable_exact() {exact_enabled = smno < pgmg_max_exact && is_bottom();}
		/** region::lu_decomposition (region_tr.cc:483-534): assemble and
		 * factor the level operator; on failure print the reference's
		 * message and fall back on Gauss-Seidel (Aexact = NULL). */
		void lu_decomposition() {
			if (!exact_enabled) return;
			pgmg_check(pgmg_lu_decomposition(hier), "lu_decomposition");
			refresh_exact();
			if (!Aexact) printf("#-->LAPACK LU decomposition failed\n#   Falling back on Gauss-Seidel\n");
			exact_enabled = Aexact != NULL;
		}
		/** Reloads Aexact / ipiv after the device refactored the bottom. */
		void refresh_exact() {
			delete[] Aexact;
			delete[] ipiv;
			Aexact = NULL;
			ipiv = NULL;
			int on = 0;
			pgmg_check(pgmg_lu_active(hier, &on), "lu_active");
			if (!on || !is_bottom()) return;
			Aexact = new double[(size_t)smno * smno];
			ipiv = new int[(size_t)smno];
			pgmg_check(pgmg_get_lu_factor(hier, Aexact, ipiv), "lu factor");
		}
		/** region::diagnostic_S / diagnostic_A / diagnostic_x
		 * (region.cc:1766-1805): the boxes, the operator rows with their
		 * sums, and x, printed in the reference's formats. */
		void diagnostic_S() {
			if (!S) p_fatal_error("No operator mirror to print", PGMG_SETUP_ERROR);
			for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++) {
				const int *Sp = S + 6 * (i + sm * (j + sn * k));
				printf("%d: (%d,%d,%d) [%d:%d] [%d:%d] [%d:%d]\n", rank, i, j, k, Sp[0], Sp[1], Sp[2], Sp[3], Sp[4], Sp[5]);
			}
		}
		void diagnostic_A() {
			if (!S) p_fatal_error("No operator mirror to print", PGMG_SETUP_ERROR);
			for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++) {
				const int l = i + sm * (j + sn * k), ss = pgmg_detail::box_size(S + 6 * l);
				const double *Amp = A[l];
				double tot = 0;
				printf("%d: (%d,%d,%d) [", rank, i + ai, j + aj, k + ak);
				int q;
				for (q = 0; q < ss - 1; q++) {
					printf("%g,", Amp[q]);
					tot += Amp[q];
				}
				printf("%g] ", Amp[q]);
				tot += Amp[q];
				printf(" tot %g\n", tot);
			}
		}
		void diagnostic_x() {
			pgmg_check(pgmg_get_x(hier, level, x0), "diagnostic_x");
			printf("rhs size (%d %d %d), solution size (%d %d %d)\n", sm, sn, so, mg, ng, og);
			for (int k = 0; k < og; k++) for (int j = 0; j < ng; j++) for (int i = 0; i < mg; i++)
				printf("(%d, %d, %d): %g\n", i, j, k, x[i + j * mg + k * mng]);
		}

		/** region::setup_matrices (region.hh:95-101): the top-level operator
		 * from the problem class.  On a lone region this also finalises the
		 * operator (there are no coarse levels to compute). */
		template <class p_class>
		void setup_matrices(p_class &pr) {
			std::vector<double> A27;
			std::vector<int> boxes;
			pgmg_detail::ingest_operator(pr, m, n, o, prd_mask(), A27, &boxes);
			pgmg_check(pgmg_set_operator27(hier, &A27[0]), "setup_matrices");
			std::vector<double>().swap(A27);
			top_boxes.swap(boxes);
			if (owns) {
				pgmg_check(pgmg_compute_rats(hier), "setup_matrices");
				mirror_operator(top_boxes);
			}
		}
		/** The top level's boxes as the problem class gave them. */
		std::vector<int> top_boxes;
		template <class p_class>
		void setup_fields(p_class &pr) {setup_x_field(pr); setup_r_field(pr);}
		/** region.hh:110-116. */
		template <class p_class>
		void setup_x_field(p_class &pr) {
			for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++)
				x0[i + mg * (j + k * ng)] = pr.x_field(i + ai, j + aj, k + ak);
			sync_x_to_device();
		}
		/** region.hh:118-126. */
		template <class p_class>
		void setup_r_field(p_class &pr) {
			double *rp = r0;
			for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++)
				*(rp++) = pr.r_field(i + ai, j + aj, k + ak);
			sync_r_to_device();
		}

		/** Mirrors <-> device. */
		void sync_from_device() {
			pgmg_check(pgmg_get_x(hier, level, x0), "sync");
			pgmg_check(pgmg_get_rhs(hier, level, r0), "sync");
		}
		void sync_to_device() {sync_x_to_device(); sync_r_to_device();}
		void sync_x_to_device() {pgmg_check(pgmg_set_x(hier, level, x0), "sync");}
		void sync_r_to_device() {pgmg_check(pgmg_set_rhs(hier, level, r0), "sync");}

		/** region::gauss_seidel (region.cc:355-542). */
		void gauss_seidel(int mode = 0) {
			gs_num++;
			pgmg_check(pgmg_gauss_seidel(hier, mode, level, 1), "gauss_seidel");
		}
		/** region::solve_exact (region_tr.cc:540-558). */
		void solve_exact(bool smooth = 0) {
			pgmg_check(pgmg_solve_exact(hier, level, smooth ? 1 : 0), "solve_exact");
		}
		/** Sums of squared residuals (region.cc:1475-1529) and of the
		 * source (region.cc:1549-1558); with one rank l2_error = l2_error_all. */
		double l2_error_all() {double v; pgmg_check(pgmg_l2_error_all(hier, level, &v), "l2_error_all"); return v;}
		double l2_error() {return l2_error_all();}
		double rhs_l2norm() {double v; pgmg_check(pgmg_rhs_l2norm(hier, level, &v), "rhs_l2norm"); return v;}
		/** Max-norm residuals (region.cc:1492-1546). */
		double l0_error_all() {
			std::vector<double> res((size_t)smno);
			pgmg_check(pgmg_residual(hier, level, &res[0]), "l0_error_all");
			double e = 0;
			for (size_t q = 0; q < res.size(); q++) if (fabs(res[q]) > e) e = fabs(res[q]);
			return e;
		}
		double l0_error() {return l0_error_all();}
		/** region::compute_residual (region.cc:1848-1854) into an array of
		 * the level's size. */
		void compute_residual(double *ext_r) {pgmg_check(pgmg_residual(hier, level, ext_r), "compute_residual");}
		/** region.cc:1672-1687. */
		void allclear_x_field() {pgmg_check(pgmg_clear_x(hier, level), "allclear_x_field");}
		void clear_x_field() {allclear_x_field();}
		void clear_r_field() {
			memset(r0, 0, sizeof(double) * (size_t)smno);
			sync_r_to_device();
		}
		/** region.cc:1690-1701. */
		void copy_r_to_x() {sync_from_device(); memcpy(x0, r0, sizeof(double) * (size_t)smno); sync_x_to_device();}
		void copy_x_to_r() {sync_from_device(); memcpy(r0, x0, sizeof(double) * (size_t)smno); sync_r_to_device();}
		/** region::setup_test (region.cc:1812-1826). */
		void setup_test(int ca, bool xfield) {
			sync_from_device();
			for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++) {
				double *pp = xfield ? x0 + (i + mg * (j + ng * k)) : r0 + (i + sm * (j + sn * k));
				switch (ca) {
					case 0: *pp = i + ai; break;
					case 1: *pp = j + aj; break;
					case 2: *pp = k + ak; break;
					case 3: *pp = ak + aj + ai + k + j + i; break;
					case 4: *pp = rank; break;
					case 5: *pp = i == 7 && j == 7 && k == 7 ? 1 : 0; break;
					case 6: *pp = 1;
				}
			}
			sync_to_device();
		}
		/** Operations on outside arrays of the level's size, for
		 * preconditioned CG: region::Ax (region.cc:1907-1916) and, at one
		 * rank without periodic axes the same, unpadded_Ax
		 * (region.cc:1880-1905). */
		void Ax(const double *in, double *out) {pgmg_check(pgmg_apply_operator(hier, level, in, out), "Ax");}
		void unpadded_Ax(const double *in, double *out) {Ax(in, out);}
		/** region::source_soln_dot (region.cc:1918-1930). */
		double source_soln_dot() {double v; pgmg_check(pgmg_rhs_x_dot(hier, level, &v), "source_soln_dot"); return v;}
		/** region::copy_rhs_to_soln (region.cc:1932-1941). */
		void copy_rhs_to_soln() {pgmg_check(pgmg_copy_rhs_to_x(hier, level), "copy_rhs_to_soln");}
		/** region::soln_diff (region.hh:226-236). */
		double soln_diff(double *ext_x) {double v; pgmg_check(pgmg_x_diff2(hier, level, ext_x, &v), "soln_diff"); return v;}
		/** region::overwrite_r_field / update_r_field (region.hh:207-224). */
		template <class p_class>
		void overwrite_r_field(p_class &pr) {
			double *rp = r0;
			for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++)
				*(rp++) = pr.reset_r_field(i + ai, j + aj, k + ak);
			sync_r_to_device();
		}
		template <class p_class>
		void update_r_field(p_class &pr) {
			pgmg_check(pgmg_get_rhs(hier, level, r0), "update_r_field");
			double *rp = r0;
			for (int k = 0; k < so; k++) for (int j = 0; j < sn; j++) for (int i = 0; i < sm; i++)
				*(rp++) += pr.update_r_field(i + ai, j + aj, k + ak);
			sync_r_to_device();
		}
		/** z-slices of x, r and the residual as a binary float matrix
		 * (region.cc:1562-1660): a header row m, 0 .. m-1, then per row j the
		 * value j followed by the m values of the row. */
		void output_x(const char *filename, int k) {
			std::vector<double> f((size_t)smno);
			pgmg_check(pgmg_get_x(hier, level, &f[0]), "output_x");
			output_slice(filename, k, &f[0]);
		}
		void output_r(const char *filename, int k) {
			std::vector<double> f((size_t)smno);
			pgmg_check(pgmg_get_rhs(hier, level, &f[0]), "output_r");
			output_slice(filename, k, &f[0]);
		}
		void output_residual(const char *filename, int k) {
			std::vector<double> f((size_t)smno);
			pgmg_check(pgmg_residual(hier, level, &f[0]), "output_residual");
			output_slice(filename, k, &f[0]);
		}
		inline int total() {return m * n * o;}
		inline int min_size() {return sm < sn ? (sm < so ? sm : so) : (sn < so ? sn : so);}
		/** The level operator as 27-entry rows (the same data as A / Am / S
		 * in the fixed box [-1,1]^3, zero where a box is smaller). */
		void get_operator27(double *A27) {pgmg_check(pgmg_get_operator27(hier, level, A27), "get_operator27");}
	private:
		bool owns;
		bool exact_enabled;
		bool is_bottom() const {
			int nl = 0;
			pgmg_num_levels(hier, &nl);
			return level == nl - 1;
		}
		void output_slice(const char *filename, int k, const double *f) {
			if (k < 0 || k >= so) p_fatal_error("Output slice out of range", 1);
			FILE *outf = fopen(filename, "wb");
			if (!outf) p_fatal_error("Can't open file", PGMG_FILE_ERROR);
			std::vector<float> buf((size_t)m + 1);
			buf[0] = (float)m;
			for (int i = 0; i < m; i++) buf[i + 1] = (float)i;
			if (fwrite(&buf[0], sizeof(float), m + 1, outf) != (size_t)m + 1) p_fatal_error("File output error", 1);
			for (int j = 0; j < n; j++) {
				buf[0] = (float)j;
				for (int i = 0; i < m; i++) buf[i + 1] = (float)f[i + (size_t)m * (j + (size_t)n * k)];
				if (fwrite(&buf[0], sizeof(float), m + 1, outf) != (size_t)m + 1) p_fatal_error("File output error", 1);
			}
			fclose(outf);
		}
		void init_fields() {
			ip = jp = kp = 0; mp = np = op = 1;
			ai = aj = ak = 0; bi = m; bj = n; bk = o;
			sm = m; sn = n; so = o; smn = sm * sn; smno = smn * so;
			mg = sm; ng = sn; og = so; mng = mg * ng; indent = 0;
			x = new double[(size_t)smno];
			x0 = x;
			r0 = new double[(size_t)smno];
			memset(x, 0, sizeof(double) * (size_t)smno);
			memset(r0, 0, sizeof(double) * (size_t)smno);
		}
		region(const region &);
		region &operator=(const region &);
};

#endif
