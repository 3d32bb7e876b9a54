/* pgmg_oracle.c - CPU restatement of PGMG's single-rank multigrid path.
 *
 * TEST INFRASTRUCTURE ONLY (see pgmg_oracle.h).  Compiled with
 * -ffp-contract=off so that a*b+c is always two roundings, as in the
 * reference's x86-64 build.
 *
 * Citations are file:line in the PGMG reference.  The reference stores a
 * variable box of coefficients per point; here every box is the clipped
 * [-1,1]^3 box with zero entries where the reference has none.  Every coarse
 * box of the reference lies inside [-1,1]^3, and a zero entry only ever adds a
 * signed zero to a sum, so values agree with the reference's (a -0.0 may read
 * as +0.0). */
#include "pgmg_oracle.h"
#include "../include/pgmg_problems.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>

#define IDX(i, j, k, m, n) ((size_t)(i) + (size_t)(m) * ((size_t)(j) + (size_t)(n) * (size_t)(k)))

/* ---------------------------------------------------------------- shape */

/* Coarse size (m+2)>>1 for a non-periodic axis: region.cc:14.  Levels are
 * added while the level has more than 1024 points, at most 16 levels:
 * multigrid.cc:26-29, common.hh:15. */
int orc_hierarchy(int m, int n, int o, int *dims) {
	int l = 0;
	dims[0] = m; dims[1] = n; dims[2] = o;
	while ((long)dims[3 * l] * dims[3 * l + 1] * dims[3 * l + 2] > 1024) {
		if (l + 1 == ORC_MAX_LEVELS) return -1;
		dims[3 * (l + 1)] = (dims[3 * l] + 2) >> 1;
		dims[3 * (l + 1) + 1] = (dims[3 * l + 1] + 2) >> 1;
		dims[3 * (l + 1) + 2] = (dims[3 * l + 2] + 2) >> 1;
		l++;
	}
	return l + 1;
}

void orc_gen_operator(int family, uint64_t seed, int m, int n, int o, double *A27) {
	int i, j, k;
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++)
		pgmg_row27(family, seed, m, n, o, i, j, k, A27 + 27 * IDX(i, j, k, m, n));
}

/* Planes k0 <= k < k1 of the operator of an m x n x o grid, into
 * A27[(k-k0), j, i, 27]: checks large grids a few planes at a time. */
void orc_gen_operator_planes(int family, uint64_t seed, int m, int n, int o, int k0, int k1, double *A27) {
	int i, j, k;
	for (k = k0; k < k1; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++)
		pgmg_row27(family, seed, m, n, o, i, j, k, A27 + 27 * IDX(i, j, k - k0, m, n));
}

void orc_gen_rhs(uint64_t seed, int m, int n, int o, double *b) {
	int i, j, k;
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++)
		b[IDX(i, j, k, m, n)] = pgmg_rhs(seed, m, n, o, i, j, k);
}

/* ------------------------------------------------- generic level kernels */

#define REAL double
#define SFX(name) name
#include "pgmg_oracle_level.inc"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f
#include "pgmg_oracle_level.inc"
#undef REAL
#undef SFX

double orc_l2_error_all(int m, int n, int o, const double *A, const double *x, const double *r) {
	/* region.cc:1475-1490: sum of squared residuals in lexicographic order. */
	double err = 0;
	int i, j, k;
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++) {
		double rem = orc_residual_at(m, n, o, A, x, r, i, j, k);
		err += rem * rem;
	}
	return err;
}

double orc_rhs_l2norm(int m, int n, int o, const double *r) {
	/* region.cc:1549-1558. */
	double s = 0;
	size_t p, N = (size_t)m * n * o;
	for (p = 0; p < N; p++) s += r[p] * r[p];
	return s;
}

void orc_residual(int m, int n, int o, const double *A, const double *x, const double *r, double *res) {
	int i, j, k;
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++)
		res[IDX(i, j, k, m, n)] = orc_residual_at(m, n, o, A, x, r, i, j, k);
}

/* ------------------------------------------------------------------ RAT */

/* grid_map, non-periodic branch: region.cc:1209-1215. */
static void grid_map(int *lo, int *hi, int ma) {
	*lo = (!(ma & 1) && *lo == ma - 1 ? *lo + 1 : *lo) >> 1;
	*hi = (*hi + 2) >> 1;
}

/* igrid_map, non-periodic branch: region.cc:1243-1258.  Returns the fine
 * index of coarse point i and the fine range [lo,hi) of its interpolation
 * support, cropped to [blo,bhi). */
static int igrid_map(int *lo, int *hi, int blo, int bhi, int i, int ma) {
	int j = i << 1;
	if (ma & 1) {
		*lo = j == 0 ? 0 : j - 1;
		*hi = j == ma - 1 ? ma : j + 2;
	} else if (j == ma) {
		*lo = ma - 1; *hi = ma; j = ma - 1;
	} else {
		*lo = j == 0 ? 0 : j - 1;
		*hi = j == ma - 2 ? ma - 1 : j + 2;
	}
	if (*lo < blo) *lo = blo;
	if (*hi > bhi) *hi = bhi;
	return j;
}

/* Galerkin coarse operator Ac = P^T Af P / 8.
 * region.cc:629-660 (compute_rats), 665-746 (scan with full/partial masks),
 * 1008-1099 (rat_fill_full/partial), 1117-1166 (A_red), 1173-1180 (A_add),
 * region.hh:460-465 (A_sca, A_mul).
 * For each fine point f in cube-scan order, A_red forms
 * At[J] = sum_g 0.125*Af(f,g)*P(g,J) over g in f's box (k,j,i order), for J
 * in the coarse box grid_map(f's box); At is then scaled by f's restriction
 * weight and added to every coarse row f restricts to. */
void orc_rat(int mf, int nf, int of, const double *Af, int mc, int nc, int oc, double *Ac) {
	int lastx = (mf - 1) >> 1, lasty = (nf - 1) >> 1, lastz = (of - 1) >> 1;
	int ic, jc, kc, li, lj, lk;
	memset(Ac, 0, sizeof(double) * 27 * (size_t)mc * nc * oc);
	for (kc = 0; kc <= lastz; kc++) for (jc = 0; jc <= lasty; jc++) for (ic = 0; ic <= lastx; ic++)
	for (lk = 0; lk < 2; lk++) for (lj = 0; lj < 2; lj++) for (li = 0; li < 2; li++) {
		int fi = 2 * ic + li, fj = 2 * jc + lj, fk = 2 * kc + lk;
		if (fi >= mf || fj >= nf || fk >= of) continue;
		/* Box of f in fine global coordinates. */
		int b0 = fi > 0 ? fi - 1 : 0, b1 = fi < mf - 1 ? fi + 2 : mf;
		int b2 = fj > 0 ? fj - 1 : 0, b3 = fj < nf - 1 ? fj + 2 : nf;
		int b4 = fk > 0 ? fk - 1 : 0, b5 = fk < of - 1 ? fk + 2 : of;
		int s0 = b0, s1 = b1, s2 = b2, s3 = b3, s4 = b4, s5 = b5;
		grid_map(&s0, &s1, mf); grid_map(&s2, &s3, nf); grid_map(&s4, &s5, of);
		double At[27];
		int vi, vj, vk, ii, jj, kk, nat = 0;
		const double *Arow = Af + 27 * IDX(fi, fj, fk, mf, nf);
		for (vk = s4; vk < s5; vk++) {
			int lok, hik, ck = igrid_map(&lok, &hik, b4, b5, vk, of);
			for (vj = s2; vj < s3; vj++) {
				int loj, hij, cj = igrid_map(&loj, &hij, b2, b3, vj, nf);
				for (vi = s0; vi < s1; vi++) {
					int loi, hii, ci = igrid_map(&loi, &hii, b0, b1, vi, mf);
					double s = 0;
					for (kk = lok; kk < hik; kk++) {
						double kfac = kk == ck ? 1 : 0.5;
						for (jj = loj; jj < hij; jj++) {
							double jkfac = jj == cj ? kfac : 0.5 * kfac;
							for (ii = loi; ii < hii; ii++)
								s += 0.125 * Arow[(kk - fk + 1) * 9 + (jj - fj + 1) * 3 + (ii - fi + 1)]
								     * (ii == ci ? jkfac : 0.5 * jkfac);
						}
					}
					At[nat++] = s;
				}
			}
		}
		/* Restriction weight of f and its coarse targets. */
		int sx = (ic == lastx) && !(mf & 1), sy = (jc == lasty) && !(nf & 1), sz = (kc == lastz) && !(of & 1);
		double w = 1.0;
		if (li && !sx) w *= 0.5;
		if (lj && !sy) w *= 0.5;
		if (lk && !sz) w *= 0.5;
		int tx[2], ty[2], tz[2], nx, ny, nz, a, bb, c;
		if (!li) { tx[0] = ic; nx = 1; } else if (sx) { tx[0] = ic + 1; nx = 1; } else { tx[0] = ic; tx[1] = ic + 1; nx = 2; }
		if (!lj) { ty[0] = jc; ny = 1; } else if (sy) { ty[0] = jc + 1; ny = 1; } else { ty[0] = jc; ty[1] = jc + 1; ny = 2; }
		if (!lk) { tz[0] = kc; nz = 1; } else if (sz) { tz[0] = kc + 1; nz = 1; } else { tz[0] = kc; tz[1] = kc + 1; nz = 2; }
		for (nat = 0; nat < (s1 - s0) * (s3 - s2) * (s5 - s4); nat++) At[nat] *= w;
		for (c = 0; c < nz; c++) for (bb = 0; bb < ny; bb++) for (a = 0; a < nx; a++) {
			int I = tx[a], J = ty[bb], K = tz[c];
			double *Arc = Ac + 27 * IDX(I, J, K, mc, nc);
			nat = 0;
			for (vk = s4; vk < s5; vk++) for (vj = s2; vj < s3; vj++) for (vi = s0; vi < s1; vi++, nat++) {
				int di = vi - I, dj = vj - J, dk = vk - K;
				if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) {
					if (At[nat] != 0) abort(); /* would lie outside [-1,1]^3 */
					continue;
				}
				Arc[(dk + 1) * 9 + (dj + 1) * 3 + (di + 1)] += At[nat];
			}
		}
	}
}

/* ------------------------------------------------------------ bottom LU */

/* Dense assembly as lu_decomposition, region_tr.cc:486-504 (column-major,
 * row = point, column = neighbour), then an unblocked right-looking LU with
 * partial pivoting choosing the first entry of largest modulus, which is
 * dgetrf's specification (region_tr.cc:518).  The usability test follows
 * region_tr.cc:526-533.  LAPACK's blocked dgetrf may round differently;
 * this factor is what the GPU kernel reproduces. */
int orc_lu_factor(int m, int n, int o, const double *A27, double *LU, int *ipiv) {
	int N = m * n * o, i, j, k, c, r, info = 0;
	memset(LU, 0, sizeof(double) * (size_t)N * N);
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++) {
		int p = (int)IDX(i, j, k, m, n), e;
		for (e = 0; e < 27; e++) {
			int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
			int qi = i + di, qj = j + dj, qk = k + dk;
			if (qi < 0 || qi >= m || qj < 0 || qj >= n || qk < 0 || qk >= o) continue;
			LU[p + (size_t)IDX(qi, qj, qk, m, n) * N] = A27[27 * (size_t)p + e];
		}
	}
	for (c = 0; c < N; c++) {
		int piv = c;
		double best = fabs(LU[c + (size_t)c * N]);
		for (r = c + 1; r < N; r++) if (fabs(LU[r + (size_t)c * N]) > best) {best = fabs(LU[r + (size_t)c * N]); piv = r;}
		ipiv[c] = piv;
		if (LU[piv + (size_t)c * N] != 0) {
			if (piv != c) for (j = 0; j < N; j++) {
				double t = LU[c + (size_t)j * N];
				LU[c + (size_t)j * N] = LU[piv + (size_t)j * N];
				LU[piv + (size_t)j * N] = t;
			}
			double rp = 1.0 / LU[c + (size_t)c * N];
			for (r = c + 1; r < N; r++) LU[r + (size_t)c * N] *= rp;
		} else if (info == 0) info = c + 1;
		for (j = c + 1; j < N; j++) {
			double u = LU[c + (size_t)j * N];
			if (u != 0) for (r = c + 1; r < N; r++) LU[r + (size_t)j * N] -= LU[r + (size_t)c * N] * u;
		}
	}
	double eig_hi = fabs(LU[0]), eig_lo = fabs(LU[(size_t)N * N - 1]);
	if (info != 0 || eig_hi == 0 || eig_lo / eig_hi < DBL_EPSILON * 1e4) return 0;
	return 1;
}

/* dgetrs('N') for one right-hand side (region_tr.cc:544-555): row
 * interchanges, then column-oriented unit-lower and upper triangular
 * substitutions. */
void orc_lu_solve(int N, const double *LU, const int *ipiv, const double *b, double *x) {
	int c, r;
	memcpy(x, b, sizeof(double) * N);
	for (c = 0; c < N; c++) if (ipiv[c] != c) {double t = x[c]; x[c] = x[ipiv[c]]; x[ipiv[c]] = t;}
	for (c = 0; c < N; c++) {
		double v = x[c];
		if (v != 0) for (r = c + 1; r < N; r++) x[r] -= v * LU[r + (size_t)c * N];
	}
	for (c = N - 1; c >= 0; c--) {
		double v = x[c] / LU[c + (size_t)c * N];
		x[c] = v;
		if (v != 0) for (r = 0; r < c; r++) x[r] -= v * LU[r + (size_t)c * N];
	}
}

/* ------------------------------------------------------------ hierarchy */

struct orc_mg {
	int nl, smoother, v_down, v_up, lu_ok;
	int dims[3 * ORC_MAX_LEVELS];
	double *A[ORC_MAX_LEVELS], *x[ORC_MAX_LEVELS], *r[ORC_MAX_LEVELS];
	float *Af[ORC_MAX_LEVELS], *xf[ORC_MAX_LEVELS], *rf[ORC_MAX_LEVELS];
	double *LU;
	int *ipiv;
};

#define LM(mg, l) (mg)->dims[3 * (l)]
#define LN(mg, l) (mg)->dims[3 * (l) + 1]
#define LO(mg, l) (mg)->dims[3 * (l) + 2]
#define LPTS(mg, l) ((size_t)LM(mg, l) * LN(mg, l) * LO(mg, l))

/* multigrid ctor + setup_matrices: multigrid.hh:282-305, multigrid.cc:17-85,
 * 138-152.  The bottom LU starts disabled, as it is in every reference solve
 * (its constructor-time factorisation meets uninitialised coefficients and
 * falls back); orc_mg_enable_lu mirrors re-running enable_exact() and
 * lu_decomposition() after the operators exist. */
orc_mg *orc_mg_create(int m, int n, int o, const double *A27) {
	orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
	int l;
	mg->nl = orc_hierarchy(m, n, o, mg->dims);
	mg->v_down = mg->v_up = 2;
	for (l = 0; l < mg->nl; l++) {
		size_t N = LPTS(mg, l);
		mg->A[l] = (double *)malloc(sizeof(double) * 27 * N);
		mg->x[l] = (double *)calloc(N, sizeof(double));
		mg->r[l] = (double *)calloc(N, sizeof(double));
	}
	memcpy(mg->A[0], A27, sizeof(double) * 27 * LPTS(mg, 0));
	for (l = 1; l < mg->nl; l++)
		orc_rat(LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->A[l - 1], LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l]);
	return mg;
}

void orc_mg_destroy(orc_mg *mg) {
	int l;
	for (l = 0; l < mg->nl; l++) {
		free(mg->A[l]); free(mg->x[l]); free(mg->r[l]);
		free(mg->Af[l]); free(mg->xf[l]); free(mg->rf[l]);
	}
	free(mg->LU); free(mg->ipiv);
	free(mg);
}

int orc_mg_levels(const orc_mg *mg) {return mg->nl;}
void orc_mg_set_smoother(orc_mg *mg, int s) {mg->smoother = s;}
void orc_mg_set_sweeps(orc_mg *mg, int d, int u) {mg->v_down = d; mg->v_up = u;}

int orc_mg_enable_lu(orc_mg *mg) {
	int b = mg->nl - 1;
	size_t N = LPTS(mg, b);
	if (N >= 2048) return 0; /* pgmg_max_exact, common.hh:65; region_tr.cc:473 */
	free(mg->LU); free(mg->ipiv);
	mg->LU = (double *)malloc(sizeof(double) * N * N);
	mg->ipiv = (int *)malloc(sizeof(int) * N);
	mg->lu_ok = orc_lu_factor(LM(mg, b), LN(mg, b), LO(mg, b), mg->A[b], mg->LU, mg->ipiv);
	if (!mg->lu_ok) {free(mg->LU); free(mg->ipiv); mg->LU = NULL; mg->ipiv = NULL;}
	return mg->lu_ok;
}

void orc_mg_set_rhs(orc_mg *mg, const double *b) {memcpy(mg->r[0], b, sizeof(double) * LPTS(mg, 0));}
void orc_mg_set_x(orc_mg *mg, int l, const double *x) {memcpy(mg->x[l], x, sizeof(double) * LPTS(mg, l));}
void orc_mg_get_x(const orc_mg *mg, int l, double *x) {memcpy(x, mg->x[l], sizeof(double) * LPTS(mg, l));}
void orc_mg_get_r(const orc_mg *mg, int l, double *r) {memcpy(r, mg->r[l], sizeof(double) * LPTS(mg, l));}
void orc_mg_get_A(const orc_mg *mg, int l, double *A) {memcpy(A, mg->A[l], sizeof(double) * 27 * LPTS(mg, l));}

/* multigrid::gauss_seidel, multigrid.cc:180-182 -> region.cc:355-542. */
void orc_mg_gauss_seidel(orc_mg *mg, int mode, int l) {
	if (mg->smoother == ORC_SMOOTH_COLOUR8)
		orc_gs_colour8(LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l], mg->x[l], mg->r[l], mode);
	else
		orc_gs_lex(LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l], mg->x[l], mg->r[l], mode);
}

/* multigrid::restriction, multigrid.cc:211-224: coarse r from the fine
 * residual, then clear_x_field (region.cc:1679-1682) on the coarse level. */
void orc_mg_restriction(orc_mg *mg, int l) {
	orc_restrict(LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->A[l - 1], mg->x[l - 1], mg->r[l - 1],
		     LM(mg, l), LN(mg, l), LO(mg, l), mg->r[l]);
	memset(mg->x[l], 0, sizeof(double) * LPTS(mg, l));
}

/* multigrid::interpolation, multigrid.cc:196-206. */
void orc_mg_interpolation(orc_mg *mg, int l) {
	orc_interp(LM(mg, l), LN(mg, l), LO(mg, l), mg->x[l], LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->x[l - 1]);
}

/* region::solve_exact, region_tr.cc:540-558: LU solve if enabled, else
 * pgmg_gs_exact=24 sweeps (common.hh:62). */
void orc_mg_solve_exact(orc_mg *mg, int l) {
	int q;
	if (l == mg->nl - 1 && mg->LU != NULL) {
		orc_lu_solve((int)LPTS(mg, l), mg->LU, mg->ipiv, mg->r[l], mg->x[l]);
	} else for (q = 0; q < 24; q++) orc_mg_gauss_seidel(mg, 0, l);
}

/* multigrid::v_cycle, multigrid.cc:88-106: V(v_down+L, v_up+L). */
void orc_mg_v_cycle(orc_mg *mg, int mode) {
	int i, q, L = mg->nl;
	for (i = 0; i < L - 1; i++) {
		for (q = 0; q < mg->v_down + i; q++) orc_mg_gauss_seidel(mg, mode, i);
		orc_mg_restriction(mg, i + 1);
	}
	orc_mg_solve_exact(mg, L - 1);
	for (i = L - 1; i > 0; i--) {
		orc_mg_interpolation(mg, i);
		for (q = 0; q < mg->v_up + i - 1; q++) orc_mg_gauss_seidel(mg, mode, i - 1);
	}
}

double orc_mg_l2_error_all(orc_mg *mg, int l) {
	return orc_l2_error_all(LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l], mg->x[l], mg->r[l]);
}
double orc_mg_rhs_l2norm(orc_mg *mg, int l) {return orc_rhs_l2norm(LM(mg, l), LN(mg, l), LO(mg, l), mg->r[l]);}

int orc_mg_solve(orc_mg *mg, double tol, int maxc, double *hist) {
	double bn = orc_mg_rhs_l2norm(mg, 0), e = orc_mg_l2_error_all(mg, 0);
	int c = 0;
	if (hist) hist[0] = e;
	while (c < maxc && !(sqrt(e / bn) < tol)) {
		orc_mg_v_cycle(mg, 0);
		c++;
		e = orc_mg_l2_error_all(mg, 0);
		if (hist) hist[c] = e;
	}
	return c;
}

/* ----------------------------------------------------- mixed precision */

static void fp32_cycle(orc_mg *mg) {
	int i, q, L = mg->nl;
	for (i = 0; i < L - 1; i++) {
		for (q = 0; q < mg->v_down + i; q++) {
			if (mg->smoother == ORC_SMOOTH_COLOUR8) orc_gs_colour8_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->Af[i], mg->xf[i], mg->rf[i], 0);
			else orc_gs_lex_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->Af[i], mg->xf[i], mg->rf[i], 0);
		}
		orc_restrict_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->Af[i], mg->xf[i], mg->rf[i],
			       LM(mg, i + 1), LN(mg, i + 1), LO(mg, i + 1), mg->rf[i + 1]);
		memset(mg->xf[i + 1], 0, sizeof(float) * LPTS(mg, i + 1));
	}
	{
		int b = L - 1;
		size_t N = LPTS(mg, b), p;
		if (mg->LU != NULL) {
			double *rb = (double *)malloc(sizeof(double) * N), *xb = (double *)malloc(sizeof(double) * N);
			for (p = 0; p < N; p++) rb[p] = (double)mg->rf[b][p];
			orc_lu_solve((int)N, mg->LU, mg->ipiv, rb, xb);
			for (p = 0; p < N; p++) mg->xf[b][p] = (float)xb[p];
			free(rb); free(xb);
		} else for (q = 0; q < 24; q++) {
			if (mg->smoother == ORC_SMOOTH_COLOUR8) orc_gs_colour8_f(LM(mg, b), LN(mg, b), LO(mg, b), mg->Af[b], mg->xf[b], mg->rf[b], 0);
			else orc_gs_lex_f(LM(mg, b), LN(mg, b), LO(mg, b), mg->Af[b], mg->xf[b], mg->rf[b], 0);
		}
	}
	for (i = L - 1; i > 0; i--) {
		orc_interp_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->xf[i], LM(mg, i - 1), LN(mg, i - 1), LO(mg, i - 1), mg->xf[i - 1]);
		for (q = 0; q < mg->v_up + i - 1; q++) {
			if (mg->smoother == ORC_SMOOTH_COLOUR8) orc_gs_colour8_f(LM(mg, i - 1), LN(mg, i - 1), LO(mg, i - 1), mg->Af[i - 1], mg->xf[i - 1], mg->rf[i - 1], 0);
			else orc_gs_lex_f(LM(mg, i - 1), LN(mg, i - 1), LO(mg, i - 1), mg->Af[i - 1], mg->xf[i - 1], mg->rf[i - 1], 0);
		}
	}
}

/* Defect correction: res = b - A x in FP64, r32 = (float)res, e32 from one
 * FP32 V-cycle started at zero, x += (double)e32; stop when
 * sqrt(sum res^2 / sum b^2) < tol.  Operators are the FP64 hierarchy
 * rounded to FP32 (the RAT itself is always FP64). */
int orc_mg_solve_mixed(orc_mg *mg, double tol, int maxc, double *hist) {
	int l, c = 0;
	size_t p, N0 = LPTS(mg, 0);
	for (l = 0; l < mg->nl; l++) {
		size_t N = LPTS(mg, l);
		if (!mg->Af[l]) {
			mg->Af[l] = (float *)malloc(sizeof(float) * 27 * N);
			mg->xf[l] = (float *)calloc(N, sizeof(float));
			mg->rf[l] = (float *)calloc(N, sizeof(float));
		}
		for (p = 0; p < 27 * N; p++) mg->Af[l][p] = (float)mg->A[l][p];
	}
	double bn = orc_mg_rhs_l2norm(mg, 0);
	double *res = (double *)malloc(sizeof(double) * N0);
	for (;;) {
		orc_residual(LM(mg, 0), LN(mg, 0), LO(mg, 0), mg->A[0], mg->x[0], mg->r[0], res);
		double e = 0;
		for (p = 0; p < N0; p++) e += res[p] * res[p];
		if (hist) hist[c] = e;
		if (c >= maxc || sqrt(e / bn) < tol) break;
		for (p = 0; p < N0; p++) {mg->rf[0][p] = (float)res[p]; mg->xf[0][p] = 0.0f;}
		fp32_cycle(mg);
		for (p = 0; p < N0; p++) mg->x[0][p] += (double)mg->xf[0][p];
		c++;
	}
	free(res);
	return c;
}
