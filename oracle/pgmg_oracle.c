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

/* Coarse size on a periodic axis: (m+1)>>1 (region.cc:14). */
int orc_hierarchy_p(int m, int n, int o, int prd, int *dims) {
	int l = 0, a;
	dims[0] = m; dims[1] = n; dims[2] = o;
	while ((long)dims[3 * l] * dims[3 * l + 1] * dims[3 * l + 2] > 1024) {
		if (l + 1 == ORC_MAX_LEVELS) return -1;
		for (a = 0; a < 3; a++) dims[3 * (l + 1) + a] = (dims[3 * l + a] + (((prd >> a) & 1) ? 1 : 2)) >> 1;
		l++;
	}
	return l + 1;
}

void orc_gen_operator_p(int family, uint64_t seed, int m, int n, int o, int prd, double *A27) {
	int i, j, k;
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++)
		pgmg_row27_prd(family, seed, m, n, o, prd, i, j, k, A27 + 27 * IDX(i, j, k, m, n));
}

/* Neighbour (i,j,k) of a grid with periodic axes prd: -1 if it lies outside
 * a non-periodic axis, else its linear index in *q and 1 if it was reached
 * across a periodic boundary (a ghost value), 0 if not. */
static int orc_nb(int m, int n, int o, int prd, int i, int j, int k, size_t *q) {
	int g = 0;
	if (i < 0 || i >= m) { if (!(prd & 1)) return -1; i = (i + m) % m; g = 1; }
	if (j < 0 || j >= n) { if (!(prd & 2)) return -1; j = (j + n) % n; g = 1; }
	if (k < 0 || k >= o) { if (!(prd & 4)) return -1; k = (k + o) % o; g = 1; }
	*q = IDX(i, j, k, m, n);
	return g;
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

double orc_l2_error_all_p(int m, int n, int o, int prd, const double *A, const double *x, const double *r) {
	double err = 0;
	int i, j, k;
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++) {
		double rem = orc_residual_at_p(m, n, o, prd, A, x, r, i, j, k);
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

/* One axis of the interpolation P on the unwrapped index line.  A periodic
 * axis repeats with period M (fine) and Mc (coarse): coarse J sits on fine
 * 2(J mod Mc) + M floor(J/Mc) (for odd M the last coarse point, fine M-1,
 * and coarse 0 of the next period, fine M, are neighbours; igrid_map,
 * region.cc:1233-1242).  On a non-periodic axis coarse J sits on fine 2J,
 * except the last coarse point of an even M, on fine M-1 (region.cc:1243-1255).
 * A fine point on a coarse position takes that coarse value alone; any
 * other fine point half of each coarse neighbour. */
static int p_fpos(int J, int M, int Mc, int prd) {
	if (prd) {
		int q = J >= 0 ? J / Mc : -((-J + Mc - 1) / Mc);
		return 2 * (J - q * Mc) + M * q;
	}
	return (!(M & 1) && J == Mc - 1) ? M - 1 : 2 * J;
}
static double p_weight(int g, int J, int M, int Mc, int prd) {
	int c, f = p_fpos(J, M, Mc, prd);
	if (prd) c = ((g % M + M) % M & 1) == 0;
	else c = (g & 1) == 0 || (!(M & 1) && g == M - 1);
	if (c) return g == f ? 1.0 : 0.0;
	return (g - f == 1 || f - g == 1) ? 0.5 : 0.0;
}

/* Galerkin coarse operator Ac = P^T Af P / 8 on a grid with periodic axes,
 * on the unwrapped index line of every periodic axis, where each coarse box
 * stays within [-1,1]^3 (the reference's periodic grid_map/igrid_map,
 * region.cc:1201-1242).  Entry d of coarse row I is
 *   sum over fine f with P(f,I) != 0, in (k,j,i) order, of
 *   P(f,I) * [sum over f's box g in (k,j,i) order of 0.125*Af(f,g)*P(g,I+d)],
 * the bracket formed as A_red forms At (region.cc:1117-1166) and scaled as
 * A_mul does (region.hh:460-465).  The reference adds the fine points in
 * cube-scan order and the wrapped ones last, through its RAT strips, so its
 * coarse entries agree with these to rounding, not bitwise. */
void orc_rat_p(int mf, int nf, int of, int prd, const double *Af, int mc, int nc, int oc, double *Ac) {
	const int P[3] = {prd & 1, (prd >> 1) & 1, (prd >> 2) & 1};
	const int Mf[3] = {mf, nf, of}, Mc[3] = {mc, nc, oc};
	int I[3], d, a;
	for (I[2] = 0; I[2] < oc; I[2]++) for (I[1] = 0; I[1] < nc; I[1]++) for (I[0] = 0; I[0] < mc; I[0]++)
	for (d = 0; d < 27; d++) {
		int J[3] = {I[0] + d % 3 - 1, I[1] + (d / 3) % 3 - 1, I[2] + d / 9 - 1}, ok = 1, fc[3];
		double acc = 0;
		for (a = 0; a < 3; a++) {
			if (!P[a] && (J[a] < 0 || J[a] >= Mc[a])) ok = 0;
			fc[a] = p_fpos(I[a], Mf[a], Mc[a], P[a]);
		}
		if (ok) {
			int f[3];
			for (f[2] = fc[2] - 1; f[2] <= fc[2] + 1; f[2]++)
			for (f[1] = fc[1] - 1; f[1] <= fc[1] + 1; f[1]++)
			for (f[0] = fc[0] - 1; f[0] <= fc[0] + 1; f[0]++) {
				double w = 1;
				int fw[3], e;
				for (a = 0; a < 3; a++) {
					if (!P[a] && (f[a] < 0 || f[a] >= Mf[a])) w = 0;
					else w *= p_weight(f[a], I[a], Mf[a], Mc[a], P[a]);
					fw[a] = ((f[a] % Mf[a]) + Mf[a]) % Mf[a];
				}
				if (w == 0) continue;
				const double *Arow = Af + 27 * IDX(fw[0], fw[1], fw[2], mf, nf);
				double sum = 0;
				for (e = 0; e < 27; e++) {
					int g[3] = {f[0] + e % 3 - 1, f[1] + (e / 3) % 3 - 1, f[2] + e / 9 - 1};
					double pw = 1;
					for (a = 0; a < 3; a++) {
						if (!P[a] && (g[a] < 0 || g[a] >= Mf[a])) pw = 0;
						else pw *= p_weight(g[a], J[a], Mf[a], Mc[a], P[a]);
					}
					if (pw != 0) sum += 0.125 * Arow[e] * pw;
				}
				acc += sum * w;
			}
		}
		Ac[27 * IDX(I[0], I[1], I[2], mc, nc) + d] = acc;
	}
}

/* ------------------------------------------------------------ bottom LU */

/* Dense assembly as lu_decomposition, region_tr.cc:486-504 (column-major,
 * row = point, column = neighbour), then an unblocked right-looking LU with
 * partial pivoting choosing the first entry of largest modulus, which is
 * dgetrf's specification (region_tr.cc:518).  The usability test follows
 * region_tr.cc:526-533.  LAPACK's blocked dgetrf may round differently;
 * this factor is what the GPU kernel reproduces. */
static int lu_factor_dense(int N, double *LU, int *ipiv) {
	int c, r, j, info = 0;
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

int orc_lu_factor(int m, int n, int o, const double *A27, double *LU, int *ipiv) {
	return orc_lu_factor_p(m, n, o, 0, A27, LU, ipiv);
}

/* With periodic axes the column of a wrapped neighbour is its wrapped
 * index (step_mod, region.hh:466; region_tr.cc:493-501).  Entries are
 * stored, not added, in box order, so on an axis of size 1 or 2, where two
 * offsets reach one column, the later one is kept, as in the reference. */
int orc_lu_factor_p(int m, int n, int o, int prd, const double *A27, double *LU, int *ipiv) {
	int N = m * n * o, i, j, k;
	memset(LU, 0, sizeof(double) * (size_t)N * N);
	for (k = 0; k < o; k++) for (j = 0; j < n; j++) for (i = 0; i < m; i++) {
		int p = (int)IDX(i, j, k, m, n), e;
		for (e = 0; e < 27; e++) {
			size_t q;
			if (orc_nb(m, n, o, prd, i + e % 3 - 1, j + (e / 3) % 3 - 1, k + e / 9 - 1, &q) < 0) continue;
			LU[p + q * N] = A27[27 * (size_t)p + e];
		}
	}
	return lu_factor_dense(N, LU, ipiv);
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
	int nl, smoother, v_down, v_up, lu_ok, prd;
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
	return orc_mg_create_p(m, n, o, 0, A27);
}

/* The same with periodic axes prd (bit 0 x, bit 1 y, bit 2 z). */
orc_mg *orc_mg_create_p(int m, int n, int o, int prd, const double *A27) {
	orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
	int l;
	mg->prd = prd;
	mg->nl = orc_hierarchy_p(m, n, o, prd, mg->dims);
	mg->v_down = mg->v_up = 2;
	for (l = 0; l < mg->nl; l++) {
		size_t N = LPTS(mg, l);
		mg->A[l] = (double *)malloc(sizeof(double) * 27 * N);
		mg->x[l] = (double *)calloc(N, sizeof(double));
		mg->r[l] = (double *)calloc(N, sizeof(double));
	}
	memcpy(mg->A[0], A27, sizeof(double) * 27 * LPTS(mg, 0));
	for (l = 1; l < mg->nl; l++) {
		if (prd) orc_rat_p(LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), prd, mg->A[l - 1], LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l]);
		else orc_rat(LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->A[l - 1], LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l]);
	}
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
	mg->lu_ok = orc_lu_factor_p(LM(mg, b), LN(mg, b), LO(mg, b), mg->prd, mg->A[b], mg->LU, mg->ipiv);
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
	if (mg->prd) {
		if (mg->smoother == ORC_SMOOTH_COLOUR8)
			orc_gs_colour8_p(LM(mg, l), LN(mg, l), LO(mg, l), mg->prd, mg->A[l], mg->x[l], mg->r[l], mode);
		else
			orc_gs_lex_p(LM(mg, l), LN(mg, l), LO(mg, l), mg->prd, mg->A[l], mg->x[l], mg->r[l], mode);
	} else if (mg->smoother == ORC_SMOOTH_COLOUR8)
		orc_gs_colour8(LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l], mg->x[l], mg->r[l], mode);
	else
		orc_gs_lex(LM(mg, l), LN(mg, l), LO(mg, l), mg->A[l], mg->x[l], mg->r[l], mode);
}

/* multigrid::restriction, multigrid.cc:211-224: coarse r from the fine
 * residual, then clear_x_field (region.cc:1679-1682) on the coarse level. */
void orc_mg_restriction(orc_mg *mg, int l) {
	if (mg->prd) orc_restrict_p(LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->prd, mg->A[l - 1], mg->x[l - 1], mg->r[l - 1],
				    LM(mg, l), LN(mg, l), LO(mg, l), mg->r[l]);
	else orc_restrict(LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->A[l - 1], mg->x[l - 1], mg->r[l - 1],
		     LM(mg, l), LN(mg, l), LO(mg, l), mg->r[l]);
	memset(mg->x[l], 0, sizeof(double) * LPTS(mg, l));
}

/* multigrid::interpolation, multigrid.cc:196-206. */
void orc_mg_interpolation(orc_mg *mg, int l) {
	if (mg->prd) orc_interp_p(LM(mg, l), LN(mg, l), LO(mg, l), mg->prd, mg->x[l], LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->x[l - 1]);
	else orc_interp(LM(mg, l), LN(mg, l), LO(mg, l), mg->x[l], LM(mg, l - 1), LN(mg, l - 1), LO(mg, l - 1), mg->x[l - 1]);
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
	if (mg->prd) return orc_l2_error_all_p(LM(mg, l), LN(mg, l), LO(mg, l), mg->prd, mg->A[l], mg->x[l], mg->r[l]);
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

static void f32_gs(orc_mg *mg, int l) {
	if (mg->prd) {
		if (mg->smoother == ORC_SMOOTH_COLOUR8) orc_gs_colour8_p_f(LM(mg, l), LN(mg, l), LO(mg, l), mg->prd, mg->Af[l], mg->xf[l], mg->rf[l], 0);
		else orc_gs_lex_p_f(LM(mg, l), LN(mg, l), LO(mg, l), mg->prd, mg->Af[l], mg->xf[l], mg->rf[l], 0);
	} else if (mg->smoother == ORC_SMOOTH_COLOUR8) orc_gs_colour8_f(LM(mg, l), LN(mg, l), LO(mg, l), mg->Af[l], mg->xf[l], mg->rf[l], 0);
	else orc_gs_lex_f(LM(mg, l), LN(mg, l), LO(mg, l), mg->Af[l], mg->xf[l], mg->rf[l], 0);
}

static void fp32_cycle(orc_mg *mg) {
	int i, q, L = mg->nl;
	for (i = 0; i < L - 1; i++) {
		for (q = 0; q < mg->v_down + i; q++) f32_gs(mg, i);
		if (mg->prd) orc_restrict_p_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->prd, mg->Af[i], mg->xf[i], mg->rf[i],
					      LM(mg, i + 1), LN(mg, i + 1), LO(mg, i + 1), mg->rf[i + 1]);
		else orc_restrict_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->Af[i], mg->xf[i], mg->rf[i],
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
		} else for (q = 0; q < 24; q++) f32_gs(mg, b);
	}
	for (i = L - 1; i > 0; i--) {
		if (mg->prd) orc_interp_p_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->prd, mg->xf[i], LM(mg, i - 1), LN(mg, i - 1), LO(mg, i - 1), mg->xf[i - 1]);
		else orc_interp_f(LM(mg, i), LN(mg, i), LO(mg, i), mg->xf[i], LM(mg, i - 1), LN(mg, i - 1), LO(mg, i - 1), mg->xf[i - 1]);
		for (q = 0; q < mg->v_up + i - 1; q++) f32_gs(mg, i - 1);
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
		if (mg->prd) orc_residual_p(LM(mg, 0), LN(mg, 0), LO(mg, 0), mg->prd, mg->A[0], mg->x[0], mg->r[0], res);
		else orc_residual(LM(mg, 0), LN(mg, 0), LO(mg, 0), mg->A[0], mg->x[0], mg->r[0], res);
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

int orc_mg_get_Af(const orc_mg *mg, int l, float *A27) {
	if (!mg->Af[l]) return 0;
	memcpy(A27, mg->Af[l], sizeof(float) * 27 * LPTS(mg, l));
	return 1;
}
