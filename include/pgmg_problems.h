/* pgmg_problems.h - deterministic synthetic problem families.
 *
 * Every coefficient and source value is a pure function of (seed, field,
 * linear index), computed with integer hashing and IEEE +,-,*,/ and sqrt only,
 * in a fixed operation order.  Host C, host C++ and gfx950 device code that
 * include this header and compile it without floating-point contraction
 * produce bit-identical operators, independent of traversal order, so a
 * problem can be generated on the device for large grids and on the host for
 * checking.
 *
 * Operators are rows of the 27-entry stencil format: entry
 * e = (dk+1)*9 + (dj+1)*3 + (di+1) couples point (i,j,k) to (i+di,j+dj,k+dk).
 * This is the box order of the variable-stencil format in region.hh:95-101
 * (k outer, i inner), with the box fixed to [-1,1]^3 and clipped entries set
 * to zero.  An axis is either non-periodic, with zero Dirichlet data
 * outside the grid (clipped entries), or periodic (bit 0 of prd for x, bit 1
 * for y, bit 2 for z), where the row keeps every entry and an offset that
 * leaves the grid couples to the point wrapped around the axis.
 *
 * Families:
 *  PGMG_FAM_FACE7_UNIFORM  -div(a grad u), 7-point, per-face a ~ U[1,10]
 *  PGMG_FAM_FACE7_LOGUNI   -div(a grad u), 7-point, per-face a = 10^U[-2,2]
 *  PGMG_FAM_Q1_ANISO       Q1 trilinear finite elements, per-cell tensor
 *                          K = R diag(1,10^u1,10^u2) R^T, u1,u2 ~ U[-2,0],
 *                          R a uniformly random rotation; 27-point rows
 * The source is b ~ U[-1,1] per point for every family. */
#ifndef PGMG_PROBLEMS_H
#define PGMG_PROBLEMS_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define PGMG_HD __host__ __device__
#else
#define PGMG_HD
#endif

#if defined(__clang__)
#define PGMG_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define PGMG_NO_CONTRACT
#endif

enum {
	PGMG_FAM_FACE7_UNIFORM = 1,
	PGMG_FAM_FACE7_LOGUNI = 2,
	PGMG_FAM_Q1_ANISO = 3
};

/* Random streams (the "field" argument of pgmg_u01). */
enum {
	PGMG_STREAM_AX = 1, PGMG_STREAM_AY = 2, PGMG_STREAM_AZ = 3,
	PGMG_STREAM_RHS = 4, PGMG_STREAM_U1 = 5, PGMG_STREAM_U2 = 6,
	PGMG_STREAM_QUAT = 8
};

/* splitmix64 finaliser. */
PGMG_HD static inline uint64_t pgmg_mix64(uint64_t z) {
	z += 0x9e3779b97f4a7c15ull;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

/* Uniform double in [0,1) with 53 random bits, keyed by (seed,stream,index). */
PGMG_HD static inline double pgmg_u01(uint64_t seed, uint64_t stream, uint64_t idx) {
	uint64_t h = pgmg_mix64(pgmg_mix64(seed ^ (stream << 56) ^ (stream * 0x632be59bd9b4e019ull)) + idx);
	return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

/* 10^u for |u| <= 8, from IEEE +,-,* only: 10^u = 2^n * exp(f ln 2) with
 * t = u log2(10), n = round(t), |f| <= 1/2, and a degree-16 Taylor series in
 * Horner form.  Its accuracy is irrelevant (it only defines a problem); it is
 * written this way so that every compiler evaluates it identically. */
PGMG_HD static inline double pgmg_exp10_det(double u) {
	PGMG_NO_CONTRACT
	const double log2_10 = 3.321928094887362347870319429489;
	const double ln2 = 0.693147180559945309417232121458;
	double t = u * log2_10;
	int n = (int)(t >= 0 ? t + 0.5 : t - 0.5);
	double f = (t - (double)n) * ln2;
	double s = 1.0;
	int q;
	for (q = 16; q >= 1; q--) s = 1.0 + s * (f / (double)q);
	if (n > 0) { for (q = 0; q < n; q++) s = s * 2.0; }
	else { for (q = 0; q < -n; q++) s = s * 0.5; }
	return s;
}

/* Source term b(i,j,k) ~ U[-1,1]. */
PGMG_HD static inline double pgmg_rhs(uint64_t seed, int m, int n, int o, int i, int j, int k) {
	PGMG_NO_CONTRACT
	(void)o;
	uint64_t idx = (uint64_t)i + (uint64_t)m * ((uint64_t)j + (uint64_t)n * (uint64_t)k);
	return 2.0 * pgmg_u01(seed, PGMG_STREAM_RHS, idx) - 1.0;
}

/* Face coefficient of the 7-point families.  Axis 0: x-face fi in [0,m]
 * between points fi-1 and fi; axis 1 and 2 likewise. */
PGMG_HD static inline double pgmg_face_coeff(int family, uint64_t seed, int m, int n, int o,
		int axis, int fi, int fj, int fk) {
	PGMG_NO_CONTRACT
	uint64_t idx, stream;
	double u;
	(void)o;
	if (axis == 0) { idx = (uint64_t)fi + (uint64_t)(m + 1) * ((uint64_t)fj + (uint64_t)n * (uint64_t)fk); stream = PGMG_STREAM_AX; }
	else if (axis == 1) { idx = (uint64_t)fi + (uint64_t)m * ((uint64_t)fj + (uint64_t)(n + 1) * (uint64_t)fk); stream = PGMG_STREAM_AY; }
	else { idx = (uint64_t)fi + (uint64_t)m * ((uint64_t)fj + (uint64_t)n * (uint64_t)fk); stream = PGMG_STREAM_AZ; }
	u = pgmg_u01(seed, stream, idx);
	if (family == PGMG_FAM_FACE7_LOGUNI) return pgmg_exp10_det(4.0 * u - 2.0);
	return 1.0 + 9.0 * u;
}

/* Conductivity tensor of cell (ci,cj,ck), ci in [0,m], as the 6 entries
 * K00,K11,K22,K01,K02,K12 of R diag(1,10^u1,10^u2) R^T. */
PGMG_HD static inline void pgmg_cell_tensor(uint64_t seed, int m, int n, int ci, int cj, int ck, double K[6]) {
	PGMG_NO_CONTRACT
	uint64_t idx = (uint64_t)ci + (uint64_t)(m + 1) * ((uint64_t)cj + (uint64_t)(n + 1) * (uint64_t)ck);
	double l1 = pgmg_exp10_det(-2.0 * pgmg_u01(seed, PGMG_STREAM_U1, idx));
	double l2 = pgmg_exp10_det(-2.0 * pgmg_u01(seed, PGMG_STREAM_U2, idx));
	/* Uniform unit quaternion by rejection from [-1,1]^4; tries are keyed
	 * by (index, attempt) so the draw is order independent. */
	double a = 1, b = 0, c = 0, d = 0, q2 = 1;
	int t;
	for (t = 0; t < 64; t++) {
		uint64_t base = idx * 256u + (uint64_t)t * 4u;
		double ta = 2.0 * pgmg_u01(seed, PGMG_STREAM_QUAT, base) - 1.0;
		double tb = 2.0 * pgmg_u01(seed, PGMG_STREAM_QUAT, base + 1) - 1.0;
		double tc = 2.0 * pgmg_u01(seed, PGMG_STREAM_QUAT, base + 2) - 1.0;
		double td = 2.0 * pgmg_u01(seed, PGMG_STREAM_QUAT, base + 3) - 1.0;
		double s = ((ta * ta + tb * tb) + tc * tc) + td * td;
		if (s > 1e-6 && s <= 1.0) { a = ta; b = tb; c = tc; d = td; q2 = s; break; }
	}
	/* Rotation matrix of the (unnormalised) quaternion, divided by |q|^2. */
	double R[3][3];
	R[0][0] = (((a * a + b * b) - c * c) - d * d) / q2;
	R[0][1] = 2.0 * (b * c - a * d) / q2;
	R[0][2] = 2.0 * (b * d + a * c) / q2;
	R[1][0] = 2.0 * (b * c + a * d) / q2;
	R[1][1] = (((a * a - b * b) + c * c) - d * d) / q2;
	R[1][2] = 2.0 * (c * d - a * b) / q2;
	R[2][0] = 2.0 * (b * d - a * c) / q2;
	R[2][1] = 2.0 * (c * d + a * b) / q2;
	R[2][2] = (((a * a - b * b) - c * c) + d * d) / q2;
	double lam[3];
	lam[0] = 1.0; lam[1] = l1; lam[2] = l2;
	int p, q, r;
	double Kf[3][3];
	for (p = 0; p < 3; p++) for (q = 0; q < 3; q++) {
		double s = 0.0;
		for (r = 0; r < 3; r++) s = s + (R[p][r] * lam[r]) * R[q][r];
		Kf[p][q] = s;
	}
	/* Symmetrise exactly: use the upper triangle. */
	K[0] = Kf[0][0]; K[1] = Kf[1][1]; K[2] = Kf[2][2];
	K[3] = Kf[0][1]; K[4] = Kf[0][2]; K[5] = Kf[1][2];
}

/* Entry (a,b) of the Q1 element stiffness matrix on the unit cube for a
 * constant tensor K: sum_{p,q} K_pq int d_p phi_a d_q phi_b, with local
 * nodes a,b in {0..7} (bit 0 = x, bit 1 = y, bit 2 = z).  The 1-D factors
 * are int psi psi = 1/3 or 1/6, int psi' psi' = +1 or -1 and
 * int psi'_a psi_b = (2a-1)/2. */
PGMG_HD static inline double pgmg_q1_entry(const double K[6], int a, int b) {
	PGMG_NO_CONTRACT
	double Kf[3][3];
	Kf[0][0] = K[0]; Kf[1][1] = K[1]; Kf[2][2] = K[2];
	Kf[0][1] = Kf[1][0] = K[3]; Kf[0][2] = Kf[2][0] = K[4]; Kf[1][2] = Kf[2][1] = K[5];
	double s = 0.0;
	int p, q, d;
	for (p = 0; p < 3; p++) for (q = 0; q < 3; q++) {
		double prod = 1.0;
		for (d = 0; d < 3; d++) {
			int ad = (a >> d) & 1, bd = (b >> d) & 1;
			double f;
			if (d == p && d == q) f = (ad == bd) ? 1.0 : -1.0;
			else if (d == p) f = ad ? 0.5 : -0.5;
			else if (d == q) f = bd ? 0.5 : -0.5;
			else f = (ad == bd) ? (1.0 / 3.0) : (1.0 / 6.0);
			prod = prod * f;
		}
		s = s + Kf[p][q] * prod;
	}
	return s;
}

/* Row (i,j,k) of the operator, in 27-entry order, on a grid whose periodic
 * axes are the set bits of prd.  On a non-periodic axis entries that leave
 * the grid are zero.  On a periodic axis the face (7-point) or cell (Q1)
 * index 0 is the same face or cell as index m, so the coefficients wrap
 * consistently and the operator stays symmetric. */
PGMG_HD static inline void pgmg_row27_prd(int family, uint64_t seed, int m, int n, int o, int prd,
		int i, int j, int k, double out[27]) {
	PGMG_NO_CONTRACT
	const int px = prd & 1, py = (prd >> 1) & 1, pz = (prd >> 2) & 1;
	int e;
	for (e = 0; e < 27; e++) out[e] = 0.0;
	if (family == PGMG_FAM_Q1_ANISO) {
		/* Cell c in [0,m] spans points c-1 and c; point i lies in cells i
		 * (as local 1) and i+1 (as local 0).  Cells are visited in
		 * (k,j,i) order. */
		int dc0, dc1, dc2;
		for (dc2 = 0; dc2 < 2; dc2++) for (dc1 = 0; dc1 < 2; dc1++) for (dc0 = 0; dc0 < 2; dc0++) {
			int ci = i + dc0, cj = j + dc1, ck = k + dc2;
			double K[6];
			pgmg_cell_tensor(seed, m, n, (px && ci == 0) ? m : ci, (py && cj == 0) ? n : cj,
					 (pz && ck == 0) ? o : ck, K);
			int la = (1 - dc0) | ((1 - dc1) << 1) | ((1 - dc2) << 2);
			int lb;
			for (lb = 0; lb < 8; lb++) {
				int bi = ci - 1 + (lb & 1), bj = cj - 1 + ((lb >> 1) & 1), bk = ck - 1 + ((lb >> 2) & 1);
				if ((!px && (bi < 0 || bi >= m)) || (!py && (bj < 0 || bj >= n)) || (!pz && (bk < 0 || bk >= o))) continue;
				int idx = (bk - k + 1) * 9 + (bj - j + 1) * 3 + (bi - i + 1);
				out[idx] = out[idx] + pgmg_q1_entry(K, la, lb);
			}
		}
		return;
	}
	double axl = pgmg_face_coeff(family, seed, m, n, o, 0, (px && i == 0) ? m : i, j, k);
	double axh = pgmg_face_coeff(family, seed, m, n, o, 0, i + 1, j, k);
	double ayl = pgmg_face_coeff(family, seed, m, n, o, 1, i, (py && j == 0) ? n : j, k);
	double ayh = pgmg_face_coeff(family, seed, m, n, o, 1, i, j + 1, k);
	double azl = pgmg_face_coeff(family, seed, m, n, o, 2, i, j, (pz && k == 0) ? o : k);
	double azh = pgmg_face_coeff(family, seed, m, n, o, 2, i, j, k + 1);
	out[13] = ((((axl + axh) + ayl) + ayh) + azl) + azh;
	if (px || i > 0) out[12] = -axl;
	if (px || i < m - 1) out[14] = -axh;
	if (py || j > 0) out[10] = -ayl;
	if (py || j < n - 1) out[16] = -ayh;
	if (pz || k > 0) out[4] = -azl;
	if (pz || k < o - 1) out[22] = -azh;
}

/* Row (i,j,k) of the operator on a non-periodic grid. */
PGMG_HD static inline void pgmg_row27(int family, uint64_t seed, int m, int n, int o,
		int i, int j, int k, double out[27]) {
	pgmg_row27_prd(family, seed, m, n, o, 0, i, j, k, out);
}

#endif
