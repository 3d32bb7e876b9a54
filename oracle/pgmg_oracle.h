/* pgmg_oracle.h - CPU restatement of PGMG's single-rank multigrid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU library is
 * compared with; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product never links or calls it.
 *
 * Layout: level arrays are point-major with p = i + m*(j + n*k); operators
 * hold 27 entries per point, e = (dk+1)*9 + (dj+1)*3 + (di+1), zero where the
 * neighbour lies outside the grid on a non-periodic axis (zero Dirichlet).
 * The _p functions take the periodic axes prd (bit 0 x, bit 1 y, bit 2 z).
 * All arithmetic is done in the reference's order without contraction, so
 * results are bit-identical to the reference where it is deterministic
 * (see the per-function notes in pgmg_oracle.c). */
#ifndef PGMG_ORACLE_H
#define PGMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 16

/* Hierarchy shape: level sizes until m*n*o <= 1024. Returns the level count. */
int orc_hierarchy(int m, int n, int o, int *dims /* 3*ORC_MAX_LEVELS */);

/* Synthetic problems (include/pgmg_problems.h). */
void orc_gen_operator(int family, uint64_t seed, int m, int n, int o, double *A27);
void orc_gen_operator_planes(int family, uint64_t seed, int m, int n, int o, int k0, int k1, double *A27);
void orc_gen_rhs(uint64_t seed, int m, int n, int o, double *b);

/* Level kernels, FP64. */
void orc_gs_lex(int m, int n, int o, const double *A, double *x, const double *r, int mode);
void orc_gs_colour8(int m, int n, int o, const double *A, double *x, const double *r, int mode);
void orc_residual(int m, int n, int o, const double *A, const double *x, const double *r, double *res);
double orc_l2_error_all(int m, int n, int o, const double *A, const double *x, const double *r);
double orc_rhs_l2norm(int m, int n, int o, const double *r);
void orc_restrict(int mf, int nf, int of, const double *Af, const double *xf, const double *rf,
		  int mc, int nc, int oc, double *rc);
void orc_interp(int mc, int nc, int oc, const double *xc, int mf, int nf, int of, double *xf);
void orc_rat(int mf, int nf, int of, const double *Af, int mc, int nc, int oc, double *Ac);

/* Level kernels, FP32 (mixed-precision inner cycle). */
void orc_gs_lex_f(int m, int n, int o, const float *A, float *x, const float *r, int mode);
void orc_gs_colour8_f(int m, int n, int o, const float *A, float *x, const float *r, int mode);
void orc_restrict_f(int mf, int nf, int of, const float *Af, const float *xf, const float *rf,
		    int mc, int nc, int oc, float *rc);
void orc_interp_f(int mc, int nc, int oc, const float *xc, int mf, int nf, int of, float *xf);

/* Bottom solve: dense LU with partial pivoting.  orc_lu_factor returns 1 if
 * the factor is usable and 0 if the reference would fall back to 24
 * Gauss-Seidel sweeps (zero pivot or |U_nn|/|U_11| < eps*1e4). */
int orc_lu_factor(int m, int n, int o, const double *A27, double *LU, int *ipiv);
void orc_lu_solve(int N, const double *LU, const int *ipiv, const double *b, double *x);

/* Periodic axes. */
int orc_hierarchy_p(int m, int n, int o, int prd, int *dims);
void orc_gen_operator_p(int family, uint64_t seed, int m, int n, int o, int prd, double *A27);
void orc_gs_lex_p(int m, int n, int o, int prd, const double *A, double *x, const double *r, int mode);
void orc_gs_colour8_p(int m, int n, int o, int prd, const double *A, double *x, const double *r, int mode);
void orc_residual_p(int m, int n, int o, int prd, const double *A, const double *x, const double *r, double *res);
double orc_l2_error_all_p(int m, int n, int o, int prd, const double *A, const double *x, const double *r);
void orc_restrict_p(int mf, int nf, int of, int prd, const double *Af, const double *xf, const double *rf,
		    int mc, int nc, int oc, double *rc);
void orc_interp_p(int mc, int nc, int oc, int prd, const double *xc, int mf, int nf, int of, double *xf);
void orc_rat_p(int mf, int nf, int of, int prd, const double *Af, int mc, int nc, int oc, double *Ac);
int orc_lu_factor_p(int m, int n, int o, int prd, const double *A27, double *LU, int *ipiv);
void orc_gs_lex_p_f(int m, int n, int o, int prd, const float *A, float *x, const float *r, int mode);
void orc_gs_colour8_p_f(int m, int n, int o, int prd, const float *A, float *x, const float *r, int mode);
void orc_restrict_p_f(int mf, int nf, int of, int prd, const float *Af, const float *xf, const float *rf,
		      int mc, int nc, int oc, float *rc);
void orc_interp_p_f(int mc, int nc, int oc, int prd, const float *xc, int mf, int nf, int of, float *xf);
/* Whole hierarchy. */
typedef struct orc_mg orc_mg;
enum { ORC_SMOOTH_LEX = 0, ORC_SMOOTH_COLOUR8 = 1 };
orc_mg *orc_mg_create(int m, int n, int o, const double *A27);
orc_mg *orc_mg_create_p(int m, int n, int o, int prd, const double *A27);
void orc_mg_destroy(orc_mg *mg);
int orc_mg_levels(const orc_mg *mg);
void orc_mg_set_smoother(orc_mg *mg, int smoother);
void orc_mg_set_sweeps(orc_mg *mg, int v_down, int v_up);
int orc_mg_enable_lu(orc_mg *mg);
void orc_mg_set_rhs(orc_mg *mg, const double *b);
void orc_mg_set_x(orc_mg *mg, int level, const double *x);
void orc_mg_get_x(const orc_mg *mg, int level, double *x);
void orc_mg_get_r(const orc_mg *mg, int level, double *r);
void orc_mg_get_A(const orc_mg *mg, int level, double *A27);
void orc_mg_gauss_seidel(orc_mg *mg, int mode, int level);
void orc_mg_restriction(orc_mg *mg, int level);
void orc_mg_interpolation(orc_mg *mg, int level);
void orc_mg_solve_exact(orc_mg *mg, int level);
void orc_mg_v_cycle(orc_mg *mg, int mode);
double orc_mg_l2_error_all(orc_mg *mg, int level);
double orc_mg_rhs_l2norm(orc_mg *mg, int level);
/* V-cycles from the current x until sqrt(l2_error_all/rhs_l2norm) < tol;
 * res_hist (maxc+1 entries, may be NULL) receives l2_error_all before and
 * after each cycle.  Returns the number of cycles. */
int orc_mg_solve(orc_mg *mg, double tol, int maxc, double *res_hist);
/* Mixed precision: FP64 defect correction around an FP32 V-cycle whose
 * operators are the FP64 ones rounded to FP32; bottom solve in FP64. */
int orc_mg_solve_mixed(orc_mg *mg, double tol, int maxc, double *res_hist);
/* The FP32 operator of level l that orc_mg_solve_mixed made (27 entries per
 * point).  Returns 0 if no mixed solve has run yet, else 1. */
int orc_mg_get_Af(const orc_mg *mg, int level, float *A27);

#ifdef __cplusplus
}
#endif
#endif
