/* pgmg_capi.h - C ABI of the MI355X (gfx950) multigrid library libpgmg_amd.so.
 *
 * This is the boundary the drop-in C++ headers in include/pgmg/ call, and
 * what any FFI (ctypes, cgo, JNI) binds.  Plain C types only: opaque handles,
 * ints, doubles and host pointers.  Host pointers are borrowed for the call's
 * duration.  Every function returns 0 on success or a PGMG_ERR_* code;
 * pgmg_last_error() gives the message of the calling thread's last failure.
 * The reference signals the same conditions through p_fatal_error ->
 * MPI_Abort (common.cc:52-55) with the codes of common.hh:19-26; here each
 * PGMG_ERR_<X> equals that enum's value plus one.
 *
 * One hierarchy lives on one GPU and owns one HIP stream.  Calls on one
 * handle must come from one thread at a time; distinct handles are
 * independent.
 *
 * Layout of host arrays: point-major, p = i + m*(j + n*k).  Operators carry
 * 27 entries per point, e = (dk+1)*9 + (dj+1)*3 + (di+1): the reference's box
 * order (region.hh:95-101) with the box fixed to [-1,1]^3 and zero entries
 * where a neighbour lies outside the grid on a non-periodic axis (zero
 * Dirichlet data outside).  On a periodic axis every entry is kept, and an
 * offset that leaves the grid couples to the wrapped point; such a grid
 * coarsens as (m+1)>>1 on that axis (region.cc:13-14).  Batches and serving
 * take non-periodic grids only. */
#ifndef PGMG_CAPI_H
#define PGMG_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGMG_CAPI_VERSION 2

/* Status codes: reference enum (common.hh:19-26) + 1, then new ones. */
enum {
	PGMG_OK = 0,
	PGMG_ERR_FILE = 1,
	PGMG_ERR_MEMORY = 2,
	PGMG_ERR_INTERNAL = 3,
	PGMG_ERR_MULTIGRID = 4,
	PGMG_ERR_MATH = 5,
	PGMG_ERR_SETUP = 6,
	PGMG_ERR_HIP = 7,          /* a HIP runtime call failed */
	PGMG_ERR_ARG = 8,          /* invalid argument */
	PGMG_ERR_UNSUPPORTED = 9,  /* e.g. periodic axes in a batch */
	PGMG_ERR_ADMISSION = 10,   /* serving: request cannot be admitted */
	PGMG_ERR_NOT_CONVERGED = 11,
	PGMG_ERR_NAN = 12
};

enum { PGMG_PREC_FP64 = 0, PGMG_PREC_MIXED = 1 };
/* Smoother: EXACT sweeps hyperplanes i+2j+4k in order and reproduces the
 * reference's lexicographic Gauss-Seidel bit for bit; COLOUR8 updates the
 * eight parity classes (i&1)+2(j&1)+4(k&1) in turn. */
enum { PGMG_SMOOTH_EXACT = 0, PGMG_SMOOTH_COLOUR8 = 1 };
/* Bottom solve: GS24 = 24 sweeps (the reference's effective default,
 * region_tr.cc:557); LU = dense LU factor and triangular solves
 * (region_tr.cc:483-555). */
enum { PGMG_BOTTOM_GS24 = 0, PGMG_BOTTOM_LU = 1 };

typedef struct {
	int m, n, o;              /* global grid size, geometry.hh:12-17 */
	int x_prd, y_prd, z_prd;  /* periodic axes (0 or 1; 0 for batches) */
} pgmg_dims;

typedef struct {
	int precision;   /* PGMG_PREC_* */
	int smoother;    /* PGMG_SMOOTH_* */
	int bottom;      /* PGMG_BOTTOM_* */
	int v_down;      /* multigrid.hh:122, default 2 */
	int v_up;        /* multigrid.hh:128, default 2 */
	int v_bottom;    /* multigrid.hh:125, default 20 (fmg only; kept for API parity) */
	int use_graph;   /* capture V-cycles in a hipGraph */
	int max_levels;  /* 0 = the reference's rule; 1 = top level only (a lone
	                    region, as region(p, gm, com) in gs_parallel.cc:25) */
} pgmg_opts;

typedef struct pgmg_hier pgmg_hier;

const char *pgmg_last_error(void);
int pgmg_version(void);
/* Default options: FP64, COLOUR8, GS24, V(2,2), graphs on. */
void pgmg_default_opts(pgmg_opts *opts);
/* Levels of the hierarchy for a grid, without allocating: multigrid.cc:17-85
 * (coarse size (m+2)>>1, or (m+1)>>1 on a periodic axis, region.cc:13-14;
 * stop at m*n*o <= 1024). dims_out holds 3*16 ints. */
int pgmg_hierarchy_shape(const pgmg_dims *dims, int *nlevels, int *dims_out);
/* Device bytes a hierarchy with these options would hold (with the exact
 * smoother, including the plane-skewed copies that pgmg_compute_rats makes
 * of its large levels when they fit). */
int pgmg_hier_footprint(const pgmg_dims *dims, const pgmg_opts *opts, uint64_t *bytes);

/* multigrid<p_class>::multigrid(pr, gm, com), multigrid.hh:282-295, and
 * setup_hierarchy, multigrid.cc:17-85: allocate every level on the device. */
int pgmg_hier_create(const pgmg_dims *dims, const pgmg_opts *opts, pgmg_hier **out);
/* multigrid::~multigrid, multigrid.cc:7-14. */
int pgmg_hier_destroy(pgmg_hier *h);
/* nlevels / mlevels, multigrid.hh:131-135. */
int pgmg_num_levels(const pgmg_hier *h, int *nlevels);
/* reg[level]->m,n,o, region.hh:30-35. */
int pgmg_level_dims(const pgmg_hier *h, int level, int dims[3]);

/* region::setup_matrices, region.hh:95-101 (top level): host 27-entry rows. */
int pgmg_set_operator27(pgmg_hier *h, const double *A27);
/* Device-side generation of a synthetic problem (include/pgmg_problems.h):
 * the top-level operator and the source. */
int pgmg_generate_problem(pgmg_hier *h, int family, uint64_t seed);
/* multigrid::compute_rats, multigrid.cc:138-152 (Galerkin operators on every
 * coarse level, region.cc:629-660), then the bottom LU when opts.bottom is
 * LU (region_tr.cc:472-534), then FP32 copies when precision is MIXED. */
int pgmg_compute_rats(pgmg_hier *h);
/* Reports whether the bottom LU factor is in use (0 = it fell back to GS24
 * by the rule of region_tr.cc:526-533). */
int pgmg_lu_active(const pgmg_hier *h, int *active);
/* region::enable_exact + region::lu_decomposition of the bottom level
 * (region_tr.cc:472-534), called by a user after the operators exist: the
 * dense bottom operator is assembled and factored now.  If the factor passes
 * the test of region_tr.cc:526-533, the bottom solve uses it from then on
 * and pgmg_compute_rats factors again (as compute_rats does while Aexact is
 * set, region.cc:659); otherwise the bottom keeps 24 sweeps (Aexact = NULL).
 * A bottom of pgmg_max_exact (2048) points or more is never factored. */
int pgmg_lu_decomposition(pgmg_hier *h);
/* The factor in use, as dgetrf leaves it (region_tr.cc:518): N*N doubles,
 * column-major, L below the unit diagonal and U on and above it, and the N
 * 1-based pivot rows.  N = the bottom level's point count. */
int pgmg_get_lu_factor(pgmg_hier *h, double *LU, int *ipiv);

/* region::setup_r_field / setup_x_field, region.hh:110-126, and read-back of
 * reg[level]->r0 / x0. */
int pgmg_set_rhs(pgmg_hier *h, int level, const double *b);
int pgmg_get_rhs(pgmg_hier *h, int level, double *b);
int pgmg_set_x(pgmg_hier *h, int level, const double *x);
int pgmg_get_x(pgmg_hier *h, int level, double *x);
/* reg[level]->A expanded to 27-entry rows. */
int pgmg_get_operator27(pgmg_hier *h, int level, double *A27);
/* region::allclear_x_field, region.cc:1672-1675. */
int pgmg_clear_x(pgmg_hier *h, int level);

/* multigrid::gauss_seidel(mode, level), multigrid.cc:180-182 ->
 * region::gauss_seidel, region.cc:355-542, repeated nsweeps times. */
int pgmg_gauss_seidel(pgmg_hier *h, int mode, int level, int nsweeps);
/* multigrid::restriction(level), multigrid.cc:211-224: r at level from the
 * residual at level-1, then x at level cleared. */
int pgmg_restriction(pgmg_hier *h, int level);
/* multigrid::interpolation(level), multigrid.cc:196-206: x at level-1 +=
 * P x at level. */
int pgmg_interpolation(pgmg_hier *h, int level);
/* multigrid::solve_exact(level, smooth), multigrid.hh:265-267. */
int pgmg_solve_exact(pgmg_hier *h, int level, int smooth);
/* multigrid::v_cycle(mode), multigrid.cc:88-106. */
int pgmg_v_cycle(pgmg_hier *h, int mode);
/* The public fields v_down / v_up (multigrid.hh:122-128) changed after
 * construction: later cycles use the new counts. */
int pgmg_set_sweeps(pgmg_hier *h, int v_down, int v_up);
/* multigrid::l2_error_all(level), multigrid.cc:246-248 (sum of squared
 * residuals, no square root) and rhs_l2norm(level), multigrid.cc:259-261. */
int pgmg_l2_error_all(pgmg_hier *h, int level, double *out);
int pgmg_rhs_l2norm(pgmg_hier *h, int level, double *out);
/* region::compute_residual, region.cc:1848-1854: the residual r - A x of a
 * level into a host array (point-major). */
int pgmg_residual(pgmg_hier *h, int level, double *res);
/* Operations on outside arrays, for preconditioned CG (multigrid.hh:222-251).
 * Arrays are point-major, p = i + m*(j + n*k), at the level's size.
 * region::Ax / unpadded_Ax, region.cc:1880-1916: out = A in. */
int pgmg_apply_operator(pgmg_hier *h, int level, const double *in, double *out);
/* region::copy_rhs_to_soln, region.cc:1932-1941: x = r. */
int pgmg_copy_rhs_to_x(pgmg_hier *h, int level);
/* region::source_soln_dot, region.cc:1918-1930: sum of x*r. */
int pgmg_rhs_x_dot(pgmg_hier *h, int level, double *out);
/* region::soln_diff, region.hh:226-236: sum of (ext_x - x)^2. */
int pgmg_x_diff2(pgmg_hier *h, int level, const double *ext_x, double *out);
/* No counterpart in the reference: pgmg_x_diff2 for callers that cannot
 * allocate device memory (code under graph capture, tight serving or PCG
 * loops, harnesses that forbid hipMalloc).  It allocates nothing: the
 * caller's vector, the block sums and the result go through one device
 * array of 8 Mi doubles (64 MiB) that each GPU the library runs on holds
 * until the process ends.  Takes levels with points + block sums + 1 <=
 * 8 Mi doubles (about 8.3 M points; PGMG_ERR_ARG beyond); the result is bit
 * for bit pgmg_x_diff2's.  Calls on one GPU share that array: calls that
 * overlap on one GPU (several host threads) race. */
int pgmg_x_diff2_noalloc(pgmg_hier *h, int level, const double *ext_x, double *out);
/* Solve to a relative residual: V-cycles (FP64) or FP64 defect correction
 * around FP32 V-cycles (MIXED) from the current x until
 * sqrt(l2_error_all/rhs_l2norm) < tol.  hist (max_cycles+1 doubles, may be
 * NULL) receives l2_error_all before and after each cycle.  Returns
 * PGMG_ERR_NOT_CONVERGED if max_cycles are used up, PGMG_ERR_NAN on NaN. */
int pgmg_solve(pgmg_hier *h, double tol, int max_cycles, int *cycles, double *hist);
/* Wait for all work on the hierarchy's stream.  Also reports (and clears)
 * an exact sweep that gave up waiting (see pgmg_set_wait_limit). */
int pgmg_synchronize(pgmg_hier *h);
/* No counterpart in the reference.  Exact-mode sweeps of a level run in one
 * launch in which the workgroup of each k-plane waits for its neighbour
 * planes' progress; each wait gives up after `polls` polls (default 2^21,
 * about a second), marks the sweep failed and lets the launch end.  The
 * failure is returned (PGMG_ERR_HIP) by the next pgmg_synchronize,
 * pgmg_get_x, norm or solve call.  A small limit is a test of that path. */
int pgmg_set_wait_limit(pgmg_hier *h, int polls);
/* region::gs_num, region.hh:17: Gauss-Seidel sweeps performed so far. */
int pgmg_gs_count(const pgmg_hier *h, long long *count);

/* ---------------------------------------------------------------------
 * Batches and serving (no counterpart in the reference, whose ranks solve
 * one problem each).  A batch owns one HBM arena and up to max_slots
 * independent hierarchies of any sizes; pgmg_batch_step advances every
 * active slot by one cycle with one launch per (operation, level) across all
 * of them.  A slot's arithmetic is that of a lone hierarchy with the same
 * options, whatever else shares the batch.  Smoother: COLOUR8 only.
 * --------------------------------------------------------------------- */
typedef struct pgmg_batch pgmg_batch;
enum { PGMG_SLOT_FREE = 0, PGMG_SLOT_ACTIVE = 1, PGMG_SLOT_CONVERGED = 2, PGMG_SLOT_FAILED = 3 };
/* arena_bytes of HBM are reserved at creation. */
int pgmg_batch_create(const pgmg_opts *opts, uint64_t arena_bytes, int max_slots, pgmg_batch **out);
int pgmg_batch_destroy(pgmg_batch *b);
/* Admit a synthetic problem (device-generated) or a host operator + source.
 * Returns PGMG_ERR_ADMISSION when no slot or not enough arena is free.  The
 * solve stops when sqrt(l2_error_all/rhs_l2norm) < tol or after max_cycles. */
int pgmg_batch_admit(pgmg_batch *b, const pgmg_dims *dims, int family, uint64_t seed, double tol, int max_cycles,
		     int *slot);
int pgmg_batch_admit_operator(pgmg_batch *b, const pgmg_dims *dims, const double *A27, const double *rhs, double tol,
			      int max_cycles, int *slot);
/* One cycle for every active slot; reports how many remain active and how
 * many finished in this step.  When no slot is active but set-ups are
 * pending, first waits for the oldest set-up to finish. */
int pgmg_batch_step(pgmg_batch *b, int *active, int *finished);
int pgmg_batch_slot_status(pgmg_batch *b, int slot, int *state, int *cycles, double *rel_residual);
int pgmg_batch_get_x(pgmg_batch *b, int slot, double *x);
/* Return a finished slot's memory to the arena. */
int pgmg_batch_release(pgmg_batch *b, int slot);
int pgmg_batch_free_bytes(const pgmg_batch *b, uint64_t *bytes);
/* Tile shapes of the batched launches: the work units (rows or planes) per
 * workgroup chosen for each launch class, in the order GS pair sweeps,
 * residual, level-0 residual, restriction, interpolation, x += x32.  The
 * batch's first eight steps try 1, 2, 4 and 8 units per workgroup (twice
 * each) with timing events around the launches and keep the fastest per
 * unit; results do not depend on the choice.  *tuned is 1 once chosen.
 * With use_graph, the cycle is then captured as a hipGraph on the second
 * step of an unchanged plan and replayed until the plan changes (a slot
 * admitted or finished); *graph_launches counts the replays. */
enum { PGMG_BATCH_TILE_CLASSES = 6 };
int pgmg_batch_tiles(const pgmg_batch *b, int *chunks, int *tuned, long long *graph_launches);
/* Open-loop serving: requests arrive at the given times (seconds from the
 * call); each is admitted when a slot and arena space are free (FIFO), runs
 * with the others, and completes when its solution is on the host.
 * Results give admission and completion times (seconds from the call).
 * A request's solution is written to its x (m*n*o doubles, point-major, the
 * caller's host memory, p = i + m*(j + n*k)) before its done_s is taken: the
 * x0 a solve_v_cycles call leaves in reg[0] (multigrid.hh:182-219).  With x
 * NULL the solution is still copied off the device (the latency includes it)
 * and then dropped. */
typedef struct {
	double arrival_s;
	int m, n, o;
	int family;
	uint64_t seed;
	double tol;
	int max_cycles;
	double *x;
} pgmg_request;
typedef struct { double admit_s, done_s; int cycles; int status; double rel_residual; } pgmg_result;
int pgmg_serve_run(pgmg_batch *b, const pgmg_request *reqs, int n, pgmg_result *results, double *wall_s);

/* Kernel timing for roofline measurement (no counterpart in the reference):
 * while enabled, every launch of the named kernel class on level `level` is
 * bracketed by HIP events on the hierarchy's stream; pgmg_prof_read returns
 * the launch count, the summed event time in milliseconds and the number of
 * grid points those launches updated.  Class 0 disables timing.  Classes: */
enum { PGMG_KC_GS = 1, PGMG_KC_RESIDUAL = 2, PGMG_KC_RESTRICT = 3, PGMG_KC_INTERP = 4 };
int pgmg_prof_enable(pgmg_hier *h, int kernel_class, int level);
int pgmg_prof_read(pgmg_hier *h, long long *launches, double *total_ms, long long *points);

#ifdef __cplusplus
}
#endif
#endif
