// pgmg_hier.hip - host side of libpgmg_amd: the level hierarchy resident in
// HBM, the V-cycle schedule, the solve drivers and the C ABI of
// include/pgmg_capi.h.  No kernels here; see pgmg_kernels.hip.
#include "pgmg_internal.hpp"
#include "../../include/pgmg_capi.h"
#include "../../include/pgmg_problems.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace pgmg;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_err = buf;
	return code;
}

#define HIP_TRY(call)                                                                        \
	do {                                                                                 \
		hipError_t e_ = (call);                                                      \
		if (e_ != hipSuccess) return fail(PGMG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
	} while (0)

const int kMaxLevels = 16;        // mg_max_levels, common.hh:15
const int kGsExact = 24;          // pgmg_gs_exact, common.hh:62
const int kMaxExact = 2048;       // pgmg_max_exact, common.hh:65
// Levels with at most this many points are smoothed by one workgroup per
// call (all sweeps in one launch); larger levels launch per colour/plane.
// Exact mode keeps the one workgroup only for levels whose x fits in LDS;
// larger ones take the plane-parallel sweeps (k_gs_flow).
const long long kSmallColour = 20000;
// Largest FP64 skewed copy (27 coefficient planes, x and r) a level gets for
// the exact smoother; larger levels sweep the level layout.
const unsigned long long kSkewMaxBytes = 64ULL << 30;
// Polls a plane of the exact sweeps makes before it gives up waiting for a
// neighbour plane (a poll takes about a microsecond; a wait in a working
// sweep lasts a few).
const int kFlowLimit = 1 << 21;

struct Level {
	LevelView v;
	double *A = nullptr, *x = nullptr, *r = nullptr;
	float *A32 = nullptr, *x32 = nullptr, *r32 = nullptr;
	// plane-skewed FP64 copies for the exact smoother (levels that do not
	// take the one-workgroup kernel), made by pgmg_compute_rats
	SkewView sv;
	double *As = nullptr, *xs = nullptr, *rs = nullptr;
	float *As32 = nullptr, *xs32 = nullptr, *rs32 = nullptr;  // mixed precision
	long long points() const { return (long long)v.m * v.n * v.o; }
};

}  // namespace

int pgmg::set_error(int code, const char *msg) {
	g_err = msg;
	return code;
}

struct pgmg_hier {
	int device = 0;
	hipStream_t stream = nullptr;
	pgmg_opts opts;
	int nl = 0;
	Level lev[kMaxLevels];
	double *res64 = nullptr;   // residual scratch, top-level size
	double *aux64 = nullptr;   // two top-level vectors for the outside-array operations
	float *res32 = nullptr;
	double *partial = nullptr;
	int partial_cap = 0;
	double *dscalar = nullptr;  // device reduction result; the int after it is the
	                            // exact sweeps' wait flag (set when a wait gave up)
	double *hscalar = nullptr;  // pinned host mirror of both
	int *flow_prog = nullptr;   // plane progress counters of the exact sweeps
	int flow_limit = kFlowLimit;
	double *LU = nullptr;
	int *ipiv = nullptr, *lu_info = nullptr;
	int lu_N = 0, lu_active = 0;
	bool ops_ready = false, f32_ready = false;
	long long gs_count = 0;
	// graphs: [mode][precision]
	hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
	// kernel timing
	int prof_class = KC_NONE, prof_level = 0;
	std::vector<hipEvent_t> prof_ev;
	size_t prof_used = 0;
	long long prof_launches = 0, prof_points = 0;
	double prof_ms = 0;
	bool capturing = false;
};

namespace {

long long level_bytes(const LevelView &v, bool mixed) {
	long long b = (27 + 2) * v.npad * 8;
	if (mixed) b += (27 + 2) * v.npad * 4;
	return b;
}

// Level sizes: coarse m' = (m+2)>>1 on a non-periodic axis, (m+1)>>1 on a
// periodic one (region.cc:13-14), until m*n*o <= 1024 (multigrid.cc:26).
int shape(int m, int n, int o, int prd, int *dims) {
	int l = 0;
	dims[0] = m; dims[1] = n; dims[2] = o;
	while ((long long)dims[3 * l] * dims[3 * l + 1] * dims[3 * l + 2] > 1024) {
		if (l + 1 == kMaxLevels) return -1;
		for (int a = 0; a < 3; a++) dims[3 * (l + 1) + a] = (dims[3 * l + a] + (((prd >> a) & 1) ? 1 : 2)) >> 1;
		l++;
	}
	return l + 1;
}

int dims_prd(const pgmg_dims *d) {
	return (d->x_prd ? 1 : 0) | (d->y_prd ? 2 : 0) | (d->z_prd ? 4 : 0);
}

int check_dims(const pgmg_dims *d) {
	if (!d) return fail(PGMG_ERR_ARG, "null dims");
	if (d->m < 1 || d->n < 1 || d->o < 1) return fail(PGMG_ERR_ARG, "grid sizes must be positive");
	if ((long long)d->m * d->n * d->o > (1ll << 33)) return fail(PGMG_ERR_ARG, "grid too large");
	return PGMG_OK;
}

int check_level(const pgmg_hier *h, int level) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	if (level < 0 || level >= h->nl) return fail(PGMG_ERR_ARG, "level %d out of range [0,%d)", level, h->nl);
	return PGMG_OK;
}

// ---- profiling: events around launches of the selected class ----
struct ProfScope {
	pgmg_hier *h;
	bool on;
	long long pts;
	ProfScope(pgmg_hier *h_, int kc, int level, long long points) : h(h_), on(false), pts(points) {
		if (kc == KC_NONE || h->prof_class != kc || h->prof_level != level || h->capturing) return;
		if (h->prof_used + 2 > h->prof_ev.size()) {
			for (int q = 0; q < 64; q++) {
				hipEvent_t e;
				// timing only: no system-scope fence, whose cache write-back
				// would otherwise land between the timed launches
				if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
				h->prof_ev.push_back(e);
			}
		}
		on = true;
		hipEventRecord(h->prof_ev[h->prof_used], h->stream);
	}
	~ProfScope() {
		if (!on) return;
		hipEventRecord(h->prof_ev[h->prof_used + 1], h->stream);
		h->prof_used += 2;
		h->prof_launches++;
		h->prof_points += pts;
	}
};

void prof_collect(pgmg_hier *h) {
	if (h->prof_used == 0) return;
	hipStreamSynchronize(h->stream);
	for (size_t q = 0; q < h->prof_used; q += 2) {
		float ms = 0;
		if (hipEventElapsedTime(&ms, h->prof_ev[q], h->prof_ev[q + 1]) == hipSuccess) h->prof_ms += ms;
	}
	h->prof_used = 0;
}

// ---- level operations, templated on the arithmetic type ----
template <typename T> T *lA(Level &L);
template <typename T> T *lx(Level &L);
template <typename T> T *lr(Level &L);
template <> double *lA<double>(Level &L) { return L.A; }
template <> double *lx<double>(Level &L) { return L.x; }
template <> double *lr<double>(Level &L) { return L.r; }
template <> float *lA<float>(Level &L) { return L.A32; }
template <> float *lx<float>(Level &L) { return L.x32; }
template <> float *lr<float>(Level &L) { return L.r32; }
// plane-skewed copies of a level (null where none were made)
template <typename T> T *sA(Level &L);
template <typename T> T *sx(Level &L);
template <typename T> T *sr(Level &L);
template <> double *sA<double>(Level &L) { return L.As; }
template <> double *sx<double>(Level &L) { return L.xs; }
template <> double *sr<double>(Level &L) { return L.rs; }
template <> float *sA<float>(Level &L) { return L.As32; }
template <> float *sx<float>(Level &L) { return L.xs32; }
template <> float *sr<float>(Level &L) { return L.rs32; }
template <typename T> T *res_scratch(pgmg_hier *h);
template <> double *res_scratch<double>(pgmg_hier *h) { return h->res64; }
template <> float *res_scratch<float>(pgmg_hier *h) { return h->res32; }

// One workgroup smooths a small level when that beats per-class (or
// per-hyperplane) launches: with x in LDS, for exact mode, and for forward
// colour sweeps when each thread holds one point with its coefficients in
// registers.  (A level of up to 1024 point pairs per pass, 17^3, is faster on
// the colour-pair launches alone: the single workgroup is bound by its one
// CU's L2 bandwidth for the coefficients.  In a batch the small levels of
// many slots run side by side, and the pair-staged form is kept there.)
// Exact forward sweeps of a non-periodic level of more than 1024 points go
// to the one-launch plane-parallel sweeps where those take the level, even
// when x fits in LDS: 17^3 FP64, six sweeps, 0.47 ms against 0.77 ms in one
// workgroup.  (Periodic levels sweep once per launch, with the ghosts
// refreshed in between, and keep the one workgroup.)
template <typename T>
bool gs_use_small(const pgmg_hier *h, const Level &L, int mode) {
	const bool exact = h->opts.smoother == PGMG_SMOOTH_EXACT;
	const bool lds = small_lds_bytes(L.v, (int)sizeof(T)) <= kSmallLdsBytes;
	if (exact) {
		if (!lds) return false;
		if (mode || L.v.prd || L.points() <= 1024) return true;
		return flow_max_sweeps<T>(make_skew_view(L.v), false) < 1;
	}
	if (mode == 0) return lds && L.points() <= 1024;
	return L.points() <= kSmallColour;
}

// Whether the level-0 down-sweeps of a cycle start with a forward FP32
// colour sweep that reads no x (then x need not be cleared before it).  With
// no down-sweep at level 0 (v_down = 0) nothing overwrites x first, so it
// must be cleared.
bool gs_from_zero_ok(const pgmg_hier *h, int level, int mode) {
	if (h->lev[level].v.prd || h->opts.v_down + level < 1) return false;
	return h->opts.smoother != PGMG_SMOOTH_EXACT && mode == 0 && !gs_use_small<float>(h, h->lev[level], mode);
}

int *flow_err(pgmg_hier *h) { return (int *)(h->dscalar + 1); }

// Forward exact sweeps on a level's plane-skewed copies: one launch per
// call where all the level's planes fit on the device at once, else one
// launch per hyperplane.
template <typename T>
void skew_sweeps(pgmg_hier *h, Level &L, int nsweeps) {
	if (!h->flow_prog || !launch_gs_flow<T>(h->stream, L.sv, L.v.prd != 0, sA<T>(L), sx<T>(L), sr<T>(L), nsweeps,
						 h->flow_prog, flow_err(h), h->flow_limit))
		launch_gs_planes_skew<T>(h->stream, L.sv, sA<T>(L), sx<T>(L), sr<T>(L), nsweeps);
}

// from_zero: x is 0 on entry (and may not be cleared in memory); only with
// gs_from_zero_ok.
template <typename T>
void op_gs(pgmg_hier *h, int level, int mode, int nsweeps, bool from_zero = false) {
	Level &L = h->lev[level];
	const bool exact = h->opts.smoother == PGMG_SMOOTH_EXACT;
	h->gs_count += nsweeps;
	if (nsweeps <= 0) return;
	const bool use_small = gs_use_small<T>(h, L, mode);
	if (L.v.prd && !use_small) {
		// Periodic axes: the ghost copies are refreshed before each pass
		// (region.cc:361, 457), then the pass runs on the per-point kernels,
		// which read the halo like any neighbour.  (The one-workgroup kernel
		// refreshes them itself.)
		const int tmax = (L.v.m - 1) + 2 * (L.v.n - 1) + 4 * (L.v.o - 1);
		// The colour-pair row kernels read the halo point m from memory where
		// a row ends, except in the point out of a row of ppl*q+1 points,
		// which takes it as zero: with a periodic x axis such rows use the
		// per-class kernel.
		const bool pair_ok = !(L.v.prd & 1) || !row_lanes(L.v.m, (int)sizeof(T)).ext;
		if (exact && !mode && sA<T>(L) && h->ops_ready) {
			// plane-skewed copies, ghosts refreshed on them before each sweep
			launch_skew_vec<T>(h->stream, L.v, L.sv, lx<T>(L), sx<T>(L), true);
			launch_skew_vec<T>(h->stream, L.v, L.sv, lr<T>(L), sr<T>(L), true);
			for (int s = 0; s < nsweeps; s++) {
				launch_skew_halo_wrap<T>(h->stream, L.v, L.sv, sx<T>(L));
				ProfScope ps(h, KC_GS, level, L.points());
				skew_sweeps<T>(h, L, 1);
			}
			launch_skew_vec<T>(h->stream, L.v, L.sv, sx<T>(L), lx<T>(L), false);
			return;
		}
		for (int s = 0; s < nsweeps; s++)
			for (int pass = 0; pass < (mode ? 2 : 1); pass++) {
				ProfScope ps(h, KC_GS, level, L.points());
				launch_halo_wrap<T>(h->stream, L.v, lx<T>(L));
				if (exact) {
					for (int tt = 0; tt <= tmax; tt++)
						launch_gs_plane<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), pass ? tmax - tt : tt);
				} else if (pass == 0 && pair_ok) {
					for (int q = 0; q < 4; q++)
						launch_gs_pair<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), q, false);
				} else {
					for (int c = 0; c < 8; c++)
						launch_gs_colour<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), pass ? 7 - c : c);
				}
			}
		return;
	}
	if (use_small) {
		ProfScope ps(h, KC_GS, level, L.points() * nsweeps * (mode ? 2 : 1));
		launch_gs_small<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), nsweeps, mode, exact ? 1 : 0);
		return;
	}
	const int tmax = (L.v.m - 1) + 2 * (L.v.n - 1) + 4 * (L.v.o - 1);
	if (exact && !mode) {
		// forward sweeps of one call, pipelined eight hyperplanes apart
		// the copies match the operator from pgmg_compute_rats until an
		// operator is set again (ops_ready false)
		if (sA<T>(L) && h->ops_ready) {
			// on the plane-skewed copies: x and r in, x out (the timing
			// hook brackets the sweeps alone)
			launch_skew_vec<T>(h->stream, L.v, L.sv, lx<T>(L), sx<T>(L), true);
			launch_skew_vec<T>(h->stream, L.v, L.sv, lr<T>(L), sr<T>(L), true);
			{
				ProfScope ps(h, KC_GS, level, L.points() * nsweeps);
				skew_sweeps<T>(h, L, nsweeps);
			}
			launch_skew_vec<T>(h->stream, L.v, L.sv, sx<T>(L), lx<T>(L), false);
			return;
		}
		ProfScope ps(h, KC_GS, level, L.points() * nsweeps);
		launch_gs_planes<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), nsweeps);
		return;
	}
	for (int s = 0; s < nsweeps; s++) {
		for (int pass = 0; pass < (mode ? 2 : 1); pass++) {
			if (exact) {
				ProfScope ps(h, KC_GS, level, L.points());
				for (int tt = 0; tt <= tmax; tt++)
					launch_gs_plane<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), pass ? tmax - tt : tt);
			} else if (pass == 0) {
				// forward: colour pairs (0,1), (2,3), (4,5), (6,7)
				for (int q = 0; q < 4; q++) {
					const int cj = q & 1, ck = q >> 1;
					// a sweep from zero reads less x than a regular one: its
					// launches stay out of the per-launch byte count
					const bool fz = from_zero && s == 0;
					ProfScope ps(h, fz ? KC_NONE : KC_GS, level,
						     (long long)L.v.m * ((L.v.n - cj + 1) / 2) * ((L.v.o - ck + 1) / 2));
					launch_gs_pair<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), q, fz);
				}
			} else {
				// backward: colours 7..0
				for (int c = 7; c >= 0; c--) {
					const int ci = c & 1, cj = (c >> 1) & 1, ck = c >> 2;
					ProfScope ps(h, KC_GS, level, (long long)((L.v.m - ci + 1) / 2) * ((L.v.n - cj + 1) / 2) * ((L.v.o - ck + 1) / 2));
					launch_gs_colour<T>(h->stream, L.v, lA<T>(L), lx<T>(L), lr<T>(L), c);
				}
			}
		}
	}
}

template <typename T>
void op_restriction(pgmg_hier *h, int level) {
	Level &F = h->lev[level - 1], &C = h->lev[level];
	{
		ProfScope ps(h, KC_RESIDUAL, level - 1, F.points());
		launch_residual<T, T>(h->stream, F.v, lA<T>(F), lx<T>(F), lr<T>(F), res_scratch<T>(h), (double *)nullptr);
	}
	ProfScope ps(h, KC_RESTRICT, level, C.points());
	launch_restrict<T>(h->stream, F.v, C.v, res_scratch<T>(h), lr<T>(C), lx<T>(C));
}

template <typename T>
void op_interpolation(pgmg_hier *h, int level) {
	Level &C = h->lev[level], &F = h->lev[level - 1];
	ProfScope ps(h, KC_INTERP, level, F.points());
	launch_interp<T>(h->stream, C.v, F.v, lx<T>(C), lx<T>(F));
}

template <typename T>
void op_solve_exact(pgmg_hier *h, int level, bool smooth) {
	Level &L = h->lev[level];
	if (level == h->nl - 1 && h->lu_active && !smooth) {
		launch_lu_solve<T>(h->stream, L.v, h->lu_N, h->LU, h->ipiv, lr<T>(L), lx<T>(L));
	} else {
		op_gs<T>(h, level, 0, kGsExact);
	}
}

// multigrid::v_cycle, multigrid.cc:88-106.  top_zero: the level-0 x is 0 on
// entry without having been cleared (the mixed-precision correction).
template <typename T>
void op_v_cycle(pgmg_hier *h, int mode, bool top_zero = false) {
	const int L = h->nl;
	for (int i = 0; i < L - 1; i++) {
		op_gs<T>(h, i, mode, h->opts.v_down + i, top_zero && i == 0);
		op_restriction<T>(h, i + 1);
	}
	op_solve_exact<T>(h, L - 1, false);
	for (int i = L - 1; i > 0; i--) {
		op_interpolation<T>(h, i);
		op_gs<T>(h, i - 1, mode, h->opts.v_up + i - 1);
	}
}

int ensure_partial(pgmg_hier *h, int n) {
	if (n <= h->partial_cap) return PGMG_OK;
	if (h->partial) hipFree(h->partial);
	h->partial = nullptr;
	HIP_TRY(hipMalloc(&h->partial, sizeof(double) * (size_t)n));
	h->partial_cap = n;
	return PGMG_OK;
}

// The exact sweeps' wait flag as read back into hscalar[1]: clears it and
// reports the error if it was set.
int flow_flag_result(pgmg_hier *h) {
	int e = 0;
	memcpy(&e, h->hscalar + 1, sizeof(int));
	if (!e) return PGMG_OK;
	launch_fill<int>(h->stream, flow_err(h), 1, 0);
	// the failed launch's last workgroup cleared the counters; clear them
	// here as well before the next sweep
	if (h->flow_prog) launch_fill<int>(h->stream, h->flow_prog, h->lev[0].v.o + 2, 0);
	hipStreamSynchronize(h->stream);
	return fail(PGMG_ERR_HIP, "exact sweep: a plane gave up waiting for its neighbour after %d polls; "
		    "the level's x is incomplete", h->flow_limit);
}

int check_flow(pgmg_hier *h) {
	HIP_TRY(hipMemcpyAsync(h->hscalar + 1, h->dscalar + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
	HIP_TRY(hipStreamSynchronize(h->stream));
	return flow_flag_result(h);
}

int read_scalar(pgmg_hier *h, double *out) {
	HIP_TRY(hipMemcpyAsync(h->hscalar, h->dscalar, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
	HIP_TRY(hipStreamSynchronize(h->stream));
	*out = *h->hscalar;
	return flow_flag_result(h);
}

// Fine residual norm (FP64); optionally write the FP32 residual for the
// mixed-precision inner cycle.
int residual_norm(pgmg_hier *h, int level, float *res32_out, double *out, float *zero32 = nullptr) {
	Level &L = h->lev[level];
	int n = residual_partial_count(L.v);
	int rc = ensure_partial(h, n);
	if (rc) return rc;
	{
		ProfScope ps(h, KC_RESIDUAL, level, L.points());
		if (res32_out) launch_residual<double, float>(h->stream, L.v, L.A, L.x, L.r, res32_out, h->partial, zero32);
		else launch_residual<double, double>(h->stream, L.v, L.A, L.x, L.r, (double *)nullptr, h->partial);
	}
	launch_reduce(h->stream, h->partial, n, h->dscalar);
	return read_scalar(h, out);
}

int run_cycle(pgmg_hier *h, int mode, bool fp32) {
	const int pi = fp32 ? 1 : 0, mi = mode ? 1 : 0;
	if (h->opts.use_graph && h->prof_class == KC_NONE) {
		if (!h->graph[mi][pi]) {
			hipGraph_t g;
			HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
			h->capturing = true;
			long long gc = h->gs_count;
			if (fp32) op_v_cycle<float>(h, mode, gs_from_zero_ok(h, 0, mode)); else op_v_cycle<double>(h, mode);
			h->gs_count = gc;
			h->capturing = false;
			HIP_TRY(hipStreamEndCapture(h->stream, &g));
			hipError_t e = hipGraphInstantiate(&h->graph[mi][pi], g, nullptr, nullptr, 0);
			hipGraphDestroy(g);
			if (e != hipSuccess) { h->graph[mi][pi] = nullptr; return fail(PGMG_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
		}
		// Count sweeps as an eager cycle would.
		long long per = 0;
		for (int i = 0; i < h->nl - 1; i++) per += (h->opts.v_down + i) + (h->opts.v_up + i);
		if (!(h->lu_active)) per += kGsExact;
		h->gs_count += per;
		HIP_TRY(hipGraphLaunch(h->graph[mi][pi], h->stream));
	} else {
		if (fp32) op_v_cycle<float>(h, mode, gs_from_zero_ok(h, 0, mode)); else op_v_cycle<double>(h, mode);
	}
	return PGMG_OK;
}

void free_graphs(pgmg_hier *h) {
	for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++)
		if (h->graph[a][b]) { hipGraphExecDestroy(h->graph[a][b]); h->graph[a][b] = nullptr; }
}

}  // namespace

// ======================================================================
// C ABI
// ======================================================================
extern "C" {

const char *pgmg_last_error(void) { return g_err.c_str(); }
int pgmg_version(void) { return PGMG_CAPI_VERSION; }

void pgmg_default_opts(pgmg_opts *o) {
	o->precision = PGMG_PREC_FP64;
	o->smoother = PGMG_SMOOTH_COLOUR8;
	o->bottom = PGMG_BOTTOM_GS24;
	o->v_down = 2; o->v_up = 2; o->v_bottom = 20;
	o->use_graph = 1;
	o->max_levels = 0;
}

int pgmg_hierarchy_shape(const pgmg_dims *d, int *nlevels, int *dims_out) {
	int rc = check_dims(d);
	if (rc) return rc;
	int tmp[3 * kMaxLevels];
	int nl = shape(d->m, d->n, d->o, dims_prd(d), tmp);
	if (nl < 0) return fail(PGMG_ERR_MULTIGRID, "Maximum number of multigrid levels reached");
	if (nlevels) *nlevels = nl;
	if (dims_out) memcpy(dims_out, tmp, sizeof(int) * 3 * nl);
	return PGMG_OK;
}

int pgmg_hier_footprint(const pgmg_dims *d, const pgmg_opts *opts, uint64_t *bytes) {
	int rc = check_dims(d);
	if (rc) return rc;
	int dims[3 * kMaxLevels];
	int nl = shape(d->m, d->n, d->o, dims_prd(d), dims);
	if (nl < 0) return fail(PGMG_ERR_MULTIGRID, "Maximum number of multigrid levels reached");
	if (opts && opts->max_levels > 0 && opts->max_levels < nl) nl = opts->max_levels;
	const bool mixed = opts && opts->precision == PGMG_PREC_MIXED;
	uint64_t b = 0;
	for (int l = 0; l < nl; l++) b += (uint64_t)level_bytes(make_level_view(dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]), mixed);
	LevelView top = make_level_view(d->m, d->n, d->o);
	top.prd = dims_prd(d);
	b += (uint64_t)top.npad * (8 + (mixed ? 4 : 0));
	b += (uint64_t)residual_partial_count(top) * 8;
	if (opts && opts->smoother == PGMG_SMOOTH_EXACT) {
		// plane-skewed copies of the levels the one-workgroup kernel does not take
		for (int l = 0; l < nl; l++) {
			const LevelView v = make_level_view(dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]);
			if (small_lds_bytes(v, 8) <= kSmallLdsBytes && (long long)v.m * v.n * v.o <= 1024) continue;
			const unsigned long long n = (unsigned long long)make_skew_view(v).size;
			if (29ULL * n * 8 > kSkewMaxBytes) continue;
			b += (uint64_t)n * 29 * ((opts->precision == PGMG_PREC_MIXED) ? 12 : 8);
		}
	}
	if (opts && opts->bottom == PGMG_BOTTOM_LU) {
		LevelView bot = make_level_view(dims[3 * (nl - 1)], dims[3 * (nl - 1) + 1], dims[3 * (nl - 1) + 2]);
		long long N = (long long)bot.m * bot.n * bot.o;
		b += (uint64_t)(N * N * 8 + N * 4);
	}
	*bytes = b;
	return PGMG_OK;
}

int pgmg_hier_create(const pgmg_dims *d, const pgmg_opts *opts, pgmg_hier **out) {
	if (!out) return fail(PGMG_ERR_ARG, "null output");
	*out = nullptr;
	int rc = check_dims(d);
	if (rc) return rc;
	pgmg_opts o;
	if (opts) o = *opts; else pgmg_default_opts(&o);
	if (o.precision != PGMG_PREC_FP64 && o.precision != PGMG_PREC_MIXED) return fail(PGMG_ERR_ARG, "bad precision");
	if (o.smoother != PGMG_SMOOTH_EXACT && o.smoother != PGMG_SMOOTH_COLOUR8) return fail(PGMG_ERR_ARG, "bad smoother");
	if (o.bottom != PGMG_BOTTOM_GS24 && o.bottom != PGMG_BOTTOM_LU) return fail(PGMG_ERR_ARG, "bad bottom");
	if (o.v_down < 0 || o.v_up < 0) return fail(PGMG_ERR_ARG, "negative sweep counts");
	if (o.max_levels < 0) return fail(PGMG_ERR_ARG, "negative max_levels");
	int dims[3 * kMaxLevels];
	int nl = shape(d->m, d->n, d->o, dims_prd(d), dims);
	if (nl < 0) return fail(PGMG_ERR_MULTIGRID, "Maximum number of multigrid levels reached");
	if (o.max_levels > 0 && o.max_levels < nl) nl = o.max_levels;
	int ndev = 0;
	HIP_TRY(hipGetDeviceCount(&ndev));
	if (ndev < 1) return fail(PGMG_ERR_HIP, "no HIP device");
	pgmg_hier *h = new pgmg_hier();
	h->opts = o;
	h->nl = nl;
	hipGetDevice(&h->device);
	const bool mixed = o.precision == PGMG_PREC_MIXED;
#define ALLOC(ptr, bytes)                                                                     \
	do {                                                                                  \
		hipError_t e_ = hipMalloc((void **)&(ptr), (bytes));                          \
		if (e_ != hipSuccess) { pgmg_hier_destroy(h); return fail(PGMG_ERR_MEMORY, "device allocation of %zu bytes failed: %s", (size_t)(bytes), hipGetErrorString(e_)); } \
	} while (0)
	if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
		delete h;
		return fail(PGMG_ERR_HIP, "stream creation failed");
	}
	for (int l = 0; l < nl; l++) {
		Level &L = h->lev[l];
		L.v = make_level_view(dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]);
		L.v.cache_coef = 1;  // one level runs at a time
		L.v.prd = dims_prd(d);
		const size_t np = (size_t)L.v.npad;
		ALLOC(L.A, 27 * np * 8);
		ALLOC(L.x, np * 8);
		ALLOC(L.r, np * 8);
		launch_fill<double>(h->stream, L.A, 27 * (long long)np, 0.0);
		launch_fill<double>(h->stream, L.x, (long long)np, 0.0);
		launch_fill<double>(h->stream, L.r, (long long)np, 0.0);
		if (mixed) {
			ALLOC(L.A32, 27 * np * 4);
			ALLOC(L.x32, np * 4);
			ALLOC(L.r32, np * 4);
			launch_fill<float>(h->stream, L.A32, 27 * (long long)np, 0.0f);
			launch_fill<float>(h->stream, L.x32, (long long)np, 0.0f);
			launch_fill<float>(h->stream, L.r32, (long long)np, 0.0f);
		}
	}
	const size_t np0 = (size_t)h->lev[0].v.npad;
	ALLOC(h->res64, np0 * 8);
	launch_fill<double>(h->stream, h->res64, (long long)np0, 0.0);
	if (mixed) {
		ALLOC(h->res32, np0 * 4);
		launch_fill<float>(h->stream, h->res32, (long long)np0, 0.0f);
	}
	ALLOC(h->dscalar, 16);
	launch_fill<double>(h->stream, h->dscalar, 2, 0.0);
	if (hipHostMalloc((void **)&h->hscalar, 16) != hipSuccess) { pgmg_hier_destroy(h); return fail(PGMG_ERR_MEMORY, "pinned alloc failed"); }
	rc = ensure_partial(h, residual_partial_count(h->lev[0].v));
	if (rc) { pgmg_hier_destroy(h); return rc; }
#undef ALLOC
	if (hipStreamSynchronize(h->stream) != hipSuccess) { pgmg_hier_destroy(h); return fail(PGMG_ERR_HIP, "initialisation failed"); }
	*out = h;
	return PGMG_OK;
}

int pgmg_hier_destroy(pgmg_hier *h) {
	if (!h) return PGMG_OK;
	if (h->stream) hipStreamSynchronize(h->stream);
	free_graphs(h);
	for (int l = 0; l < kMaxLevels; l++) {
		Level &L = h->lev[l];
		hipFree(L.A); hipFree(L.x); hipFree(L.r);
		hipFree(L.A32); hipFree(L.x32); hipFree(L.r32);
		hipFree(L.As); hipFree(L.xs); hipFree(L.rs);
		hipFree(L.As32); hipFree(L.xs32); hipFree(L.rs32);
	}
	hipFree(h->res64); hipFree(h->res32); hipFree(h->partial); hipFree(h->dscalar); hipFree(h->aux64);
	if (h->hscalar) hipHostFree(h->hscalar);
	hipFree(h->LU); hipFree(h->ipiv); hipFree(h->lu_info); hipFree(h->flow_prog);
	for (auto e : h->prof_ev) hipEventDestroy(e);
	if (h->stream) hipStreamDestroy(h->stream);
	delete h;
	return PGMG_OK;
}

int pgmg_num_levels(const pgmg_hier *h, int *nl) {
	if (!h || !nl) return fail(PGMG_ERR_ARG, "null argument");
	*nl = h->nl;
	return PGMG_OK;
}

int pgmg_level_dims(const pgmg_hier *h, int level, int dims[3]) {
	int rc = check_level(h, level);
	if (rc) return rc;
	dims[0] = h->lev[level].v.m; dims[1] = h->lev[level].v.n; dims[2] = h->lev[level].v.o;
	return PGMG_OK;
}

static int upload_op(pgmg_hier *h, int level, const double *A27) {
	Level &L = h->lev[level];
	const long long N = L.points();
	const long long chunk = 1 << 20;
	double *stage = nullptr;
	HIP_TRY(hipMalloc(&stage, sizeof(double) * 27 * (size_t)std::min(N, chunk)));
	for (long long p0 = 0; p0 < N; p0 += chunk) {
		const long long cnt = std::min(chunk, N - p0);
		hipError_t e = hipMemcpyAsync(stage, A27 + 27 * p0, sizeof(double) * 27 * cnt, hipMemcpyHostToDevice, h->stream);
		if (e != hipSuccess) { hipFree(stage); return fail(PGMG_ERR_HIP, "upload: %s", hipGetErrorString(e)); }
		launch_pack_op(h->stream, L.v, stage, L.A, p0, cnt);
		hipStreamSynchronize(h->stream);
	}
	hipFree(stage);
	return PGMG_OK;
}

int pgmg_set_operator27(pgmg_hier *h, const double *A27) {
	if (!h || !A27) return fail(PGMG_ERR_ARG, "null argument");
	int rc = upload_op(h, 0, A27);
	if (rc) return rc;
	h->ops_ready = false;
	free_graphs(h);
	return PGMG_OK;
}

int pgmg_generate_problem(pgmg_hier *h, int family, uint64_t seed) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	if (family != PGMG_FAM_FACE7_UNIFORM && family != PGMG_FAM_FACE7_LOGUNI && family != PGMG_FAM_Q1_ANISO)
		return fail(PGMG_ERR_ARG, "unknown problem family %d", family);
	launch_generate(h->stream, h->lev[0].v, family, seed, h->lev[0].A, h->lev[0].r);
	HIP_TRY(hipStreamSynchronize(h->stream));
	h->ops_ready = false;
	free_graphs(h);
	return PGMG_OK;
}

// enable_exact + lu_decomposition of the bottom level (region_tr.cc:472-534):
// dense assembly and factor; lu_active is whether the factor passed the test
// of region_tr.cc:526-533 (else the bottom falls back on 24 sweeps).
static int factor_bottom(pgmg_hier *h) {
	h->lu_active = 0;
	Level &B = h->lev[h->nl - 1];
	const long long N = B.points();
	if (N >= kMaxExact) return PGMG_OK;  // enable_exact's size rule, region_tr.cc:472-478
	if (!h->LU) {
		HIP_TRY(hipMalloc(&h->LU, sizeof(double) * (size_t)(N * N)));
		HIP_TRY(hipMalloc(&h->ipiv, sizeof(int) * (size_t)N));
		HIP_TRY(hipMalloc(&h->lu_info, sizeof(int) * 2));
	}
	h->lu_N = (int)N;
	launch_fill<double>(h->stream, h->LU, N * N, 0.0);
	launch_lu_assemble(h->stream, B.v, B.A, h->LU);
	launch_lu_factor(h->stream, (int)N, h->LU, h->ipiv, h->lu_info);
	int info[2] = {0, 1};
	HIP_TRY(hipMemcpyAsync(info, h->lu_info, sizeof(info), hipMemcpyDeviceToHost, h->stream));
	HIP_TRY(hipStreamSynchronize(h->stream));
	h->lu_active = info[1] == 0;
	return PGMG_OK;
}

int pgmg_compute_rats(pgmg_hier *h) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	free_graphs(h);
	for (int l = 1; l < h->nl; l++) launch_rat(h->stream, h->lev[l - 1].v, h->lev[l].v, h->lev[l - 1].A, h->lev[l].A);
	h->lu_active = 0;
	if (h->opts.bottom == PGMG_BOTTOM_LU) {
		int rc = factor_bottom(h);
		if (rc) return rc;
	}
	if (h->opts.precision == PGMG_PREC_MIXED) {
		for (int l = 0; l < h->nl; l++) launch_round_f32(h->stream, h->lev[l].A, h->lev[l].A32, 27 * h->lev[l].v.npad);
	}
	if (h->opts.smoother == PGMG_SMOOTH_EXACT) {
		if (!h->flow_prog) {
			// one progress counter per plane of the largest level, and the
			// count of workgroups that have left
			const long long nc = h->lev[0].v.o + 2;
			HIP_TRY(hipMalloc(&h->flow_prog, sizeof(int) * (size_t)nc));
			launch_fill<int>(h->stream, h->flow_prog, nc, 0);
		}
		// plane-skewed copies for the exact sweeps of large levels (about 1.5
		// times a level's size each: skipped, and the level layout swept
		// instead, when they would not fit)
		for (int l = 0; l < h->nl; l++) {
			Level &L = h->lev[l];
			if (gs_use_small<double>(h, L, 0) &&
			    (h->opts.precision != PGMG_PREC_MIXED || gs_use_small<float>(h, L, 0)))
				continue;
			if (!L.As) {
				const SkewView sv = make_skew_view(L.v);
				const size_t n = (size_t)sv.size;
				size_t fr = 0, tot = 0;
				if (29ULL * n * 8 > kSkewMaxBytes) continue;
				if (hipMemGetInfo(&fr, &tot) != hipSuccess || 29 * n * 8 > fr / 2) continue;
				if (hipMalloc(&L.As, 27 * n * 8) != hipSuccess || hipMalloc(&L.xs, n * 8) != hipSuccess ||
				    hipMalloc(&L.rs, n * 8) != hipSuccess) {
					hipFree(L.As); hipFree(L.xs); hipFree(L.rs);
					L.As = L.xs = L.rs = nullptr;
					(void)hipGetLastError();
					continue;
				}
				L.sv = sv;
				launch_fill<double>(h->stream, L.As, 27 * (long long)n, 0.0);
				launch_fill<double>(h->stream, L.xs, (long long)n, 0.0);
				launch_fill<double>(h->stream, L.rs, (long long)n, 0.0);
			}
			launch_skew_pack_op<double>(h->stream, L.v, L.sv, L.A, L.As);
			if (h->opts.precision != PGMG_PREC_MIXED) continue;
			if (!L.As32) {
				const size_t n = (size_t)L.sv.size;
				size_t fr = 0, tot = 0;
				if (hipMemGetInfo(&fr, &tot) != hipSuccess || 29 * n * 4 > fr / 2) continue;
				if (hipMalloc(&L.As32, 27 * n * 4) != hipSuccess || hipMalloc(&L.xs32, n * 4) != hipSuccess ||
				    hipMalloc(&L.rs32, n * 4) != hipSuccess) {
					hipFree(L.As32); hipFree(L.xs32); hipFree(L.rs32);
					L.As32 = L.xs32 = L.rs32 = nullptr;
					(void)hipGetLastError();
					continue;
				}
				launch_fill<float>(h->stream, L.As32, 27 * (long long)n, 0.0f);
				launch_fill<float>(h->stream, L.xs32, (long long)n, 0.0f);
				launch_fill<float>(h->stream, L.rs32, (long long)n, 0.0f);
			}
			launch_skew_pack_op<float>(h->stream, L.v, L.sv, L.A32, L.As32);
		}
	}
	HIP_TRY(hipStreamSynchronize(h->stream));
	h->ops_ready = true;
	return PGMG_OK;
}

int pgmg_lu_decomposition(pgmg_hier *h) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	free_graphs(h);
	int rc = factor_bottom(h);
	if (rc) return rc;
	// as Aexact is kept (and refactored by compute_rats, region.cc:659)
	// exactly when the factor succeeded
	h->opts.bottom = h->lu_active ? PGMG_BOTTOM_LU : PGMG_BOTTOM_GS24;
	return PGMG_OK;
}

int pgmg_get_lu_factor(pgmg_hier *h, double *LU, int *ipiv) {
	if (!h || !LU || !ipiv) return fail(PGMG_ERR_ARG, "null argument");
	if (!h->lu_active) return fail(PGMG_ERR_ARG, "no bottom LU factor is in use");
	const size_t N = (size_t)h->lu_N;
	HIP_TRY(hipMemcpyAsync(LU, h->LU, sizeof(double) * N * N, hipMemcpyDeviceToHost, h->stream));
	HIP_TRY(hipMemcpyAsync(ipiv, h->ipiv, sizeof(int) * N, hipMemcpyDeviceToHost, h->stream));
	HIP_TRY(hipStreamSynchronize(h->stream));
	for (size_t q = 0; q < N; q++) ipiv[q] += 1;  // LAPACK's 1-based rows
	return PGMG_OK;
}

int pgmg_lu_active(const pgmg_hier *h, int *active) {
	if (!h || !active) return fail(PGMG_ERR_ARG, "null argument");
	*active = h->lu_active;
	return PGMG_OK;
}

static int vec_io(pgmg_hier *h, int level, double *dptr, const double *hin, double *hout) {
	Level &L = h->lev[level];
	const long long N = L.points();
	double *stage = nullptr;
	HIP_TRY(hipMalloc(&stage, sizeof(double) * (size_t)N));
	hipError_t e = hipSuccess;
	if (hin) {
		e = hipMemcpyAsync(stage, hin, sizeof(double) * N, hipMemcpyHostToDevice, h->stream);
		if (e == hipSuccess) launch_pack_vec(h->stream, L.v, stage, dptr, 0, N);
	} else {
		launch_unpack_vec(h->stream, L.v, dptr, stage, 0, N);
		e = hipMemcpyAsync(hout, stage, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream);
	}
	if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
	hipFree(stage);
	if (e != hipSuccess) return fail(PGMG_ERR_HIP, "vector transfer: %s", hipGetErrorString(e));
	return PGMG_OK;
}

int pgmg_set_rhs(pgmg_hier *h, int level, const double *b) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!b) return fail(PGMG_ERR_ARG, "null array");
	return vec_io(h, level, h->lev[level].r, b, nullptr);
}
int pgmg_get_rhs(pgmg_hier *h, int level, double *b) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!b) return fail(PGMG_ERR_ARG, "null array");
	return vec_io(h, level, h->lev[level].r, nullptr, b);
}
int pgmg_set_x(pgmg_hier *h, int level, const double *x) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!x) return fail(PGMG_ERR_ARG, "null array");
	return vec_io(h, level, h->lev[level].x, x, nullptr);
}
int pgmg_get_x(pgmg_hier *h, int level, double *x) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!x) return fail(PGMG_ERR_ARG, "null array");
	rc = vec_io(h, level, h->lev[level].x, nullptr, x);
	return rc ? rc : check_flow(h);
}

int pgmg_get_operator27(pgmg_hier *h, int level, double *A27) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!A27) return fail(PGMG_ERR_ARG, "null array");
	Level &L = h->lev[level];
	const long long N = L.points(), chunk = 1 << 20;
	double *stage = nullptr;
	HIP_TRY(hipMalloc(&stage, sizeof(double) * 27 * (size_t)std::min(N, chunk)));
	for (long long p0 = 0; p0 < N; p0 += chunk) {
		const long long cnt = std::min(chunk, N - p0);
		launch_unpack_op(h->stream, L.v, L.A, stage, p0, cnt);
		hipError_t e = hipMemcpyAsync(A27 + 27 * p0, stage, sizeof(double) * 27 * cnt, hipMemcpyDeviceToHost, h->stream);
		if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
		if (e != hipSuccess) { hipFree(stage); return fail(PGMG_ERR_HIP, "download: %s", hipGetErrorString(e)); }
	}
	hipFree(stage);
	return PGMG_OK;
}

int pgmg_clear_x(pgmg_hier *h, int level) {
	int rc = check_level(h, level);
	if (rc) return rc;
	launch_fill<double>(h->stream, h->lev[level].x, h->lev[level].v.npad, 0.0);
	HIP_TRY(hipStreamSynchronize(h->stream));
	return PGMG_OK;
}

int pgmg_gauss_seidel(pgmg_hier *h, int mode, int level, int nsweeps) {
	int rc = check_level(h, level);
	if (rc) return rc;
	op_gs<double>(h, level, mode, nsweeps);
	HIP_TRY(hipGetLastError());
	return PGMG_OK;
}

int pgmg_restriction(pgmg_hier *h, int level) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (level < 1) return fail(PGMG_ERR_ARG, "restriction needs level >= 1");
	op_restriction<double>(h, level);
	HIP_TRY(hipGetLastError());
	return PGMG_OK;
}

int pgmg_interpolation(pgmg_hier *h, int level) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (level < 1) return fail(PGMG_ERR_ARG, "interpolation needs level >= 1");
	op_interpolation<double>(h, level);
	HIP_TRY(hipGetLastError());
	return PGMG_OK;
}

int pgmg_solve_exact(pgmg_hier *h, int level, int smooth) {
	int rc = check_level(h, level);
	if (rc) return rc;
	op_solve_exact<double>(h, level, smooth != 0);
	HIP_TRY(hipGetLastError());
	return PGMG_OK;
}

int pgmg_v_cycle(pgmg_hier *h, int mode) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	if (!h->ops_ready) return fail(PGMG_ERR_SETUP, "operators not set up (compute_rats)");
	int rc = run_cycle(h, mode, false);
	if (rc) return rc;
	HIP_TRY(hipGetLastError());
	return PGMG_OK;
}

int pgmg_set_sweeps(pgmg_hier *h, int v_down, int v_up) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	if (v_down < 0 || v_up < 0) return fail(PGMG_ERR_ARG, "negative sweep counts");
	if (v_down != h->opts.v_down || v_up != h->opts.v_up) {
		free_graphs(h);
		h->opts.v_down = v_down;
		h->opts.v_up = v_up;
	}
	return PGMG_OK;
}

int pgmg_l2_error_all(pgmg_hier *h, int level, double *out) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!out) return fail(PGMG_ERR_ARG, "null output");
	return residual_norm(h, level, nullptr, out);
}

int pgmg_rhs_l2norm(pgmg_hier *h, int level, double *out) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!out) return fail(PGMG_ERR_ARG, "null output");
	Level &L = h->lev[level];
	const int n = ((L.v.m + 63) / 64) * ((L.v.n + 3) / 4) * L.v.o;
	rc = ensure_partial(h, n);
	if (rc) return rc;
	launch_sumsq<double>(h->stream, L.v, L.r, h->partial);
	launch_reduce(h->stream, h->partial, n, h->dscalar);
	return read_scalar(h, out);
}

int pgmg_solve(pgmg_hier *h, double tol, int max_cycles, int *cycles, double *hist) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	if (!h->ops_ready) return fail(PGMG_ERR_SETUP, "operators not set up (compute_rats)");
	double bn = 0;
	int rc = pgmg_rhs_l2norm(h, 0, &bn);
	if (rc) return rc;
	const bool mixed = h->opts.precision == PGMG_PREC_MIXED;
	Level &T0 = h->lev[0];
	int c = 0;
	double e = 0;
	for (;;) {
		// mixed: the FP64 defect goes to r32, and the FP32 correction x32 is
		// cleared in the same pass unless the cycle's first sweep starts from
		// zero without reading it
		float *zero32 = mixed && !gs_from_zero_ok(h, 0, 0) ? T0.x32 : nullptr;
		rc = residual_norm(h, 0, mixed ? T0.r32 : nullptr, &e, zero32);
		if (rc) return rc;
		if (hist) hist[c] = e;
		if (std::isnan(e)) { if (cycles) *cycles = c; return fail(PGMG_ERR_NAN, "NaN residual after %d cycles", c); }
		if (sqrt(e / bn) < tol) break;
		if (c >= max_cycles) { if (cycles) *cycles = c; return fail(PGMG_ERR_NOT_CONVERGED, "not converged in %d cycles", c); }
		if (mixed) {
			// defect correction: r32 holds (float)(b - A x), x32 = 0 (in
			// memory or by the first sweep); one FP32 V-cycle; x += x32.
			rc = run_cycle(h, 0, true);
			if (rc) return rc;
			launch_add_f32_to_f64(h->stream, T0.v, T0.x32, T0.x);
		} else {
			rc = run_cycle(h, 0, false);
			if (rc) return rc;
		}
		c++;
	}
	if (cycles) *cycles = c;
	HIP_TRY(hipGetLastError());
	return PGMG_OK;
}

int pgmg_residual(pgmg_hier *h, int level, double *res) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!res) return fail(PGMG_ERR_ARG, "null array");
	Level &L = h->lev[level];
	launch_residual<double, double>(h->stream, L.v, L.A, L.x, L.r, h->res64, (double *)nullptr);
	return vec_io(h, level, h->res64, nullptr, res);
}

// Two level-sized device vectors with zero halos for the operations on
// outside arrays (allocated on first use).
static int aux_vectors(pgmg_hier *h, int level, double **a, double **b) {
	const long long np0 = h->lev[0].v.npad;
	if (!h->aux64) HIP_TRY(hipMalloc(&h->aux64, sizeof(double) * 2 * (size_t)np0));
	const long long np = h->lev[level].v.npad;
	launch_fill<double>(h->stream, h->aux64, np, 0.0);
	launch_fill<double>(h->stream, h->aux64 + np0, np, 0.0);
	*a = h->aux64;
	*b = h->aux64 + np0;
	return PGMG_OK;
}

int pgmg_apply_operator(pgmg_hier *h, int level, const double *in, double *out) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!in || !out) return fail(PGMG_ERR_ARG, "null array");
	double *din, *dout;
	rc = aux_vectors(h, level, &din, &dout);
	if (rc) return rc;
	rc = vec_io(h, level, din, in, nullptr);
	if (rc) return rc;
	Level &L = h->lev[level];
	launch_halo_wrap<double>(h->stream, L.v, din);
	launch_apply(h->stream, L.v, L.A, din, dout);
	return vec_io(h, level, dout, nullptr, out);
}

int pgmg_copy_rhs_to_x(pgmg_hier *h, int level) {
	int rc = check_level(h, level);
	if (rc) return rc;
	Level &L = h->lev[level];
	HIP_TRY(hipMemcpyAsync(L.x, L.r, sizeof(double) * (size_t)L.v.npad, hipMemcpyDeviceToDevice, h->stream));
	HIP_TRY(hipStreamSynchronize(h->stream));
	return PGMG_OK;
}

int pgmg_rhs_x_dot(pgmg_hier *h, int level, double *out) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!out) return fail(PGMG_ERR_ARG, "null output");
	Level &L = h->lev[level];
	const int n = ((L.v.m + 63) / 64) * ((L.v.n + 3) / 4) * L.v.o;
	rc = ensure_partial(h, n);
	if (rc) return rc;
	launch_pair_sum(h->stream, L.v, L.x, L.r, 0, h->partial);
	launch_reduce(h->stream, h->partial, n, h->dscalar);
	return read_scalar(h, out);
}

int pgmg_x_diff2(pgmg_hier *h, int level, const double *ext_x, double *out) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!ext_x || !out) return fail(PGMG_ERR_ARG, "null argument");
	double *dext, *unused;
	rc = aux_vectors(h, level, &dext, &unused);
	if (rc) return rc;
	rc = vec_io(h, level, dext, ext_x, nullptr);
	if (rc) return rc;
	Level &L = h->lev[level];
	const int n = ((L.v.m + 63) / 64) * ((L.v.n + 3) / 4) * L.v.o;
	rc = ensure_partial(h, n);
	if (rc) return rc;
	launch_pair_sum(h->stream, L.v, dext, L.x, 1, h->partial);
	launch_reduce(h->stream, h->partial, n, h->dscalar);
	return read_scalar(h, out);
}

// ---- pgmg_x_diff2_noalloc: no device allocation ----
// Device memory of the entry point: the caller's vector (point-major), the
// block sums and the result, in one array that every GPU the library runs
// on holds from the first use until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the current
// device (the array's name is a host-side stand-in).
const long long kScratchDoubles = 8LL << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
const int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kScratchDevices] = {};

// The scratch array's address on device `dev`, looked up once per device
// with that device current.
static int scratch_on(int dev, double **out) {
	if (dev < 0 || dev >= kScratchDevices) return fail(PGMG_ERR_ARG, "device %d beyond the scratch table", dev);
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		int cur = -1;
		HIP_TRY(hipGetDevice(&cur));
		if (cur != dev) HIP_TRY(hipSetDevice(dev));
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (cur != dev) (void)hipSetDevice(cur);
		if (e != hipSuccess || !p) return fail(PGMG_ERR_HIP, "scratch address: %s", hipGetErrorString(e));
		g_scratch_ptr[dev] = (double *)p;
	}
	*out = g_scratch_ptr[dev];
	return PGMG_OK;
}

int pgmg_x_diff2_noalloc(pgmg_hier *h, int level, const double *ext_x, double *out) {
	int rc = check_level(h, level);
	if (rc) return rc;
	if (!ext_x || !out) return fail(PGMG_ERR_ARG, "null argument");
	Level &L = h->lev[level];
	const long long N = L.points();
	const long long nparts = (long long)((L.v.m + 63) / 64) * ((L.v.n + 3) / 4) * L.v.o;
	// the vector, one sum per block, the result
	if (N + nparts + 1 > kScratchDoubles)
		return fail(PGMG_ERR_ARG, "level %d has %lld points: the allocation-free soln_diff takes levels whose "
			    "points plus %lld block sums fit %lld doubles; use pgmg_x_diff2", level, N, nparts, kScratchDoubles - 1);
	double *sp = nullptr;
	rc = scratch_on(h->device, &sp);
	if (rc) return rc;
	double *ext = sp, *partial = sp + N, *res = sp + N + nparts;
	HIP_TRY(hipMemcpyAsync(ext, ext_x, sizeof(double) * (size_t)N, hipMemcpyHostToDevice, h->stream));
	launch_diff2_compact(h->stream, L.v, ext, L.x, partial);
	launch_reduce(h->stream, partial, (int)nparts, res);
	HIP_TRY(hipMemcpyAsync(h->hscalar, res, sizeof(double), hipMemcpyDeviceToHost, h->stream));
	HIP_TRY(hipStreamSynchronize(h->stream));
	*out = *h->hscalar;
	return PGMG_OK;
}

int pgmg_synchronize(pgmg_hier *h) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	HIP_TRY(hipStreamSynchronize(h->stream));
	return check_flow(h);
}

int pgmg_set_wait_limit(pgmg_hier *h, int polls) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	if (polls < 0) return fail(PGMG_ERR_ARG, "negative wait limit");
	if (polls != h->flow_limit) {
		free_graphs(h);  // the limit is a launch argument
		h->flow_limit = polls;
	}
	return PGMG_OK;
}

int pgmg_gs_count(const pgmg_hier *h, long long *count) {
	if (!h || !count) return fail(PGMG_ERR_ARG, "null argument");
	*count = h->gs_count;
	return PGMG_OK;
}

int pgmg_prof_enable(pgmg_hier *h, int kernel_class, int level) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	if (kernel_class < KC_NONE || kernel_class > KC_INTERP) return fail(PGMG_ERR_ARG, "bad kernel class");
	prof_collect(h);
	h->prof_class = kernel_class;
	h->prof_level = level;
	h->prof_launches = 0;
	h->prof_points = 0;
	h->prof_ms = 0;
	return PGMG_OK;
}

int pgmg_prof_read(pgmg_hier *h, long long *launches, double *total_ms, long long *points) {
	if (!h) return fail(PGMG_ERR_ARG, "null hierarchy");
	prof_collect(h);
	if (points) *points = h->prof_points;
	if (launches) *launches = h->prof_launches;
	if (total_ms) *total_ms = h->prof_ms;
	return PGMG_OK;
}

}  // extern "C"
