// pgmg_batch.hip - batches of independent hierarchies and the serving loop
// (C ABI pgmg_batch_*, pgmg_serve_run).
//
// All slots live in one HBM arena reserved at creation; a slot's arrays are
// one contiguous arena block.  A plan (work lists per operation and level,
// see WorkList) is rebuilt on the host whenever the set of active slots
// changes and copied to the device; every cycle then issues one launch per
// (operation, level) for all slots at once.  New slots are set up (operator
// generation, Galerkin levels, LU, FP32 copies, source norm) on a second
// stream while the others keep cycling, and join at the next step.
#include "pgmg_internal.hpp"
#include "../../include/pgmg_capi.h"
#include "../../include/pgmg_problems.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

using namespace pgmg;

namespace {

int bfail(int code, const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	return set_error(code, buf);
}

#define BHIP(call)                                                                           \
	do {                                                                                 \
		hipError_t e_ = (call);                                                      \
		if (e_ != hipSuccess) return bfail(PGMG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
	} while (0)

const int kGs24 = 24;
const uint64_t kAlign = 256;

uint64_t al(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, cap) with coalescing.
struct Arena {
	char *base = nullptr;
	uint64_t cap = 0, used = 0;
	std::map<uint64_t, uint64_t> free_;  // offset -> size
	void init(char *b, uint64_t c) { base = b; cap = c; free_.clear(); free_[0] = c; used = 0; }
	bool alloc(uint64_t sz, uint64_t &off) {
		sz = al(sz);
		for (auto it = free_.begin(); it != free_.end(); ++it) {
			if (it->second >= sz) {
				off = it->first;
				const uint64_t rest = it->second - sz;
				free_.erase(it);
				if (rest) free_[off + sz] = rest;
				used += sz;
				return true;
			}
		}
		return false;
	}
	void release(uint64_t off, uint64_t sz) {
		sz = al(sz);
		used -= sz;
		auto it = free_.emplace(off, sz).first;
		auto nx = std::next(it);
		if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
		if (it != free_.begin()) {
			auto pv = std::prev(it);
			if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
		}
	}
};

enum { ST_FREE = 0, ST_SETUP = 1, ST_ACTIVE = 2, ST_DONE = 3, ST_FAILED = 4 };

struct SlotHost {
	int state = ST_FREE;
	int nl = 0;
	int dims[3 * kBatchMaxLevels];
	uint64_t off = 0, bytes = 0;
	double tol = 0;
	int max_cycles = 0, cycles = 0;
	double bn = 0, rel = 1;
	bool want_lu = false;
	uint64_t seq = 0;  // admission order
	hipEvent_t ready = nullptr;
	SlotDev dev;
};

int shape_dims(int m, int n, int o, int *dims) {
	int l = 0;
	dims[0] = m; dims[1] = n; dims[2] = o;
	while ((long long)dims[3 * l] * dims[3 * l + 1] * dims[3 * l + 2] > 1024) {
		if (l + 1 == kBatchMaxLevels) return -1;
		for (int a = 0; a < 3; a++) dims[3 * (l + 1) + a] = (dims[3 * l + a] + 2) >> 1;
		l++;
	}
	return l + 1;
}

// Lists of the plan.
enum { L_PAIR0 = 0, L_PAIR1, L_PAIR2, L_PAIR3, L_SMALL, L_RES, L_RESTRICT, L_INTERP, L_LU, L_GS24, L_PER_LEVEL };
enum { L_TOP_ROWS = 0, L_TOP_ONE, L_TOP_PTS, L_TOP_COUNT };

// Tile shapes of the batched launches: work units per workgroup, one choice
// per launch class, measured on the batch's own first steps.  Trial t
// (t = 0 .. kTuneSteps-1, two rounds over the candidates) runs the launches
// between two convergence checks with candidate t % kNumChunks, each
// bracketed by timing events; each class then keeps the chunk with the
// least time per work unit (the smaller of its rounds).  A unit's arithmetic
// does not depend on the chunk, so the trial steps are ordinary steps with
// the same results.  (Running the trials again whenever the active slots'
// level-0 rows halved or doubled was tried on the 256-problem batch: 14.4 s
// against 14.0 s, the trial steps costing more than the later choices
// gained.)
enum { TC_PAIR = 0, TC_RES, TC_RESTOP, TC_RESTRICT, TC_INTERP, TC_ADD, TC_COUNT };
const int kChunks[] = {1, 2, 4, 8};
const int kNumChunks = 4;
const int kTuneSteps = 2 * kNumChunks;

struct TileTuner {
	int step = 0;           // trial steps done
	bool done = false;
	int best[TC_COUNT] = {1, 1, 1, 1, 1, 1};
	double ms[TC_COUNT][kNumChunks][2] = {};        // per candidate and round
	long long units[TC_COUNT][kNumChunks][2] = {};
	struct Mark { int cls, cand, round; long long units; hipEvent_t a, b; };
	std::vector<Mark> marks;    // this step's timed launches
	std::vector<hipEvent_t> pool;
	size_t used = 0;
};

}  // namespace

struct pgmg_batch {
	pgmg_opts opts;
	bool mixed = false;
	int max_slots = 0;
	hipStream_t stream = nullptr, setup = nullptr;
	Arena arena;
	std::vector<SlotHost> slots;
	SlotDev *d_slots = nullptr;
	int *d_plan = nullptr;
	std::vector<int> h_plan;
	int list_stride = 0;
	std::vector<int> list_n;  // slots per list
	bool plan_dirty = true;
	double *d_norms = nullptr, *h_norms = nullptr;
	double *d_aux = nullptr, *h_aux = nullptr;  // per slot: source norm, LU flag
	int *d_info = nullptr;
	int max_nl = 0;
	uint64_t admitted = 0;
	TileTuner tune;
	// the batched cycle as a hipGraph, captured once the tile shapes are
	// chosen and the plan has stayed the same for a step; dropped when the
	// plan changes
	hipGraphExec_t cycle_graph = nullptr;
	int plan_steps = 0;   // consecutive steps on the current plan
	long long graph_launches = 0;
};

namespace {

void drop_graph(pgmg_batch *b) {
	if (b->cycle_graph) { hipGraphExecDestroy(b->cycle_graph); b->cycle_graph = nullptr; }
	b->plan_steps = 0;
}

int list_index(int level, int kind) { return L_TOP_COUNT + level * L_PER_LEVEL + kind; }
int n_lists() { return L_TOP_COUNT + kBatchMaxLevels * L_PER_LEVEL; }

WorkList wlist(pgmg_batch *b, int li) {
	WorkList w;
	w.slot = b->d_plan + (size_t)li * b->list_stride;
	w.prefix = w.slot + b->max_slots;
	w.n = b->list_n[li];
	w.total = b->h_plan[(size_t)li * b->list_stride + b->max_slots + w.n];
	return w;
}

void plan_add(pgmg_batch *b, int li, int slot, int units) {
	int *base = &b->h_plan[(size_t)li * b->list_stride];
	int &n = b->list_n[li];
	base[n] = slot;
	int *prefix = base + b->max_slots;
	prefix[n + 1] = prefix[n] + units;
	n++;
}

int build_plan(pgmg_batch *b) {
	const int el = b->mixed ? 4 : 8;  // the cycle's arithmetic
	std::fill(b->list_n.begin(), b->list_n.end(), 0);
	for (int li = 0; li < n_lists(); li++) b->h_plan[(size_t)li * b->list_stride + b->max_slots] = 0;
	b->max_nl = 0;
	for (int s = 0; s < b->max_slots; s++) {
		SlotHost &S = b->slots[s];
		if (S.state != ST_ACTIVE) continue;
		const SlotDev &D = S.dev;
		b->max_nl = std::max(b->max_nl, S.nl);
		plan_add(b, L_TOP_ROWS, s, batch_units(BK_ROWS, D.lv[0], 8));
		plan_add(b, L_TOP_ONE, s, 1);
		plan_add(b, L_TOP_PTS, s, batch_units(BK_FLAT, D.lv[0], 8));
		for (int l = 0; l < S.nl; l++) {
			const LevelView &L = D.lv[l];
			if (l < S.nl - 1) {
				// as the lone hierarchy's forward colour sweeps
				if (batch_level_small(L, el))
					plan_add(b, list_index(l, L_SMALL), s, 1);
				else for (int q = 0; q < 4; q++) plan_add(b, list_index(l, L_PAIR0 + q), s, batch_units(BK_PAIR0 + q, L, el));
				plan_add(b, list_index(l, L_RES), s, batch_units(BK_ROWS, L, el));
				plan_add(b, list_index(l + 1, L_RESTRICT), s, batch_restrict_units(L, D.lv[l + 1]));
				plan_add(b, list_index(l + 1, L_INTERP), s, batch_units(BK_CUBES, L, el));
			} else if (D.lu_active) {
				plan_add(b, list_index(l, L_LU), s, 1);
			} else {
				plan_add(b, list_index(l, L_GS24), s, 1);
			}
		}
	}
	BHIP(hipMemcpyAsync(b->d_plan, b->h_plan.data(), sizeof(int) * b->h_plan.size(), hipMemcpyHostToDevice, b->stream));
	b->plan_dirty = false;
	drop_graph(b);  // its launches were sized for the old lists
	return PGMG_OK;
}

int chunk_of(pgmg_batch *b, int cls) {
	return b->tune.done ? b->tune.best[cls] : kChunks[b->tune.step % kNumChunks];
}

// Runs launch(chunk) for class cls; during the trial steps it is bracketed
// by timing events (units: the list's work units).
template <typename F>
void tuned(pgmg_batch *b, int cls, const WorkList &wl, F launch) {
	if (wl.n <= 0) return;
	const int c = chunk_of(b, cls);
	TileTuner &t = b->tune;
	if (t.done) { launch(c); return; }
	if (t.used + 2 > t.pool.size()) {
		for (int q = 0; q < 64; q++) {
			hipEvent_t e;
			if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) { launch(c); return; }
			t.pool.push_back(e);
		}
	}
	TileTuner::Mark mk{cls, t.step % kNumChunks, t.step / kNumChunks, wl.total, t.pool[t.used], t.pool[t.used + 1]};
	t.used += 2;
	hipEventRecord(mk.a, b->stream);
	launch(c);
	hipEventRecord(mk.b, b->stream);
	t.marks.push_back(mk);
}

// Once the stream is synchronised: books the timed launches under their
// candidates, moves to the next trial, and after the last one picks the
// chunks.
void tune_collect(pgmg_batch *b) {
	TileTuner &t = b->tune;
	if (t.done) return;
	for (const auto &mk : t.marks) {
		float ms = 0;
		if (hipEventElapsedTime(&ms, mk.a, mk.b) == hipSuccess) {
			t.ms[mk.cls][mk.cand][mk.round] += ms;
			t.units[mk.cls][mk.cand][mk.round] += mk.units;
		}
	}
	t.marks.clear();
	t.used = 0;
	if (++t.step < kTuneSteps) return;
	for (int c = 0; c < TC_COUNT; c++) {
		double best = -1;
		for (int q = 0; q < kNumChunks; q++) {
			double per = -1;
			for (int r = 0; r < 2; r++) {
				if (t.units[c][q][r] <= 0) continue;
				const double v = t.ms[c][q][r] / (double)t.units[c][q][r];
				if (per < 0 || v < per) per = v;
			}
			if (per >= 0 && (best < 0 || per < best)) { best = per; t.best[c] = kChunks[q]; }
		}
	}
	t.done = true;
}

// from_zero: level 0 of the mixed correction, whose pair-list slots did not
// clear x32 (top_zero); the first sweep reads none of it.
template <typename T>
void batched_gs(pgmg_batch *b, int l, int sweeps, bool from_zero = false) {
	if (sweeps <= 0) return;
	for (int s = 0; s < sweeps; s++)
		for (int q = 0; q < 4; q++) {
			const WorkList wl = wlist(b, list_index(l, L_PAIR0 + q));
			tuned(b, TC_PAIR, wl, [&](int c) {
				launch_b_gs_pair<T>(b->stream, b->d_slots, wl, l, q, from_zero && s == 0, c);
			});
		}
	launch_b_gs_small<T>(b->stream, b->d_slots, wlist(b, list_index(l, L_SMALL)), l, sweeps);
}

// One V-cycle (multigrid.cc:88-106) for every active slot, level by level.
template <typename T>
void batched_cycle(pgmg_batch *b) {
	const int L = b->max_nl;
	for (int l = 0; l < L; l++) {
		batched_gs<T>(b, l, b->opts.v_down + l, b->mixed && l == 0);
		const WorkList wr = wlist(b, list_index(l, L_RES));
		tuned(b, TC_RES, wr, [&](int c) { launch_b_residual<T>(b->stream, b->d_slots, wr, l, c); });
		if (l + 1 < kBatchMaxLevels) {
			const WorkList wt = wlist(b, list_index(l + 1, L_RESTRICT));
			tuned(b, TC_RESTRICT, wt, [&](int c) { launch_b_restrict<T>(b->stream, b->d_slots, wt, l + 1, c); });
		}
		launch_b_lu_solve<T>(b->stream, b->d_slots, wlist(b, list_index(l, L_LU)), l);
		launch_b_gs_small<T>(b->stream, b->d_slots, wlist(b, list_index(l, L_GS24)), l, kGs24);
	}
	for (int l = L - 1; l > 0; l--) {
		const WorkList wi = wlist(b, list_index(l, L_INTERP));
		tuned(b, TC_INTERP, wi, [&](int c) { launch_b_interp<T>(b->stream, b->d_slots, wi, l, c); });
		batched_gs<T>(b, l - 1, b->opts.v_up + l - 1);
	}
	if (b->mixed) {
		// x += (double) x32 on level 0
		const WorkList wa = wlist(b, L_TOP_PTS);
		tuned(b, TC_ADD, wa, [&](int c) { launch_b_add_x(b->stream, b->d_slots, wa, c); });
	}
}

// The cycle of a step (and in mixed precision x += x32): eager during the
// tile trials and on a plan's first step, then as a hipGraph captured on
// the plan's second step and replayed until the plan changes (its launches
// hold the lists' sizes).  The graph holds kernel launches only.
int run_cycle(pgmg_batch *b) {
	auto eager = [&]() { if (b->mixed) batched_cycle<float>(b); else batched_cycle<double>(b); };
	if (!b->opts.use_graph || !b->tune.done) { eager(); return PGMG_OK; }
	if (!b->cycle_graph && ++b->plan_steps >= 2) {
		// A capture only records: if it cannot be made (the stream already
		// capturing, instantiation refused), the batch runs eagerly from
		// here on, this cycle included.
		hipGraph_t g = nullptr;
		hipGraphExec_t ge = nullptr;
		bool ok = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
		if (ok) {
			eager();
			ok = hipStreamEndCapture(b->stream, &g) == hipSuccess && g;
			ok = ok && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess;
			if (g) hipGraphDestroy(g);
		}
		if (!ok) {
			(void)hipGetLastError();
			b->opts.use_graph = 0;
			eager();
			return PGMG_OK;
		}
		b->cycle_graph = ge;
	}
	if (b->cycle_graph) {
		BHIP(hipGraphLaunch(b->cycle_graph, b->stream));
		b->graph_launches++;
	} else {
		eager();
	}
	return PGMG_OK;
}

uint64_t slot_bytes(const pgmg_batch *b, int nl, const int *dims) {
	uint64_t t = 0;
	for (int l = 0; l < nl; l++) {
		const LevelView v = make_level_view(dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]);
		t += al(27 * v.npad * 8) + 2 * al(v.npad * 8);
		if (b->mixed) t += al(27 * v.npad * 4) + 2 * al(v.npad * 4);
	}
	const LevelView top = make_level_view(dims[0], dims[1], dims[2]);
	t += b->mixed ? al(top.npad * 4) : al(top.npad * 8);
	t += al(sizeof(double) * (uint64_t)(residual_blocks(top, 8) + 64));
	if (b->opts.bottom == PGMG_BOTTOM_LU) {
		const uint64_t N = (uint64_t)dims[3 * (nl - 1)] * dims[3 * (nl - 1) + 1] * dims[3 * (nl - 1) + 2];
		t += al(N * N * 8) + al(N * 4);
	}
	return t;
}

// Carves the slot's arrays out of its arena block and queues its set-up on
// the set-up stream.  family == 0: A27/rhs are host arrays.
int slot_setup(pgmg_batch *b, int s, const pgmg_dims *d, int family, uint64_t seed, const double *A27, const double *rhs) {
	SlotHost &S = b->slots[s];
	SlotDev &D = S.dev;
	memset(&D, 0, sizeof D);
	D.nl = S.nl;
	// level 0 on the colour-pair kernels with at least one down-sweep: the
	// mixed correction's first sweep starts from zero without x32 being
	// cleared (with v_down = 0 the residual pass clears it)
	D.top_zero = b->mixed && S.nl > 1 && b->opts.v_down >= 1 &&
			     !batch_level_small(make_level_view(S.dims[0], S.dims[1], S.dims[2]), 4) ? 1 : 0;
	char *p = b->arena.base + S.off;
	auto take = [&](uint64_t bytes) { char *q = p; p += al(bytes); return (void *)q; };
	for (int l = 0; l < S.nl; l++) {
		D.lv[l] = make_level_view(S.dims[3 * l], S.dims[3 * l + 1], S.dims[3 * l + 2]);
		const uint64_t np = (uint64_t)D.lv[l].npad;
		D.A[l] = (double *)take(27 * np * 8);
		D.x[l] = (double *)take(np * 8);
		D.r[l] = (double *)take(np * 8);
		if (b->mixed) {
			D.A32[l] = (float *)take(27 * np * 4);
			D.x32[l] = (float *)take(np * 4);
			D.r32[l] = (float *)take(np * 4);
		}
	}
	const uint64_t np0 = (uint64_t)D.lv[0].npad;
	if (b->mixed) D.res32 = (float *)take(np0 * 4); else D.res64 = (double *)take(np0 * 8);
	D.partial = (double *)take(sizeof(double) * (uint64_t)(residual_blocks(D.lv[0], 8) + 64));
	const int bl = S.nl - 1;
	const long long N = (long long)D.lv[bl].m * D.lv[bl].n * D.lv[bl].o;
	if (S.want_lu) {
		D.LU = (double *)take((uint64_t)(N * N) * 8);
		D.ipiv = (int *)take((uint64_t)N * 4);
		D.lu_n = (int)N;
	}
	hipStream_t st = b->setup;
	launch_fill<double>(st, (double *)(b->arena.base + S.off), (long long)(S.bytes / 8), 0.0);
	if (family) {
		launch_generate(st, D.lv[0], family, seed, D.A[0], D.r[0]);
	} else {
		// host operator and source: staged through the level-0 x and r32
		// arrays is not possible for every size, so copy synchronously.
		const long long N0 = (long long)D.lv[0].m * D.lv[0].n * D.lv[0].o;
		double *stage = nullptr;
		BHIP(hipMalloc(&stage, sizeof(double) * 27 * (size_t)N0));
		BHIP(hipMemcpyAsync(stage, A27, sizeof(double) * 27 * N0, hipMemcpyHostToDevice, st));
		launch_pack_op(st, D.lv[0], stage, D.A[0], 0, N0);
		BHIP(hipMemcpyAsync(stage, rhs, sizeof(double) * N0, hipMemcpyHostToDevice, st));
		launch_pack_vec(st, D.lv[0], stage, D.r[0], 0, N0);
		BHIP(hipStreamSynchronize(st));
		hipFree(stage);
	}
	for (int l = 1; l < S.nl; l++) launch_rat(st, D.lv[l - 1], D.lv[l], D.A[l - 1], D.A[l]);
	if (S.want_lu) {
		launch_lu_assemble(st, D.lv[bl], D.A[bl], D.LU);
		launch_lu_factor(st, (int)N, D.LU, D.ipiv, b->d_info + 2 * s);
	}
	if (b->mixed)
		for (int l = 0; l < S.nl; l++) launch_round_f32(st, D.A[l], D.A32[l], 27 * D.lv[l].npad);
	const int np_ = ((D.lv[0].m + 63) / 64) * ((D.lv[0].n + 3) / 4) * D.lv[0].o;
	double *tmp = nullptr;
	BHIP(hipMallocAsync((void **)&tmp, sizeof(double) * (size_t)np_, st));
	launch_sumsq<double>(st, D.lv[0], D.r[0], tmp);
	launch_reduce(st, tmp, np_, b->d_aux + 2 * s);
	BHIP(hipFreeAsync(tmp, st));
	BHIP(hipMemcpyAsync(b->h_aux + 2 * s, b->d_aux + 2 * s, sizeof(double), hipMemcpyDeviceToHost, st));
	if (S.want_lu) BHIP(hipMemcpyAsync((int *)(b->h_aux + 2 * s + 1), b->d_info + 2 * s, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
	BHIP(hipMemcpyAsync(b->d_slots + s, &S.dev, sizeof(SlotDev), hipMemcpyHostToDevice, st));
	BHIP(hipEventRecord(S.ready, st));
	S.state = ST_SETUP;
	return PGMG_OK;
}

// Slots whose set-up finished become active.
int activate_ready(pgmg_batch *b) {
	for (int s = 0; s < b->max_slots; s++) {
		SlotHost &S = b->slots[s];
		if (S.state != ST_SETUP) continue;
		hipError_t q = hipEventQuery(S.ready);
		if (q == hipErrorNotReady) continue;
		if (q != hipSuccess) return bfail(PGMG_ERR_HIP, "slot set-up: %s", hipGetErrorString(q));
		S.bn = b->h_aux[2 * s];
		if (S.want_lu) {
			const int *info = (const int *)(b->h_aux + 2 * s + 1);
			S.dev.lu_active = info[1] == 0;
			BHIP(hipMemcpyAsync(b->d_slots + s, &S.dev, sizeof(SlotDev), hipMemcpyHostToDevice, b->stream));
		}
		S.state = ST_ACTIVE;
		S.cycles = 0;
		b->plan_dirty = true;
	}
	return PGMG_OK;
}

int admit_common(pgmg_batch *b, const pgmg_dims *d, int family, uint64_t seed, const double *A27, const double *rhs,
		 double tol, int max_cycles, int *slot) {
	if (!b || !d || !slot) return bfail(PGMG_ERR_ARG, "null argument");
	if (d->x_prd || d->y_prd || d->z_prd) return bfail(PGMG_ERR_UNSUPPORTED, "periodic axes are not supported");
	if (d->m < 1 || d->n < 1 || d->o < 1) return bfail(PGMG_ERR_ARG, "grid sizes must be positive");
	int dims[3 * kBatchMaxLevels];
	const int nl = shape_dims(d->m, d->n, d->o, dims);
	if (nl < 0) return bfail(PGMG_ERR_MULTIGRID, "Maximum number of multigrid levels reached");
	int s = -1;
	for (int q = 0; q < b->max_slots; q++) if (b->slots[q].state == ST_FREE) { s = q; break; }
	if (s < 0) return bfail(PGMG_ERR_ADMISSION, "no free slot");
	SlotHost &S = b->slots[s];
	S.nl = nl;
	memcpy(S.dims, dims, sizeof(int) * 3 * nl);
	const long long Nb = (long long)dims[3 * (nl - 1)] * dims[3 * (nl - 1) + 1] * dims[3 * (nl - 1) + 2];
	S.want_lu = b->opts.bottom == PGMG_BOTTOM_LU && Nb < 2048;
	S.bytes = slot_bytes(b, nl, dims);
	if (!b->arena.alloc(S.bytes, S.off)) return bfail(PGMG_ERR_ADMISSION, "arena full (%llu bytes needed)", (unsigned long long)S.bytes);
	S.tol = tol;
	S.max_cycles = max_cycles;
	S.cycles = 0;
	S.rel = 1;
	S.seq = ++b->admitted;
	int rc = slot_setup(b, s, d, family, seed, A27, rhs);
	if (rc) { b->arena.release(S.off, S.bytes); S.state = ST_FREE; return rc; }
	*slot = s;
	return PGMG_OK;
}

}  // namespace

extern "C" {

int pgmg_batch_create(const pgmg_opts *opts, uint64_t arena_bytes, int max_slots, pgmg_batch **out) {
	if (!out) return bfail(PGMG_ERR_ARG, "null output");
	*out = nullptr;
	if (max_slots < 1 || max_slots > 4096) return bfail(PGMG_ERR_ARG, "max_slots out of range");
	pgmg_batch *b = new pgmg_batch();
	if (opts) b->opts = *opts; else pgmg_default_opts(&b->opts);
	if (b->opts.smoother != PGMG_SMOOTH_COLOUR8) { delete b; return bfail(PGMG_ERR_UNSUPPORTED, "batches use the COLOUR8 smoother"); }
	b->mixed = b->opts.precision == PGMG_PREC_MIXED;
	b->max_slots = max_slots;
	b->slots.resize(max_slots);
	b->list_stride = 2 * max_slots + 2;
	b->h_plan.assign((size_t)n_lists() * b->list_stride, 0);
	b->list_n.assign(n_lists(), 0);
	int rc = PGMG_OK;
	char *base = nullptr;
	if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
	    hipStreamCreateWithFlags(&b->setup, hipStreamNonBlocking) != hipSuccess) rc = bfail(PGMG_ERR_HIP, "stream creation failed");
	if (!rc && hipMalloc((void **)&base, arena_bytes) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "arena of %llu bytes", (unsigned long long)arena_bytes);
	if (!rc) b->arena.init(base, arena_bytes);
	if (!rc && hipMalloc((void **)&b->d_slots, sizeof(SlotDev) * max_slots) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "slot table");
	if (!rc && hipMalloc((void **)&b->d_plan, sizeof(int) * b->h_plan.size()) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "plan");
	if (!rc && hipMalloc((void **)&b->d_norms, sizeof(double) * max_slots) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "norms");
	if (!rc && hipMalloc((void **)&b->d_aux, sizeof(double) * 2 * max_slots) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "aux");
	if (!rc && hipMalloc((void **)&b->d_info, sizeof(int) * 2 * max_slots) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "info");
	if (!rc && hipHostMalloc((void **)&b->h_norms, sizeof(double) * max_slots) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "pinned");
	if (!rc && hipHostMalloc((void **)&b->h_aux, sizeof(double) * 2 * max_slots) != hipSuccess) rc = bfail(PGMG_ERR_MEMORY, "pinned");
	for (int s = 0; !rc && s < max_slots; s++)
		if (hipEventCreateWithFlags(&b->slots[s].ready, hipEventDisableTiming) != hipSuccess) rc = bfail(PGMG_ERR_HIP, "event");
	if (rc) { pgmg_batch_destroy(b); return rc; }
	*out = b;
	return PGMG_OK;
}

int pgmg_batch_destroy(pgmg_batch *b) {
	if (!b) return PGMG_OK;
	if (b->stream) hipStreamSynchronize(b->stream);
	if (b->setup) hipStreamSynchronize(b->setup);
	for (auto &S : b->slots) if (S.ready) hipEventDestroy(S.ready);
	drop_graph(b);
	for (auto e : b->tune.pool) hipEventDestroy(e);
	hipFree(b->arena.base);
	hipFree(b->d_slots); hipFree(b->d_plan); hipFree(b->d_norms); hipFree(b->d_aux); hipFree(b->d_info);
	if (b->h_norms) hipHostFree(b->h_norms);
	if (b->h_aux) hipHostFree(b->h_aux);
	if (b->stream) hipStreamDestroy(b->stream);
	if (b->setup) hipStreamDestroy(b->setup);
	delete b;
	return PGMG_OK;
}

int pgmg_batch_admit(pgmg_batch *b, const pgmg_dims *d, int family, uint64_t seed, double tol, int max_cycles, int *slot) {
	if (family != PGMG_FAM_FACE7_UNIFORM && family != PGMG_FAM_FACE7_LOGUNI && family != PGMG_FAM_Q1_ANISO)
		return bfail(PGMG_ERR_ARG, "unknown problem family %d", family);
	return admit_common(b, d, family, seed, nullptr, nullptr, tol, max_cycles, slot);
}

int pgmg_batch_admit_operator(pgmg_batch *b, const pgmg_dims *d, const double *A27, const double *rhs, double tol,
			      int max_cycles, int *slot) {
	if (!A27 || !rhs) return bfail(PGMG_ERR_ARG, "null array");
	return admit_common(b, d, 0, 0, A27, rhs, tol, max_cycles, slot);
}

int pgmg_batch_step(pgmg_batch *b, int *active, int *finished) {
	if (!b) return bfail(PGMG_ERR_ARG, "null batch");
	int rc = activate_ready(b);
	if (rc) return rc;
	{
		// Nothing to cycle but set-ups in flight: wait for the oldest one
		// (the set-up stream runs them in admission order).
		bool any_active = false;
		int oldest = -1;
		for (int s = 0; s < b->max_slots; s++) {
			const SlotHost &S = b->slots[s];
			if (S.state == ST_ACTIVE) any_active = true;
			if (S.state == ST_SETUP && (oldest < 0 || S.seq < b->slots[oldest].seq)) oldest = s;
		}
		if (!any_active && oldest >= 0) {
			BHIP(hipEventSynchronize(b->slots[oldest].ready));
			rc = activate_ready(b);
			if (rc) return rc;
		}
	}
	if (b->plan_dirty) { rc = build_plan(b); if (rc) return rc; }
	int nfin = 0, nact = 0;
	if (b->list_n[L_TOP_ROWS] > 0) {
		// residual of the current iterate: convergence test, and the FP32
		// defect for the mixed-precision cycle
		const WorkList wtop = wlist(b, L_TOP_ROWS);
		tuned(b, TC_RESTOP, wtop, [&](int c) { launch_b_residual_top(b->stream, b->d_slots, wtop, b->mixed, c); });
		launch_b_reduce(b->stream, b->d_slots, wlist(b, L_TOP_ONE), b->d_norms);
		BHIP(hipMemcpyAsync(b->h_norms, b->d_norms, sizeof(double) * b->max_slots, hipMemcpyDeviceToHost, b->stream));
		BHIP(hipStreamSynchronize(b->stream));
		// the previous step's cycle and this residual are complete
		tune_collect(b);
		for (int s = 0; s < b->max_slots; s++) {
			SlotHost &S = b->slots[s];
			if (S.state != ST_ACTIVE) continue;
			const double e = b->h_norms[s];
			S.rel = std::sqrt(e / S.bn);
			if (S.rel < S.tol) { S.state = ST_DONE; nfin++; }
			else if (std::isnan(e) || S.cycles >= S.max_cycles) { S.state = ST_FAILED; nfin++; }
		}
		if (nfin) { rc = build_plan(b); if (rc) return rc; }
		if (b->list_n[L_TOP_ROWS] > 0) {
			rc = run_cycle(b);
			if (rc) return rc;
			BHIP(hipGetLastError());
			for (auto &S : b->slots) if (S.state == ST_ACTIVE) { S.cycles++; nact++; }
		}
	}
	if (active) *active = nact;
	if (finished) *finished = nfin;
	return PGMG_OK;
}

int pgmg_batch_tiles(const pgmg_batch *b, int *chunks, int *tuned, long long *graph_launches) {
	if (!b) return bfail(PGMG_ERR_ARG, "null batch");
	if (chunks)
		for (int c = 0; c < TC_COUNT; c++) chunks[c] = b->tune.done ? b->tune.best[c] : 0;
	if (tuned) *tuned = b->tune.done ? 1 : 0;
	if (graph_launches) *graph_launches = b->graph_launches;
	return PGMG_OK;
}

int pgmg_batch_slot_status(pgmg_batch *b, int slot, int *state, int *cycles, double *rel) {
	if (!b || slot < 0 || slot >= b->max_slots) return bfail(PGMG_ERR_ARG, "bad slot");
	const SlotHost &S = b->slots[slot];
	const int map[] = {PGMG_SLOT_FREE, PGMG_SLOT_ACTIVE, PGMG_SLOT_ACTIVE, PGMG_SLOT_CONVERGED, PGMG_SLOT_FAILED};
	if (state) *state = map[S.state];
	if (cycles) *cycles = S.cycles;
	if (rel) *rel = S.rel;
	return PGMG_OK;
}

int pgmg_batch_get_x(pgmg_batch *b, int slot, double *x) {
	if (!b || slot < 0 || slot >= b->max_slots || !x) return bfail(PGMG_ERR_ARG, "bad argument");
	SlotHost &S = b->slots[slot];
	if (S.state == ST_FREE) return bfail(PGMG_ERR_ARG, "slot %d is free", slot);
	// a slot still in set-up waits for its own set-up only, not for later admissions
	if (S.state == ST_SETUP) BHIP(hipEventSynchronize(S.ready));
	const LevelView &L = S.dev.lv[0];
	const long long N = (long long)L.m * L.n * L.o;
	double *stage = nullptr;
	BHIP(hipMallocAsync((void **)&stage, sizeof(double) * (size_t)N, b->stream));
	launch_unpack_vec(b->stream, L, S.dev.x[0], stage, 0, N);
	BHIP(hipMemcpyAsync(x, stage, sizeof(double) * N, hipMemcpyDeviceToHost, b->stream));
	BHIP(hipFreeAsync(stage, b->stream));
	BHIP(hipStreamSynchronize(b->stream));
	return PGMG_OK;
}

int pgmg_batch_release(pgmg_batch *b, int slot) {
	if (!b || slot < 0 || slot >= b->max_slots) return bfail(PGMG_ERR_ARG, "bad slot");
	SlotHost &S = b->slots[slot];
	if (S.state == ST_FREE) return PGMG_OK;
	if (S.state == ST_SETUP || S.state == ST_ACTIVE) return bfail(PGMG_ERR_ARG, "slot %d is still running", slot);
	BHIP(hipStreamSynchronize(b->stream));
	b->arena.release(S.off, S.bytes);
	S.state = ST_FREE;
	return PGMG_OK;
}

int pgmg_batch_free_bytes(const pgmg_batch *b, uint64_t *bytes) {
	if (!b || !bytes) return bfail(PGMG_ERR_ARG, "null argument");
	*bytes = b->arena.cap - b->arena.used;
	return PGMG_OK;
}

}  // extern "C"

namespace {

// Result readout of the serving loop, off the cycling path: a worker thread
// copies a finished slot's solution to the host on its own stream while the
// main loop keeps stepping the batch.  A finished slot is in no plan, so no
// queued kernel touches its arrays; its arena block is released only after
// its copy has completed.
struct Readouts {
	struct Job {
		int slot;
		LevelView lv;
		const double *x;
		double *dst;  // the request's host array, or null (dropped)
		bool done = false;
		int rc = PGMG_OK;
		double t_done = 0;
	};
	std::mutex mu;
	std::condition_variable cv;
	std::deque<int> queue;  // indices into jobs
	std::vector<Job> jobs;
	bool stop = false;
	hipStream_t stream = nullptr;
	std::thread worker;
	std::vector<double> host;  // destination of results nobody asked for

	template <typename Clock>
	void run(Clock now) {
		for (;;) {
			int q;
			Job jb;
			{
				std::unique_lock<std::mutex> lk(mu);
				cv.wait(lk, [&] { return stop || !queue.empty(); });
				if (queue.empty()) return;
				q = queue.front();
				queue.pop_front();
				jb = jobs[q];
			}
			const long long N = (long long)jb.lv.m * jb.lv.n * jb.lv.o;
			double *dst = jb.dst;
			if (!dst) {
				host.resize((size_t)N);
				dst = host.data();
			}
			int rc = PGMG_OK;
			double *stage = nullptr;
			if (hipMallocAsync((void **)&stage, sizeof(double) * (size_t)N, stream) != hipSuccess) rc = PGMG_ERR_HIP;
			if (!rc) {
				launch_unpack_vec(stream, jb.lv, jb.x, stage, 0, N);
				if (hipMemcpyAsync(dst, stage, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost, stream) != hipSuccess)
					rc = PGMG_ERR_HIP;
				hipFreeAsync(stage, stream);
				if (hipStreamSynchronize(stream) != hipSuccess) rc = PGMG_ERR_HIP;
			}
			const double t = now();
			std::lock_guard<std::mutex> lk(mu);
			jobs[q].done = true;
			jobs[q].rc = rc;
			jobs[q].t_done = t;
		}
	}
};

}  // namespace

extern "C" {

int pgmg_serve_run(pgmg_batch *b, const pgmg_request *reqs, int n, pgmg_result *res, double *wall_s) {
	if (!b || (n > 0 && (!reqs || !res))) return bfail(PGMG_ERR_ARG, "null argument");
	typedef std::chrono::steady_clock clk;
	const auto t0 = clk::now();
	auto now = [t0]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
	std::vector<int> order(n);
	for (int q = 0; q < n; q++) { order[q] = q; res[q].admit_s = res[q].done_s = -1; res[q].cycles = 0; res[q].status = -1; res[q].rel_residual = 0; }
	std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return reqs[a].arrival_s < reqs[c].arrival_s; });
	std::vector<int> owner(b->max_slots, -1);     // request of each slot
	std::vector<int> reading(b->max_slots, -1);   // readout job of each slot
	Readouts ro;
	BHIP(hipStreamCreateWithFlags(&ro.stream, hipStreamNonBlocking));
	ro.jobs.reserve((size_t)n);
	ro.worker = std::thread([&ro, now]() { ro.run(now); });
	auto finish = [&](int rc) {
		{
			std::lock_guard<std::mutex> lk(ro.mu);
			ro.stop = true;
		}
		ro.cv.notify_all();
		ro.worker.join();
		hipStreamSynchronize(ro.stream);
		hipStreamDestroy(ro.stream);
		return rc;
	};
	int next = 0, done = 0;
	while (done < n) {
		const double t = now();
		// completed readouts: the result is on the host; the slot is free
		for (int s = 0; s < b->max_slots; s++) {
			if (reading[s] < 0) continue;
			bool fin;
			Readouts::Job jb;
			{
				std::lock_guard<std::mutex> lk(ro.mu);
				jb = ro.jobs[reading[s]];
				fin = jb.done;
			}
			if (!fin) continue;
			if (jb.rc) return finish(bfail(jb.rc, "result readout of slot %d failed", s));
			pgmg_result &R = res[owner[s]];
			R.done_s = jb.t_done;
			R.cycles = b->slots[s].cycles;
			R.rel_residual = b->slots[s].rel;
			R.status = b->slots[s].state == ST_DONE ? PGMG_OK : PGMG_ERR_NOT_CONVERGED;
			b->arena.release(b->slots[s].off, b->slots[s].bytes);
			b->slots[s].state = ST_FREE;
			reading[s] = owner[s] = -1;
			done++;
		}
		while (next < n && reqs[order[next]].arrival_s <= t) {
			const pgmg_request &R = reqs[order[next]];
			pgmg_dims d = {R.m, R.n, R.o, 0, 0, 0};
			int slot = -1;
			int rc = pgmg_batch_admit(b, &d, R.family, R.seed, R.tol, R.max_cycles, &slot);
			if (rc == PGMG_ERR_ADMISSION) {
				bool any = false;
				for (int q = 0; q < b->max_slots; q++) any = any || owner[q] >= 0;
				if (any) break;  // wait for capacity (FIFO)
				// the batch is empty and the request still does not fit: it
				// never will (larger than the arena, or too many levels)
			}
			if (rc) { res[order[next]].status = rc; res[order[next]].done_s = now(); done++; next++; continue; }
			owner[slot] = order[next];
			res[order[next]].admit_s = now();
			next++;
		}
		bool busy = false, cycling = false;
		for (int s = 0; s < b->max_slots; s++) {
			if (owner[s] < 0) continue;
			busy = true;
			if (reading[s] < 0) cycling = true;
		}
		if (!busy) {
			if (next < n) {
				const double wait = reqs[order[next]].arrival_s - now();
				if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
				continue;
			}
			break;
		}
		if (!cycling) {
			// only readouts in flight
			std::this_thread::sleep_for(std::chrono::microseconds(50));
			continue;
		}
		int act = 0, fin = 0;
		int rc = pgmg_batch_step(b, &act, &fin);
		if (rc) return finish(rc);
		for (int s = 0; s < b->max_slots; s++) {
			if (owner[s] < 0 || reading[s] >= 0) continue;
			const int st = b->slots[s].state;
			if (st != ST_DONE && st != ST_FAILED) continue;
			// the step synchronized after its residual: this slot's x is final
			// and the cycle just queued does not touch it
			Readouts::Job jb;
			jb.slot = s;
			jb.lv = b->slots[s].dev.lv[0];
			jb.x = b->slots[s].dev.x[0];
			jb.dst = reqs[owner[s]].x;
			std::lock_guard<std::mutex> lk(ro.mu);
			reading[s] = (int)ro.jobs.size();
			ro.jobs.push_back(jb);
			ro.queue.push_back(reading[s]);
			ro.cv.notify_one();
		}
	}
	if (wall_s) *wall_s = now();
	return finish(PGMG_OK);
}

}  // extern "C"
