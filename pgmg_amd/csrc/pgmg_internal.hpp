// pgmg_internal.hpp - device data layout and kernel launchers shared by the
// host driver (pgmg_hier.hip) and the kernels (pgmg_kernels.hip).
//
// Level layout in HBM.  Every per-point array of a level (the 27 coefficient
// planes, x, r, residual scratch) uses one padded index space: the grid plus
// a zero halo, rows padded to a multiple of 16 elements with room for
// i = -1 .. m+2.  Point (i,j,k),
// -1 <= i <= m etc., lives at base + i + px*j + pxy*k with base = 4 + px + pxy,
// so i = 0 of every row sits at a multiple of 4 elements: a pair (i, i+1),
// i even, moves with one 16-byte (FP64) or 8-byte (FP32) vector access, and
// a quad (i .. i+3), i a multiple of 4, with one 16-byte FP32 access.
// Halo entries are zero and stay zero, so a stencil at the grid edge reads
// zeros instead of branching; on a periodic axis the halo instead holds
// ghost copies of the wrapped points, refreshed (launch_halo_wrap) before
// each sweep and residual.  Coefficients: see arow0.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace pgmg {

struct LevelView {
	int m, n, o;       // grid points per axis
	int px;            // row pitch (elements)
	int pa;            // coefficient row pitch (elements): m rounded up to 4
	int tail;          // 1: rows of 4q+1 points (q >= 1), whose coefficient
	                   // blocks end with a copy of the last point's 27
	                   // entries (see arow0)
	long long pxy;     // plane pitch (elements)
	long long base;    // index of point (0,0,0)
	long long npad;    // elements per array, halo included
	int cache_coef;    // 1: levels run one at a time (a lone hierarchy), so a
	                   // level may keep coefficients in the memory-side cache;
	                   // 0: stream them (batched slots run side by side)
	int prd;           // periodic axes: bit 0 x, bit 1 y, bit 2 z
};

inline LevelView make_level_view(int m, int n, int o) {
	LevelView v;
	v.m = m; v.n = n; v.o = o;
	v.px = ((m + 7 + 15) / 16) * 16;  // i = -1 .. m+2 fit: a quad at i0 <= m-1 reads i0+3
	v.pa = ((m + 3) / 4) * 4;
	v.tail = (m >= 5 && (m & 3) == 1) ? 1 : 0;
	v.pxy = (long long)v.px * (n + 2);
	v.base = 4 + v.px + v.pxy;
	v.npad = v.pxy * (o + 2);
	v.cache_coef = 0;
	v.prd = 0;
	return v;
}

// Plane-skewed copy of a level for the exact smoother.  Within a k-plane,
// row u = i + 2j holds the points of that u (halo included: i = -1 .. m,
// j = -1 .. n) in order of j, from j = jf(u): the smallest j whose i is at
// most m when that makes rows shorter (cmp = 1; at most (m+1)/2 + 2
// positions), else -1 (rows of n + 2).  Point (i,j,k) is at
//   (u + 3)*R + (j - jf(u)) + pk*(k + 1),
// so the points of one hyperplane i + 2j + 4k = t with the same k lie side
// by side in j, and so do each of their neighbours.  Positions that are not
// grid points hold zero.  Coefficients: see sarow.
struct SkewView {
	int m, n, o;
	int R;             // row length
	int cmp;           // 1: rows start at jf(u); 0: at j = -1
	long long pk;      // k pitch: R * (m + 2n + 4) (u = -3 .. m + 2n)
	long long size;    // elements per array, halo included
};
inline SkewView make_skew_view(const LevelView &L) {
	SkewView v;
	v.m = L.m; v.n = L.n; v.o = L.o;
	const int rc = (L.m + 1) / 2 + 2, rf = L.n + 2;
	v.cmp = rc < rf ? 1 : 0;
	v.R = v.cmp ? rc : rf;
	v.pk = (long long)v.R * (L.m + 2 * L.n + 4);
	v.size = v.pk * (L.o + 2);
	return v;
}
__host__ __device__ inline long long sidx3(const SkewView &S, int i, int j, int k) {
	const int u = i + 2 * j;
	const int jf = S.cmp ? max(-1, (u - S.m + 1) >> 1) : -1;  // ceil((u - m) / 2), at least -1
	return (long long)(u + 3) * S.R + (j - jf) + S.pk * (k + 1);
}
// The 27 coefficient planes of a skewed copy are interleaved by row: row
// (u, k) of plane e sits at ((row(u, k) * 27 + e) * R, so the coefficients of
// one row of points are one contiguous block of 27 R values (a few pages,
// not 27 rows a plane apart, which a sweep over every plane at once would
// spread over thousands of pages).
__host__ __device__ inline long long sarow(const SkewView &S, int u, int k) {
	return ((long long)(u + 3) + (long long)(S.m + 2 * S.n + 4) * (k + 1)) * 27 * S.R;
}
__host__ __device__ inline long long saidx3(const SkewView &S, int e, int i, int j, int k) {
	const int u = i + 2 * j;
	const int jf = S.cmp ? max(-1, (u - S.m + 1) >> 1) : -1;
	return sarow(S, u, k) + (long long)e * S.R + (j - jf);
}

// Point (i,j,k) of a per-point array.
__host__ __device__ inline long long idx3(const LevelView &L, int i, int j, int k) {
	return L.base + i + (long long)L.px * j + L.pxy * k;
}
// Coefficients are interleaved by grid row: the 27 planes' rows of grid row
// (j,k) are one contiguous block of 27*pa elements, plane e's row at offset
// e*pa in it, entry i at +i.  A wavefront that streams one grid row therefore
// reads one contiguous 27*pa block (one or two memory pages), not 27 rows a
// plane apart.  The coefficient pitch pa is m rounded up to 4, not the x
// rows' px: coefficients have no halo, and a block of 27 px-rows would be
// streamed whole (1.45x the bytes used at m = 33, 1.05x at 257; with pa
// 1.09x and 1.01x).  A vector load at i0 <= m-1 may run past m-1 into the
// next row of the block; those entries only feed sums that are discarded.
// On levels with rows of 4q+1 points (tail), the block ends with 28 more
// entries: a copy of the 27 coefficients of the row's last point i = m-1, in
// plane order, and a zero.  The row kernels give that point to one extra
// lane per row, which reads the copy with 16-byte loads (two 64-byte lines)
// instead of 27 single entries a plane row apart (27 lines); the copies are
// refreshed by every launcher that writes coefficients (launch_coef_tail).
// Row blocks are numbered as the x rows, halo rows included, so the array
// fits in 27 per-point arrays; the halo rows' blocks are unused.
__host__ __device__ inline long long ablock(const LevelView &L) {
	return 27LL * L.pa + (L.tail ? 28 : 0);
}
__host__ __device__ inline long long arow0(const LevelView &L, int j, int k) {  // plane 0, i = 0
	return ablock(L) * ((long long)(j + 1) + (long long)(L.n + 2) * (k + 1));
}
__host__ __device__ inline long long aidx3(const LevelView &L, int e, int i, int j, int k) {
	return arow0(L, j, k) + (long long)e * L.pa + i;
}

// Lane layout of the row kernels (smoother and residual).  A lane holds
// ppl consecutive points of a grid row: 4 in FP32 (16-byte loads), 2 in
// FP64.  A row of m = ppl*k + 1 points (the 2^k+1 sizes of the hierarchy)
// takes k lanes, the last of which also does the point out (ext); other rows
// take ceil(m/ppl) lanes.  When a row needs at most 64 lanes, one wavefront
// packs rpw = 64/lpr rows; longer rows are streamed in segments by one
// wavefront (rpw = 1, lpr > 64).
struct RowLanes {
	int ppl, lpr, rpw;
	bool ext;
};
__host__ __device__ inline RowLanes row_lanes(int m, int elem) {
	RowLanes w;
	w.ppl = elem == 4 ? 4 : 2;
	w.ext = m > w.ppl && (m - 1) % w.ppl == 0;
	w.lpr = w.ext ? (m - 1) / w.ppl : (m + w.ppl - 1) / w.ppl;
	w.rpw = w.lpr <= 64 ? 64 / w.lpr : 1;
	return w;
}
// Wavefronts that cover `rows` grid rows of m points.
__host__ __device__ inline int row_waves(int rows, int m, int elem) {
	const RowLanes w = row_lanes(m, elem);
	return (rows + w.rpw - 1) / w.rpw;
}

// Segments of a long row (more than 64 lanes' worth) in the residual: 128
// points each in FP64, 256 in FP32, the last one up to one more (its last
// lane's point out).  Each segment is a wavefront of its own: no update runs
// along the row, so the segments need no order, and one wavefront per row
// made two or more dependent round trips of loads.
__host__ __device__ inline int residual_segments(int m, int elem) {
	return 1 + (m - 2) / (elem == 4 ? 256 : 128);
}
// Workgroups (4 wavefronts each) of a residual launch over all rows of L in
// arithmetic of elem bytes; an FP64 residual has one partial sum per block.
__host__ __device__ inline int residual_blocks(const LevelView &L, int elem) {
	const int rows = L.n * L.o;
	const int waves = row_lanes(L.m, elem).lpr <= 64 ? row_waves(rows, L.m, elem) : rows * residual_segments(L.m, elem);
	return (waves + 3) / 4;
}

// ---- batches of independent hierarchies (one per slot) ----
const int kBatchMaxLevels = 16;
// Device-side description of one slot's hierarchy.  Pointers of a precision
// the batch does not use are null.
struct SlotDev {
	int nl;
	int lu_active, lu_n;
	int top_zero;  // mixed: the first level-0 sweep starts from x32 = 0 without reading it
	LevelView lv[kBatchMaxLevels];
	double *A[kBatchMaxLevels], *x[kBatchMaxLevels], *r[kBatchMaxLevels];
	float *A32[kBatchMaxLevels], *x32[kBatchMaxLevels], *r32[kBatchMaxLevels];
	double *res64;     // residual scratch, top-level size
	float *res32;
	double *partial;   // per-block sums of the top-level residual
	double *LU;
	int *ipiv;
};
// A list of slots and the cumulative number of work units of a batched
// kernel over them: slot[q] owns units [prefix[q], prefix[q+1]).
struct WorkList {
	const int *slot;
	const int *prefix;
	int n;
	int total;  // prefix[n], known to the host: the launch's grid size
};
// Unit kinds: the colour-pair row groups of a level, one unit per slot,
// residual row groups, 64x4-point blocks of a level, 64x4 blocks of its
// 2x2x2 cubes.
enum { BK_PAIR0 = 0, BK_PAIR1, BK_PAIR2, BK_PAIR3, BK_ONE, BK_ROWS, BK_PTS, BK_CUBES, BK_FLAT };
// Work units of one slot for a batched kernel kind at a level; elem is the
// size of the cycle's arithmetic type (4 or 8).
int batch_units(int kind, const LevelView &L, int elem);
// Work units of a batched restriction onto C from F (interior tiles, then
// shell blocks).
int batch_restrict_units(const LevelView &F, const LevelView &C);

// Batched launchers: one launch covers the level `level` of every slot in
// the list; blocks walk the units grid-stride.  chunk: work units per
// workgroup (the grid is ceil(units / chunk)); a unit's arithmetic does not
// depend on it, so every chunk gives the same results.
// from_zero (FP32): the first sweep of level 0 of a mixed correction (slots
// with top_zero set; the others are in the small-level list).
template <typename T> void launch_b_gs_pair(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int pair,
					    bool from_zero = false, int chunk = 1);
template <typename T> void launch_b_gs_small(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int nsweeps);
template <typename T> void launch_b_residual(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int chunk = 1);
void launch_b_residual_top(hipStream_t s, const SlotDev *slots, WorkList wl, bool to_f32, int chunk = 1);
void launch_b_reduce(hipStream_t s, const SlotDev *slots, WorkList wl, double *norms);
template <typename T> void launch_b_restrict(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int chunk = 1);
template <typename T> void launch_b_interp(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int chunk = 1);
template <typename T> void launch_b_lu_solve(hipStream_t s, const SlotDev *slots, WorkList wl, int level);
void launch_b_add_x(hipStream_t s, const SlotDev *slots, WorkList wl, int chunk = 1);

// Records the calling thread's last error message (pgmg_last_error) and
// returns code.
int set_error(int code, const char *msg);

// LDS bytes the small-level smoother needs to stage x with its halo, and the
// most it uses (within every HIP device's limit for dynamic LDS).
__host__ __device__ inline long long small_lds_bytes(const LevelView &L, int elem) {
	return (long long)(L.m + 2) * (L.n + 2) * (L.o + 2) * elem;
}
const int kSmallLdsBytes = 64 * 1024;
// Point pairs (i0, i0+1), i0 even, of one colour pair (0,cj,ck), (1,cj,ck)
// at most (cj = ck = 0).
__host__ __device__ inline long long small_pairs_per_pass(const LevelView &L) {
	return (long long)((L.m + 1) / 2) * ((L.n + 1) / 2) * ((L.o + 1) / 2);
}
// A batched level goes to the one-workgroup small-level kernel when each
// thread holds one point (coefficients in registers), else to the colour-pair
// kernels, which spread a level over every CU (the pair-staged one-workgroup
// form is bound by its CU's L2 bandwidth).
inline bool batch_level_small(const LevelView &L, int elem) {
	return small_lds_bytes(L, elem) <= kSmallLdsBytes && (long long)L.m * L.n * L.o <= 1024;
}

// Kernel classes for timing.
enum { KC_NONE = 0, KC_GS = 1, KC_RESIDUAL = 2, KC_RESTRICT = 3, KC_INTERP = 4 };

// ---- launchers (pgmg_kernels.hip).  All are asynchronous on `s`. ----
// Gauss-Seidel, one colour class (ci,cj,ck) of one level.
template <typename T> void launch_gs_colour(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int colour);
// Forward Gauss-Seidel of the two colour classes (0,cj,ck) then (1,cj,ck)
// in one launch, one wavefront per grid row.
// from_zero (FP32 only): the first sweep of a call that starts from x = 0;
// x need not be cleared beforehand.
template <typename T> void launch_gs_pair(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int pair,
					  bool from_zero = false);
// Ghost copies of the wrapped points into the halo of the periodic axes of
// L (no-op without periodic axes).
template <typename T> void launch_halo_wrap(hipStream_t s, const LevelView &L, T *x);
// Gauss-Seidel on the hyperplane i+2j+4k = t.
template <typename T> void launch_gs_plane(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int t);
// Plane-skewed copies for the exact smoother: the 27 coefficient planes
// (As[saidx3(S, e, i, j, k)]) and one vector, both ways.
template <typename T> void launch_skew_pack_op(hipStream_t s, const LevelView &L, const SkewView &S, const T *A, T *As);
template <typename T> void launch_skew_vec(hipStream_t s, const LevelView &L, const SkewView &S, const T *v, T *vs,
					   bool to_skew);
// Ghost copies of the periodic axes of L on a skewed vector.
template <typename T> void launch_skew_halo_wrap(hipStream_t s, const LevelView &L, const SkewView &S, T *xs);
// nsweeps forward lexicographic sweeps on skewed copies, pipelined eight
// hyperplanes apart (bit-identical to the sweeps on the level layout).
template <typename T> void launch_gs_planes_skew(hipStream_t s, const SkewView &S, const T *As, T *xs, const T *rs,
						 int nsweeps);
// nsweeps forward lexicographic sweeps on skewed copies in one launch per
// (up to flow_max_sweeps) sweeps: each workgroup owns a plane and waits for
// the progress counters of its neighbour planes (bit-identical to separate
// sweeps).  prog: o + 1 ints, zero on entry and left zero; err: set when a
// wait gave up after `limit` polls (the sweep is then incomplete).  Returns
// false, launching nothing, when the level cannot take the kernel (all its
// planes' workgroups would not be resident at once).  prd: the level has
// periodic axes (its halo holds copies of x).
template <typename T> bool launch_gs_flow(hipStream_t s, const SkewView &S, bool prd, const T *As, T *xs,
					  const T *rs, int nsweeps, int *prog, int *err, int limit);
template <typename T> int flow_max_sweeps(const SkewView &S, bool prd);
// nsweeps forward lexicographic sweeps with the sweeps pipelined eight
// hyperplanes apart (bit-identical to nsweeps separate sweeps).
template <typename T> void launch_gs_planes(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int nsweeps);
// nsweeps Gauss-Seidel sweeps of a small level by one workgroup, with x
// staged in LDS when small_lds_bytes(L, sizeof(T)) <= kSmallLdsBytes.
template <typename T> void launch_gs_small(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r,
					   int nsweeps, int mode, int exact);
// res = r - A x on interior points (res may be null), per-block sums of
// res^2 into partial (may be null); zero_out (may be null, FP64 arithmetic
// only) is set to 0 at interior points in the same pass; returns the number
// of partial sums.
template <typename T, typename TR> int launch_residual(hipStream_t s, const LevelView &L, const T *A, const T *x,
						       const T *r, TR *res, double *partial, TR *zero_out = nullptr);
// Partial sums (workgroups) of an FP64 residual launch over level L.
int residual_partial_count(const LevelView &L);
// out = A in at interior points (in with a zero halo).
void launch_apply(hipStream_t s, const LevelView &L, const double *A, const double *in, double *out);
// Block sums of a*b (diff = 0) or (a-b)^2 (diff = 1); returns the count.
int launch_pair_sum(hipStream_t s, const LevelView &L, const double *a, const double *b, int diff, double *partial);
// The block sums of launch_pair_sum(diff = 1) with `ext` point-major
// (i + m*(j + n*k)) rather than in the level layout; returns the count.
int launch_diff2_compact(hipStream_t s, const LevelView &L, const double *ext, const double *x, double *partial);
// out[0] = sum of partial[0..n) in a fixed order.
void launch_reduce(hipStream_t s, const double *partial, int n, double *out);
// partial sums of r^2 (rhs_l2norm); returns count.
template <typename T> int launch_sumsq(hipStream_t s, const LevelView &L, const T *r, double *partial);
// rc = R res (gather form of the reference's scatter), xc = 0.
template <typename T> void launch_restrict(hipStream_t s, const LevelView &F, const LevelView &C, const T *res, T *rc, T *xc);
// xf += P xc.
template <typename T> void launch_interp(hipStream_t s, const LevelView &C, const LevelView &F, const T *xc, T *xf);
// Ac = P^T Af P / 8.
void launch_rat(hipStream_t s, const LevelView &F, const LevelView &C, const double *Af, double *Ac);
// Dense bottom LU.
void launch_lu_assemble(hipStream_t s, const LevelView &L, const double *A, double *LU);
void launch_lu_factor(hipStream_t s, int N, double *LU, int *ipiv, int *info);
template <typename T> void launch_lu_solve(hipStream_t s, const LevelView &L, int N, const double *LU, const int *ipiv,
					   const T *r, T *x);
// Problem generation on the device.
void launch_generate(hipStream_t s, const LevelView &L, int family, uint64_t seed, double *A, double *r);
// Layout conversion between host point-major arrays (staged on the device)
// and the padded level layout, for points [p0, p0+count).
void launch_pack_vec(hipStream_t s, const LevelView &L, const double *compact, double *padded, long long p0, long long count);
void launch_unpack_vec(hipStream_t s, const LevelView &L, const double *padded, double *compact, long long p0, long long count);
void launch_pack_op(hipStream_t s, const LevelView &L, const double *compact27, double *A, long long p0, long long count);
void launch_unpack_op(hipStream_t s, const LevelView &L, const double *A, double *compact27, long long p0, long long count);
// Elementwise helpers.
template <typename T> void launch_fill(hipStream_t s, T *p, long long n, T v);
void launch_round_f32(hipStream_t s, const double *src, float *dst, long long n);
void launch_add_f32_to_f64(hipStream_t s, const LevelView &L, const float *e, double *x);

}  // namespace pgmg
