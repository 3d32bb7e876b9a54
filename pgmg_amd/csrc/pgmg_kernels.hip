// pgmg_kernels.hip - gfx950 kernels of the multigrid hot path.
//
// Built with -ffp-contract=off: every a*b+c rounds twice, as the reference's
// x86-64 build does, so FP64 results reproduce the reference's bit for bit
// where the operation order is the same.  All kernels are memory-bound
// (0.225 flop/byte in FP64); none is a dense contraction, so no MFMA is used.
#include "pgmg_internal.hpp"
#include <utility>
#include "../../include/pgmg_problems.h"

namespace pgmg {

#define HIPCHECK_LAUNCH() (void)hipGetLastError()

// The streaming stencil kernels leave occupancy to the compiler: a row
// kernel issues every load of a row segment at once (about 210 VGPRs in
// FP32, 240 in FP64: two wavefronts per SIMD), so its latency is covered
// within the wavefront; forcing more wavefronts per SIMD only made the
// compiler spill.


// The hardware deals workgroups to the 8 XCDs round-robin (workgroup b runs
// on XCD b % 8), and each XCD has its own L2.  Kernels whose neighbouring
// workgroups share input rows renumber them so that every XCD receives one
// contiguous range of logical workgroups: x rows shared by adjacent grid rows
// are then fetched into one L2 instead of eight.  A bijection of [0, G).
__device__ __forceinline__ int xcd_block(int b, int G) {
	const int q = G >> 3, rmd = G & 7, x = b & 7;
	return x * q + min(x, rmd) + (b >> 3);
}

// Gauss-Seidel update of point (i,j,k) (region.cc:369-438): sum over the
// 26 off-centre entries in box order from zero, then
// x = (r - Ax) * (1/a_centre) with the reciprocal taken first.
template <typename T>
__device__ __forceinline__ void gs_point(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
					 const T *__restrict__ r, int i, int j, int k) {
	const long long px = L.px, pxy = L.pxy, p = idx3(L, i, j, k);
	const T *Ap = A + aidx3(L, 0, i, j, k);
	T ax = (T)0;
#pragma unroll
	for (int e = 0; e < 27; e++) {
		if (e == 13) continue;
		const int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
		ax = ax + Ap[e * L.pa] * x[p + di + dj * px + dk * pxy];
	}
	x[p] = (r[p] - ax) * ((T)1 / Ap[13 * L.pa]);
}

// ------------------------------------------------------- periodic halo

// Ghost copies for the periodic axes of L.  At one rank the reference's
// periodic neighbour is the rank itself, and communicate()
// (region.cc:1338-1365) copies the grid's wrapped values into its ghost
// layers before each sweep and each residual (region.cc:361, 457, 549,
// 1478); a sweep then reads, but never updates, those copies.  One thread
// per halo point of the six faces (edges and corners come up more than once
// and get the same value); a halo point outside a non-periodic axis keeps
// its zero.  Every source is a grid point, never a halo point, so the
// copies do not depend on one another.
// Halo point q of the six faces of L: its indices, and those of the grid
// point it copies; false when it lies outside a non-periodic axis.
__device__ __forceinline__ bool halo_src(const LevelView &L, long long q, int &i, int &j, int &k, int &si, int &sj,
					 int &sk) {
	const long long ex = 2LL * (L.n + 2) * (L.o + 2), ey = 2LL * (L.m + 2) * (L.o + 2);
	if (q < ex) {
		const long long t = q >> 1;
		i = (q & 1) ? L.m : -1; j = (int)(t % (L.n + 2)) - 1; k = (int)(t / (L.n + 2)) - 1;
	} else if (q < ex + ey) {
		const long long t = (q - ex) >> 1;
		j = ((q - ex) & 1) ? L.n : -1; i = (int)(t % (L.m + 2)) - 1; k = (int)(t / (L.m + 2)) - 1;
	} else {
		const long long t = (q - ex - ey) >> 1;
		k = ((q - ex - ey) & 1) ? L.o : -1; i = (int)(t % (L.m + 2)) - 1; j = (int)(t / (L.m + 2)) - 1;
	}
	si = i; sj = j; sk = k;
	if (si < 0 || si >= L.m) { if (!(L.prd & 1)) return false; si = si < 0 ? si + L.m : si - L.m; }
	if (sj < 0 || sj >= L.n) { if (!(L.prd & 2)) return false; sj = sj < 0 ? sj + L.n : sj - L.n; }
	if (sk < 0 || sk >= L.o) { if (!(L.prd & 4)) return false; sk = sk < 0 ? sk + L.o : sk - L.o; }
	return true;
}

__host__ __device__ inline long long halo_points(const LevelView &L) {
	return 2LL * ((long long)(L.n + 2) * (L.o + 2) + (long long)(L.m + 2) * (L.o + 2) + (long long)(L.m + 2) * (L.n + 2));
}

template <typename T>
__global__ __launch_bounds__(256) void k_halo_wrap(LevelView L, T *__restrict__ x) {
	const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
	int i, j, k, si, sj, sk;
	if (q >= halo_points(L) || !halo_src(L, q, i, j, k, si, sj, sk)) return;
	x[idx3(L, i, j, k)] = x[idx3(L, si, sj, sk)];
}

// The same copies by one workgroup, between barriers, on an array whose
// point (i,j,k) is at org + i + hx*j + hxy*k (a level in memory, or one staged
// in LDS).  No-op without periodic axes.
template <typename T>
__device__ __forceinline__ void halo_wrap_block(const LevelView &L, T *__restrict__ x, long long org, long long hx,
						long long hxy) {
	if (!L.prd) return;
	__syncthreads();
	const long long tot = halo_points(L);
	for (long long q = threadIdx.x; q < tot; q += blockDim.x) {
		int i, j, k, si, sj, sk;
		if (halo_src(L, q, i, j, k, si, sj, sk)) x[org + i + hx * j + hxy * k] = x[org + si + hx * sj + hxy * sk];
	}
	__syncthreads();
}

template <typename T>
void launch_halo_wrap(hipStream_t s, const LevelView &L, T *x) {
	if (!L.prd) return;
	const long long tot = halo_points(L);
	hipLaunchKernelGGL(k_halo_wrap<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, L, x);
	HIPCHECK_LAUNCH();
}

// ------------------------------------------------------------ smoothers

// One colour class: points (2a+ci, 2b+cj, 2d+ck).  Points of one class never
// share a stencil box, so all of them update in parallel.
template <typename T>
__global__ __launch_bounds__(256) void k_gs_colour(LevelView L, const T *__restrict__ A, T *__restrict__ x,
						   const T *__restrict__ r, int ci, int cj, int ck) {
	const int i = 2 * (int)(blockIdx.x * 64 + threadIdx.x) + ci;
	const int j = 2 * (int)(blockIdx.y * 4 + threadIdx.y) + cj;
	const int k = 2 * (int)blockIdx.z + ck;
	if (i >= L.m || j >= L.n || k >= L.o) return;
	gs_point<T>(L, A, x, r, i, j, k);
}

template <typename T>
void launch_gs_colour(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int colour) {
	const int ci = colour & 1, cj = (colour >> 1) & 1, ck = (colour >> 2) & 1;
	const int na = (L.m - ci + 1) / 2, nb = (L.n - cj + 1) / 2, nd = (L.o - ck + 1) / 2;
	if (na <= 0 || nb <= 0 || nd <= 0) return;
	dim3 grid((na + 63) / 64, (nb + 3) / 4, nd), block(64, 4, 1);
	hipLaunchKernelGGL(k_gs_colour<T>, grid, block, 0, s, L, A, x, r, ci, cj, ck);
	HIPCHECK_LAUNCH();
}

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

// Row-streaming kernels address every array of a grid row as a
// wavefront-uniform base (the row's start in that array, held in scalar
// registers) plus one per-lane byte offset below the row pitch, so the
// compiler emits the scalar-base form of global loads: one 32-bit offset
// register serves all 27 planes and the x rows, instead of a 64-bit
// address per plane.
template <typename V>
__device__ __forceinline__ V ld_row(const void *base, uint32_t off) {
	return *(const V *)((const char *)base + off);
}
// Raw buffer descriptor word 3 for gfx950 (no format conversion, no
// swizzle) and the cache-policy bits of buffer stores (bit 4 = sc1).
constexpr int kBufferWord3 = 0x00020000;
constexpr int kStoreSc1 = 16;
template <typename V>
__device__ __forceinline__ void st_row(void *base, uint32_t off, V v) {
	// Row results go out as buffer stores with the device-scope (sc1) cache
	// policy: in these read-dominated kernels each written byte costs
	// several times a read byte of HBM time, and this policy measured the
	// cheapest (level-0 FP32 smoother -2 %, FP64 -2 %).  The descriptor spans
	// 2 GiB from the row base; row offsets are far below that.
	typedef unsigned u4v __attribute__((ext_vector_type(4)));
	typedef unsigned u2v __attribute__((ext_vector_type(2)));
	const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, 0x7fffffff, kBufferWord3);
	if constexpr (sizeof(V) == 16) {
		u4v t; __builtin_memcpy(&t, &v, 16); __builtin_amdgcn_raw_buffer_store_b128(t, rs, (int)off, 0, kStoreSc1);
	} else if constexpr (sizeof(V) == 8) {
		u2v t; __builtin_memcpy(&t, &v, 8); __builtin_amdgcn_raw_buffer_store_b64(t, rs, (int)off, 0, kStoreSc1);
	} else {
		static_assert(sizeof(V) == 4, "row stores move 4, 8 or 16 bytes");
		unsigned t; __builtin_memcpy(&t, &v, 4); __builtin_amdgcn_raw_buffer_store_b32(t, rs, (int)off, 0, kStoreSc1);
	}
}

// Coefficient planes and r are read once per launch: their loads carry the
// non-temporal hint, so the streamed planes do not evict from L2 the x rows
// that neighbouring grid rows read again (measured on the level-0 smoother:
// FP32 113.8 -> 104.8 us per launch; x loads stay plain).
typedef double d2_nt __attribute__((ext_vector_type(2)));
template <typename V> __device__ __forceinline__ V ld_stream(const void *base, uint32_t off);
template <> __device__ __forceinline__ float2 ld_stream<float2>(const void *base, uint32_t off) {
	return *(const float2 *)((const char *)base + off);
}
typedef float f4_nt __attribute__((ext_vector_type(4)));
template <> __device__ __forceinline__ float4 ld_stream<float4>(const void *base, uint32_t off) {
	const f4_nt v = __builtin_nontemporal_load((const f4_nt *)((const char *)base + off));
	return make_float4(v.x, v.y, v.z, v.w);
}
template <> __device__ __forceinline__ double2 ld_stream<double2>(const void *base, uint32_t off) {
	const d2_nt v = __builtin_nontemporal_load((const d2_nt *)((const char *)base + off));
	return make_double2(v.x, v.y);
}

// First sweep from x = 0 (template flag Z of the FP32 row kernels): the rows of this colour pair and of the pairs
// after it have not been updated yet in the sweep, so they still hold 0;
// their x is not read (the caller need not clear x), and the terms see the
// same +0.0 a cleared array would give.
__device__ __forceinline__ bool zero_row(int cj, int ck, int dj, int dk) {
	return ((cj ^ (dj & 1)) + 2 * (ck ^ (dk & 1))) >= cj + 2 * ck;
}

// Colour classes (0,cj,ck) and then (1,cj,ck) in one launch.  One wavefront
// owns one grid row (j,k) with j = cj (mod 2), k = ck (mod 2); only that
// row is both read and written by the launch, because every other row a
// stencil reaches has the other parity in j or k.  The row is processed in
// segments of 128 points: lane l holds the pair (i0, i0+1), i0 = s0 + 2l,
// read with one vector load per coefficient plane.  Lanes update the even
// point i0, then the odd point i0+1, whose two even neighbours are the lane's
// own new value and lane l+1's (shuffle).  Lane 63's right even neighbour
// s0+128 (the first point of the next segment, or the last point of a row of
// 128q+1 points) is updated by lane 63 itself from scalar loads, with the
// evens; its value is handed to lane 0 of the next segment, which must not
// recompute it: by then its odd neighbour s0+127 has changed.  A row of
// 2^k+1 points thus takes 2^k/128 segments.  A segment's stores are issued
// after the next segment's loads (one counter tracks loads and stores in
// issue order, so a load issued after a store would wait for it); the only
// loads they overtake are the next segment's lane-0 reads of s0-1 and s0,
// which feed the even point that lane takes from the hand-over.  Per-point
// arithmetic and the colour order are those of k_gs_colour, so the result is
// bit-identical.  j and k must be wavefront-uniform.
//
// The point out's 27 coefficients and the x of the point after it in each of
// its 9 rows are loaded by lanes 0..26 and 0..8, one load each, and lane 63
// reads them from those lanes (v_readlane) instead of issuing 36 loads of its
// own: those doubled the load instructions of the segment (level-0 FP64
// smoother 216 -> 206 us).
template <typename T>
__device__ __forceinline__ T read_lane(T v, int l) {
	if constexpr (sizeof(T) == 8) {
		unsigned long long u;
		__builtin_memcpy(&u, &v, 8);
		const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffu), l);
		const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
		u = ((unsigned long long)hi << 32) | lo;
		T w;
		__builtin_memcpy(&w, &u, 8);
		return w;
	} else {
		unsigned u;
		__builtin_memcpy(&u, &v, 4);
		u = (unsigned)__builtin_amdgcn_readlane((int)u, l);
		T w;
		__builtin_memcpy(&w, &u, 4);
		return w;
	}
}
// ALL: every lane forms the point out's sum, with no branch (see below); the
// batched kernel keeps it in lane 63's branch, which holds it to 256 VGPRs.
template <typename T, bool ALL = true>
__device__ __forceinline__ void gs_pair_row(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
					    const T *__restrict__ r, int j, int k) {
	typedef typename Vec2<T>::type V2;
	const int lane = threadIdx.x;
	const long long row = idx3(L, 0, j, k);
	const long long px = L.px, pxy = L.pxy, pa = L.pa;
	const T *Ar = A + arow0(L, j, k);
	const T *rr_ = r + row;
	T *xr = x + row;
	T carry = (T)0;
	// the previous segment's results, stored after this segment's loads
	bool pE = false, pO = false, pZ = false;
	uint32_t pob = 0;
	T pe = (T)0, po = (T)0, pz = (T)0;
	for (int s0 = 0;; s0 += 128) {
		const bool last = L.m - s0 <= 129;
		const int i0 = s0 + 2 * lane;
		const bool live = i0 < L.m;
		const bool doE = live && !(lane == 0 && s0 > 0);
		const bool doO = live && i0 + 1 < L.m;
		const bool ext = lane == 63 && i0 + 2 < L.m;  // point out s0+128
		const bool wext = s0 + 128 < L.m;              // the wavefront has one
		const uint32_t ob = (uint32_t)i0 * (uint32_t)sizeof(T);
		const uint32_t oz = ob + 2u * (uint32_t)sizeof(T);
		// lanes 0..26: coefficient `lane` of the point out; lanes 0..8: x of
		// the point after it (s0+129, the halo where the row ends) in row
		// (dj, dk) = (lane % 3 - 1, lane / 3 - 1)
		T zco = (T)0, zx3 = (T)0;
		if (wext) {
			const uint32_t oz0 = (uint32_t)(s0 + 128) * (uint32_t)sizeof(T);
			if (lane < 27) zco = ld_row<T>(Ar + lane * pa, oz0);
			if (lane < 9) zx3 = ld_row<T>(xr + (lane % 3 - 1) * px + (lane / 3 - 1) * pxy + 1, oz0);
		}
		// Streamed row by row (dj,dk): the even point's sum is formed as the
		// loads arrive; the odd point's terms before its own row are summed
		// at once, its own-row coefficients are kept, and its 12 terms after
		// the own row are kept as products, to be added in box order once
		// the new even values exist.  Products are the same roundings the
		// in-order sum would form.
		T ax = (T)0, ay = (T)0, az = (T)0, ac_e = (T)1, ac_o = (T)1, ac_z = (T)1, a12 = (T)0, a14 = (T)0;
		T rx = (T)0, ry = (T)0, rz = (T)0, xh = (T)0;
		T prod[12];
		if (live) {
			const V2 rr = ld_stream<V2>(rr_, ob);
			rx = rr.x; ry = rr.y;
			if (ext) rz = ld_row<T>(rr_, oz);
#pragma unroll
			for (int rq = 0; rq < 9; rq++) {
				const int dj = rq % 3 - 1, dk = rq / 3 - 1;
				const T *xq = xr + dj * px + dk * pxy;
				const V2 am = ld_stream<V2>(Ar + (3 * rq) * pa, ob);
				const V2 a0 = ld_stream<V2>(Ar + (3 * rq + 1) * pa, ob);
				const V2 ap = ld_stream<V2>(Ar + (3 * rq + 2) * pa, ob);
				const V2 xv = ld_row<V2>(xq, ob);
				const T xm = ld_row<T>(xq - 1, ob);
				const T xp = doO ? ld_row<T>(xq + 2, ob) : (T)0;
				if (rq == 4) {
					ax = ax + am.x * xm;
					ax = ax + ap.x * xv.y;
					ac_e = a0.x; ac_o = a0.y; a12 = am.y; a14 = ap.y;
					xh = xp;  // i0+2: the halo point m when the row ends there
				} else {
					ax = ax + am.x * xm;
					ax = ax + a0.x * xv.x;
					ax = ax + ap.x * xv.y;
					if (rq < 4) {
						ay = ay + am.y * xv.x;
						ay = ay + a0.y * xv.y;
						ay = ay + ap.y * xp;
					} else {
						prod[3 * (rq - 5)] = am.y * xv.x;
						prod[3 * (rq - 5) + 1] = a0.y * xv.y;
						prod[3 * (rq - 5) + 2] = ap.y * xp;
					}
				}
				if (ALL || ext) {
					// point out i0+2 (lane 63's): neighbours i0+1 (xv.y), i0+2
					// (xp), i0+3.  With ALL every lane forms the sum, with no
					// branch, and only lane 63's is used: inside a branch these
					// products make the wavefront wait for each row step's loads
					// (nine round trips a segment instead of one).
					const T zm = read_lane(zco, 3 * rq);
					const T z0 = read_lane(zco, 3 * rq + 1);
					const T zp = read_lane(zco, 3 * rq + 2);
					const T x3 = read_lane(zx3, rq);
					az = az + zm * xv.y;
					if (rq == 4) ac_z = z0;
					else az = az + z0 * xp;
					az = az + zp * x3;
				}
			}
		}
		if (pE) st_row<T>(xr, pob, pe);
		if (pO) st_row<T>(xr + 1, pob, po);
		if (pZ) st_row<T>(xr, pob + 2u * (uint32_t)sizeof(T), pz);
		// even points: every neighbour still holds its old value
		T xe = (T)0, xz = (T)0, xo = (T)0;
		if (doE) xe = (rx - ax) * ((T)1 / ac_e);
		else if (live) xe = carry;
		if (ext) xz = (rz - az) * ((T)1 / ac_z);
		// Shuffles run with every lane active.  The odd point's right
		// neighbour is the next lane's new even point, lane 63's point out,
		// or, where the row ends (i0+2 = m), the halo value: zero on a
		// non-periodic axis, the ghost copy on a periodic one.
		const T xs = __shfl_down(xe, 1, 64);
		const T xe_next = i0 + 2 >= L.m ? xh : (lane == 63 ? xz : xs);
		if (doO) {
			// odd point i0+1: its even neighbours in the own row are new
			ay = ay + a12 * xe;
			ay = ay + a14 * xe_next;
#pragma unroll
			for (int q2 = 0; q2 < 12; q2++) ay = ay + prod[q2];
			xo = (ry - ay) * ((T)1 / ac_o);
		}
		carry = __shfl(xz, 63, 64);
		pE = doE; pO = doO; pZ = ext; pob = ob; pe = xe; po = xo; pz = xz;
		if (last) break;
	}
	if (pE) st_row<T>(xr, pob, pe);
	if (pO) st_row<T>(xr + 1, pob, po);
	if (pZ) st_row<T>(xr, pob + 2u * (uint32_t)sizeof(T), pz);
}

// FP32 variant of gs_pair_row with four points per lane: lane l holds
// i0 = s0 + 4l .. i0+3 (one 16-byte load per plane, i0 a multiple of 4 and
// so 16-byte aligned).  Even points i0, i0+2 update first from old
// neighbours; odd i0+1 then takes the lane's own new evens, odd i0+3 its own
// i0+2 and lane l+1's i0+4 (shuffle).  Segments are 256 points; lane 63's
// right even neighbour s0+256 (the first point of the next segment, or the
// last point of a row of 256q+1 points) is updated by lane 63 itself from
// scalar loads, with the evens, and handed to lane 0 of the next segment,
// which must not recompute it (its odd neighbour s0+255 has changed by
// then).  A 257-point row takes one segment, 513 points two.  A segment's
// stores are issued after the next segment's loads, as in gs_pair_row.
// Per-point arithmetic and order as in gs_point.  j and k must be
// wavefront-uniform.
template <bool Z = false>
__device__ __forceinline__ void gs_quad_row_f32(const LevelView &L, const float *__restrict__ A, float *__restrict__ x,
						const float *__restrict__ r, int j, int k) {
	const int cj = j & 1, ck = k & 1;
	const int lane = threadIdx.x;
	const long long row = idx3(L, 0, j, k);
	const long long px = L.px, pxy = L.pxy, pa = L.pa;
	const float *Ar = A + arow0(L, j, k);
	const float *rr_ = r + row;
	float *xr = x + row;
	float carry = 0.0f;
	// the previous segment's results, stored after this segment's loads
	bool pE0 = false, pE2 = false, pO1 = false, pO3 = false, pX = false;
	uint32_t pob = 0;
	float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, v3 = 0.0f, v4 = 0.0f;
	for (int s0 = 0;; s0 += 256) {
		const int i0 = s0 + 4 * lane;
		const bool last = L.m - s0 <= 257;
		const bool live = i0 < L.m;
		const bool doE0 = live && !(lane == 0 && s0 > 0);  // lane 0 takes i0 from the hand-over
		const bool doE2 = live && i0 + 2 < L.m;
		const bool doO1 = live && i0 + 1 < L.m;
		const bool doO3 = live && i0 + 3 < L.m;
		const bool ext = lane == 63 && i0 + 4 < L.m;  // even point out i0+4 = s0+256
		const uint32_t ob = (uint32_t)i0 * 4u;
		// lanes 0..26: coefficient `lane` of the point out; lanes 0..8: x of
		// the point after it (s0+257, the halo where the row ends) in row
		// (dj, dk) = (lane % 3 - 1, lane / 3 - 1), 0 in rows known to be 0.
		// Every lane reads them with v_readlane and forms the point out's sum,
		// with no branch (lane 63's is the one used): loaded by lane 63 inside
		// its branch, they made the wavefront wait for each row step's loads.
		float zco = 0.0f, zx5 = 0.0f;
		if (s0 + 256 < L.m) {
			const uint32_t oz0 = (uint32_t)(s0 + 256) * 4u;
			if (lane < 27) zco = ld_row<float>(Ar + lane * pa, oz0);
			if (lane < 9 && !(Z && zero_row(cj, ck, lane % 3 - 1, lane / 3 - 1)))
				zx5 = ld_row<float>(xr + (lane % 3 - 1) * px + (lane / 3 - 1) * pxy + 1, oz0);
		}
		// even sums e0 (i0), e2 (i0+2), ex (i0+4, ext); odd sums o1 (i0+1),
		// o3 (i0+3) up to the own row; own-row coefficients kept; the odd
		// points' 12 terms after the own row kept as products (the in-order
		// sums' roundings)
		float e0 = 0.0f, e2 = 0.0f, ex = 0.0f, o1 = 0.0f, o3 = 0.0f;
		float c0 = 1.0f, c1 = 1.0f, c2 = 1.0f, c3 = 1.0f, cx = 1.0f;
		float w1m = 0.0f, w1p = 0.0f, w3m = 0.0f, w3p = 0.0f, xh2 = 0.0f, xh4 = 0.0f;
		float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f, rx = 0.0f;
		float p1[12], p3[12];
		if (live) {
			const float4 rv = ld_stream<float4>(rr_, ob);
			r0 = rv.x; r1 = rv.y; r2 = rv.z; r3 = rv.w;
			if (ext) rx = ld_row<float>(rr_, ob + 16u);
#pragma unroll
			for (int rq = 0; rq < 9; rq++) {
				const int dj = rq % 3 - 1, dk = rq / 3 - 1;
				const float *xq = xr + dj * px + dk * pxy;
				const float4 am = ld_stream<float4>(Ar + (3 * rq) * pa, ob);
				const float4 a0 = ld_stream<float4>(Ar + (3 * rq + 1) * pa, ob);
				const float4 ap = ld_stream<float4>(Ar + (3 * rq + 2) * pa, ob);
				const bool zr = Z && zero_row(cj, ck, dj, dk);
				float4 xv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
				float xm = 0.0f, xp = 0.0f;
				if (!zr) {
					xv = ld_row<float4>(xq, ob);
					xm = ld_row<float>(xq - 1, ob);
					xp = doO3 ? ld_row<float>(xq + 4, ob) : 0.0f;
				}
				const float bm = read_lane(zco, 3 * rq);
				const float b0 = read_lane(zco, 3 * rq + 1);
				const float bp = read_lane(zco, 3 * rq + 2);
				const float x5 = read_lane(zx5, rq);  // the halo point when i0+4 = m-1
				if (rq == 4) {
					e0 = e0 + am.x * xm;
					e0 = e0 + ap.x * xv.y;
					e2 = e2 + am.z * xv.y;
					e2 = e2 + ap.z * xv.w;
					ex = ex + bm * xv.w;
					ex = ex + bp * x5;
					c0 = a0.x; c1 = a0.y; c2 = a0.z; c3 = a0.w; cx = b0;
					w1m = am.y; w1p = ap.y; w3m = am.w; w3p = ap.w;
					xh2 = xv.z; xh4 = xp;  // own-row i0+2, i0+4: the halo where the row ends
				} else {
					e0 = e0 + am.x * xm;
					e0 = e0 + a0.x * xv.x;
					e0 = e0 + ap.x * xv.y;
					e2 = e2 + am.z * xv.y;
					e2 = e2 + a0.z * xv.z;
					e2 = e2 + ap.z * xv.w;
					ex = ex + bm * xv.w;
					ex = ex + b0 * xp;
					ex = ex + bp * x5;
					if (rq < 4) {
						o1 = o1 + am.y * xv.x;
						o1 = o1 + a0.y * xv.y;
						o1 = o1 + ap.y * xv.z;
						o3 = o3 + am.w * xv.z;
						o3 = o3 + a0.w * xv.w;
						o3 = o3 + ap.w * xp;
					} else {
						p1[3 * (rq - 5)] = am.y * xv.x;
						p1[3 * (rq - 5) + 1] = a0.y * xv.y;
						p1[3 * (rq - 5) + 2] = ap.y * xv.z;
						p3[3 * (rq - 5)] = am.w * xv.z;
						p3[3 * (rq - 5) + 1] = a0.w * xv.w;
						p3[3 * (rq - 5) + 2] = ap.w * xp;
					}
				}
			}
		}
		if (pE0) st_row<float>(xr, pob, v0);
		if (pO1) st_row<float>(xr, pob + 4u, v1);
		if (pE2) st_row<float>(xr, pob + 8u, v2);
		if (pO3) st_row<float>(xr, pob + 12u, v3);
		if (pX) st_row<float>(xr, pob + 16u, v4);
		float x0 = 0.0f, x2 = 0.0f, xx = 0.0f, y1 = 0.0f, y3 = 0.0f;
		if (doE0) x0 = (r0 - e0) * (1.0f / c0);
		else if (live) x0 = carry;
		if (doE2) x2 = (r2 - e2) * (1.0f / c2);
		if (ext) xx = (rx - ex) * (1.0f / cx);
		// shuffles with every lane active; lanes past the row end hold 0
		const float xn = __shfl_down(x0, 1, 64);
		const float x4 = lane == 63 ? xx : xn;  // lane 63: the point out, or the halo
		if (doO1) {
			o1 = o1 + w1m * x0;
			o1 = o1 + w1p * (i0 + 2 >= L.m ? xh2 : x2);
#pragma unroll
			for (int q2 = 0; q2 < 12; q2++) o1 = o1 + p1[q2];
			y1 = (r1 - o1) * (1.0f / c1);
		}
		if (doO3) {
			o3 = o3 + w3m * x2;
			o3 = o3 + w3p * (i0 + 4 >= L.m ? xh4 : x4);
#pragma unroll
			for (int q2 = 0; q2 < 12; q2++) o3 = o3 + p3[q2];
			y3 = (r3 - o3) * (1.0f / c3);
		}
		pE0 = doE0; pE2 = doE2; pO1 = doO1; pO3 = doO3; pX = ext; pob = ob;
		v0 = x0; v1 = y1; v2 = x2; v3 = y3; v4 = xx;
		if (last) break;
		carry = __shfl(xx, 63, 64);
	}
	if (pE0) st_row<float>(xr, pob, v0);
	if (pO1) st_row<float>(xr, pob + 4u, v1);
	if (pE2) st_row<float>(xr, pob + 8u, v2);
	if (pO3) st_row<float>(xr, pob + 12u, v3);
	if (pX) st_row<float>(xr, pob + 16u, v4);
}

// Vectors of ppl elements moved with one load or store.
template <typename T, int N> struct VecOf;
template <> struct VecOf<float, 4> { typedef float4 type; };
template <> struct VecOf<float, 2> { typedef float2 type; };
template <> struct VecOf<double, 2> { typedef double2 type; };

template <typename T, int N>
__device__ __forceinline__ void ld_vec(const T *base, uint32_t off, T (&out)[N]) {
	typedef typename VecOf<T, N>::type V;
	const V v = ld_row<V>(base, off);
	const T *pv = reinterpret_cast<const T *>(&v);
#pragma unroll
	for (int t = 0; t < N; t++) out[t] = pv[t];
}
// Coefficients and r: read once per launch (non-temporal in FP64).
template <typename T, int N>
__device__ __forceinline__ void ld_vec_stream(const T *base, uint32_t off, T (&out)[N]) {
	typedef typename VecOf<T, N>::type V;
	const V v = ld_stream<V>(base, off);
	const T *pv = reinterpret_cast<const T *>(&v);
#pragma unroll
	for (int t = 0; t < N; t++) out[t] = pv[t];
}
// Coefficient planes 0..split-1 of a level are read with plain loads, the
// rest with the non-temporal hint.  In a lone hierarchy (L.cache_coef), a
// level whose coefficients fit the 256 MB memory-side cache (MALL) keeps
// them there from one sweep to the next (split 27); one up to 1 GB keeps a
// third of its planes there (a cyclic sweep over more than the cache would
// otherwise evict every line before its reuse); a larger one keeps one plane
// if that is at most 80 MB, and streams the rest.  Measured on 257^3: FP32
// levels 1-2 and FP64 level 2 fastest with plain loads, FP64 level 1
// (464 MB) 7 % faster with 9 planes cached, FP32 level 0 1.5 % faster with
// one (two cost level 1 its cached planes).  Batched slots run side by side,
// their levels together far larger than the cache: they stream everything
// (measured: 256-problem batch 16.3 s with the per-level rule, 13.6 s
// streamed).
template <typename T>
__device__ __forceinline__ int coef_split(const LevelView &L) {
	if (!L.cache_coef) return 0;
	const long long plane = (long long)L.m * L.n * L.o * (long long)sizeof(T);
	if (27 * plane <= (256LL << 20)) return 27;
	if (27 * plane <= (1024LL << 20)) return 9;
	return plane <= (80LL << 20) ? 1 : 0;
}
template <typename T, int N>
__device__ __forceinline__ void ld_vec_coef(const T *base, uint32_t off, T (&out)[N], bool nt) {
	if (nt) ld_vec_stream<T, N>(base, off, out);
	else ld_vec<T, N>(base, off, out);
}
template <typename T, int N>
__device__ __forceinline__ void st_vec(T *base, uint32_t off, const T (&v)[N]) {
	typedef typename VecOf<T, N>::type V;
	V o;
	T *po = reinterpret_cast<T *>(&o);
#pragma unroll
	for (int t = 0; t < N; t++) po[t] = v[t];
	st_row<V>(base, off, o);
}

// How the packed-row kernels take the coefficients of a row's point out (its
// last point m-1, done by the row's last lane):
//   PO_ROWS  FP64: where the level has point outs (a run-time test), the last
//            lane loads them from the plane rows, a plane row apart;
//   PO_NONE  FP32 levels without point outs (rows of other than 4q+1 points);
//   PO_TAIL  FP32 levels with point outs: every lane loads its row's tail
//            copy (see arow0), 16 bytes at a time, so the loads need no
//            branch: the lanes of a row share the addresses, and only the
//            last lane uses the values.  Loaded into registers inside a
//            branch, they made the wavefront wait for its loads at every row
//            step (nine round trips instead of one).
// In FP64 the tail's loads, or a second inlined variant, took the kernels
// past 256 VGPRs (one wavefront per SIMD instead of two).
enum { PO_NONE = 0, PO_ROWS = 1, PO_TAIL = 2 };

// Coefficients 3rq .. 3rq+2 of the row's point out from its tail copy (ot:
// the copy's byte offset from Ab), read 16 bytes at a time into tb as the
// planes are reached; plain loads, so a level that keeps its coefficients
// cached keeps the copy cached too.  rq must be a compile-time constant (an
// unrolled loop), so tb stays in registers.
template <typename T>
__device__ __forceinline__ void point_out_tail(const T *Ab, uint32_t ot, int rq, T (&tb)[28], T (&b)[3]) {
	constexpr int VW = 16 / (int)sizeof(T);
#pragma unroll
	for (int d = 0; d < 3; d++) {
		const int e = 3 * rq + d;
		if (e % VW == 0) {
			T v[VW];
			ld_vec<T, VW>(Ab, ot + (uint32_t)(e * (int)sizeof(T)), v);
#pragma unroll
			for (int q = 0; q < VW; q++) tb[e + q] = v[q];
		}
		b[d] = tb[e];
	}
}

// Forward colour-pair update of rows that fit one wavefront's lanes
// (row_lanes: rpw rows per wavefront, lpr lanes per row, ppl points per
// lane, the last lane of a 2^k+1 row also doing the point out).  The
// wavefront owns rows q0 .. q0+rpw-1 of the launch's colour pair; each array
// is addressed from row q0's start (scalar registers) plus a per-lane 32-bit
// offset.  Even points update from old neighbours; odd points then take
// their own lane's new evens and the next lane's first even (shuffle), or
// at a row's last lane the point out (or the halo).  Only the wavefront's
// rows are written by the launch, so each lane stores its points once, at
// the end.  Per-point arithmetic and order as in gs_point.
template <typename T, int PPL, bool Z, int PO>
__device__ __forceinline__ void gs_lanes(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
					 const T *__restrict__ r, int cj, int ck, int q0, int nrows) {
	const RowLanes W = row_lanes(L.m, (int)sizeof(T));
	const int lane = threadIdx.x;
	const int rr = lane / W.lpr, li = lane - rr * W.lpr;
	const int q = q0 + rr;
	const bool live = rr < W.rpw && q < nrows;
	const int nb = (L.n - cj + 1) >> 1;
	const int j0 = 2 * (q0 % nb) + cj, k0 = 2 * (q0 / nb) + ck;
	const long long xb0 = idx3(L, 0, j0, k0), ab0 = arow0(L, j0, k0);
	const long long px = L.px, pxy = L.pxy, pa = L.pa;
	const int split = coef_split<T>(L);
	const T *Ab = A + ab0;
	const T *rb = r + xb0;
	T *xb = x + xb0;
	int i0 = 0;
	uint32_t ox = 0, oa = 0;
	if (live) {
		const int j = 2 * (q % nb) + cj, k = 2 * (q / nb) + ck;
		i0 = PPL * li;
		ox = (uint32_t)((idx3(L, 0, j, k) - xb0 + i0) * (long long)sizeof(T));
		oa = (uint32_t)((arow0(L, j, k) - ab0 + i0) * (long long)sizeof(T));
	}
	const bool ext = (PO == PO_TAIL || (PO == PO_ROWS && W.ext)) && live && li == W.lpr - 1;  // even point i0+PPL = m-1
	const bool lastlane = li == W.lpr - 1;
	const bool needxp = live && i0 + PPL <= L.m;
	const uint32_t oe = (uint32_t)(PPL * sizeof(T));
	const uint32_t ot = oa + (uint32_t)((27 * pa - i0) * (long long)sizeof(T));  // the tail copy
	T tb[28];
	T s[PPL], cc[PPL], wm[PPL], wp[PPL], rv[PPL];
	T prod[PPL / 2][12];
	T sx = (T)0, cx = (T)1, rx = (T)0, xhalo = (T)0;
#pragma unroll
	for (int t = 0; t < PPL; t++) { s[t] = (T)0; cc[t] = (T)1; wm[t] = (T)0; wp[t] = (T)0; rv[t] = (T)0; }
	if (live) {
		ld_vec_stream<T, PPL>(rb, ox, rv);
		if (ext) rx = ld_row<T>(rb, ox + oe);
#pragma unroll
		for (int rq = 0; rq < 9; rq++) {
			const int dj = rq % 3 - 1, dk = rq / 3 - 1;
			const T *xq = xb + dj * px + dk * pxy;
			T a[3][PPL], xv[PPL];
#pragma unroll
			for (int d = 0; d < 3; d++) ld_vec_coef<T, PPL>(Ab + (3 * rq + d) * pa, oa, a[d], 3 * rq + d >= split);
			T xm = (T)0, xp = (T)0;
			if (Z && zero_row(cj, ck, dj, dk)) {
#pragma unroll
				for (int t = 0; t < PPL; t++) xv[t] = (T)0;
			} else {
				ld_vec<T, PPL>(xq, ox, xv);
				xm = ld_row<T>(xq - 1, ox);
				xp = needxp ? ld_row<T>(xq + PPL, ox) : (T)0;
			}
			T b[3] = {(T)0, (T)0, (T)0};
			if constexpr (PO == PO_TAIL) {
				point_out_tail<T>(Ab, ot, rq, tb, b);
			} else if constexpr (PO == PO_ROWS) {
				if (ext) {
#pragma unroll
					for (int d = 0; d < 3; d++) b[d] = ld_row<T>(Ab + (3 * rq + d) * pa, oa + oe);
				}
			}
			if (rq == 4) {
				// own-row value at m, where the row ends inside this lane
#pragma unroll
				for (int t = 1; t < PPL; t++)
					if (i0 + t == L.m) xhalo = xv[t];
				if (i0 + PPL == L.m) xhalo = xp;
			}
#pragma unroll
			for (int t = 0; t < PPL; t++) {
				const T vm = t == 0 ? xm : xv[t - 1 < 0 ? 0 : t - 1];
				const T v0 = xv[t];
				const T vp = t == PPL - 1 ? xp : xv[t + 1 < PPL ? t + 1 : PPL - 1];
				if ((t & 1) == 0) {
					if (rq == 4) {
						s[t] = s[t] + a[0][t] * vm;
						s[t] = s[t] + a[2][t] * vp;
						cc[t] = a[1][t];
					} else {
						s[t] = s[t] + a[0][t] * vm;
						s[t] = s[t] + a[1][t] * v0;
						s[t] = s[t] + a[2][t] * vp;
					}
				} else if (rq < 4) {
					s[t] = s[t] + a[0][t] * vm;
					s[t] = s[t] + a[1][t] * v0;
					s[t] = s[t] + a[2][t] * vp;
				} else if (rq == 4) {
					wm[t] = a[0][t]; cc[t] = a[1][t]; wp[t] = a[2][t];
				} else {
					prod[t / 2][3 * (rq - 5)] = a[0][t] * vm;
					prod[t / 2][3 * (rq - 5) + 1] = a[1][t] * v0;
					prod[t / 2][3 * (rq - 5) + 2] = a[2][t] * vp;
				}
			}
			// the point out: neighbours i0+PPL-1, itself, the halo i = m
			if (rq == 4) {
				sx = sx + b[0] * xv[PPL - 1];
				sx = sx + b[2] * (T)0;
				cx = b[1];
			} else {
				sx = sx + b[0] * xv[PPL - 1];
				sx = sx + b[1] * xp;
				sx = sx + b[2] * (T)0;
			}
		}
	}
	T xn[PPL];
#pragma unroll
	for (int t = 0; t < PPL; t++) xn[t] = (T)0;
#pragma unroll
	for (int t = 0; t < PPL; t += 2)
		if (live && i0 + t < L.m) xn[t] = (rv[t] - s[t]) * ((T)1 / cc[t]);
	T xe = (T)0;
	if (ext) xe = (rx - sx) * ((T)1 / cx);
	// shuffles with every lane active
	const T nxt = __shfl_down(xn[0], 1, 64);
	const T right = lastlane ? xe : nxt;
#pragma unroll
	for (int t = 1; t < PPL; t += 2) {
		if (live && i0 + t < L.m) {
			// right neighbour: a new even point, or the halo value (zero on a
			// non-periodic axis, the ghost copy on a periodic one) where the
			// row ends
			const T rt = i0 + t + 1 >= L.m ? xhalo : (t + 1 < PPL ? xn[t + 1 < PPL ? t + 1 : 0] : right);
			T acc = s[t];
			acc = acc + wm[t] * xn[t - 1];
			acc = acc + wp[t] * rt;
#pragma unroll
			for (int q2 = 0; q2 < 12; q2++) acc = acc + prod[t / 2][q2];
			xn[t] = (rv[t] - acc) * ((T)1 / cc[t]);
		}
	}
	if (live) {
		if (i0 + PPL <= L.m) {
			st_vec<T, PPL>(xb, ox, xn);
		} else {
#pragma unroll
			for (int t = 0; t < PPL; t++)
				if (i0 + t < L.m) st_row<T>(xb, ox + (uint32_t)(t * sizeof(T)), xn[t]);
		}
		if (ext) st_row<T>(xb, ox + oe, xe);
	}
}

// Residual of rows that fit one wavefront's lanes (layout as gs_lanes);
// returns this lane's sum of squares.  zx (may be null) is set to 0 at the
// same points.
template <typename T, typename TR, int PPL, int PO>
__device__ __forceinline__ double residual_lanes(const LevelView &L, const T *__restrict__ A, const T *__restrict__ x,
						 const T *__restrict__ r, TR *__restrict__ res, TR *__restrict__ zx,
						 int q0, int nrows) {
	const RowLanes W = row_lanes(L.m, (int)sizeof(T));
	const int lane = threadIdx.x;
	const int rr = lane / W.lpr, li = lane - rr * W.lpr;
	const int q = q0 + rr;
	const bool live = rr < W.rpw && q < nrows;
	const int j0 = q0 % L.n, k0 = q0 / L.n;
	const long long xb0 = idx3(L, 0, j0, k0), ab0 = arow0(L, j0, k0);
	const long long px = L.px, pxy = L.pxy, pa = L.pa;
	const int split = coef_split<T>(L);
	double v2 = 0.0;
	if (!live) return v2;
	const int j = q % L.n, k = q / L.n;
	const int i0 = PPL * li;
	const long long dx = idx3(L, 0, j, k) - xb0 + i0;
	const uint32_t ox = (uint32_t)(dx * (long long)sizeof(T));
	const uint32_t oa = (uint32_t)((arow0(L, j, k) - ab0 + i0) * (long long)sizeof(T));
	const bool ext = (PO == PO_TAIL || (PO == PO_ROWS && W.ext)) && li == W.lpr - 1;
	const bool needxp = i0 + PPL <= L.m;
	const uint32_t oe = (uint32_t)(PPL * sizeof(T));
	const uint32_t ot = oa + (uint32_t)((27 * pa - i0) * (long long)sizeof(T));  // the tail copy
	const uint32_t oz = oa + (uint32_t)((L.m - 1 - i0) * (long long)sizeof(T));  // the row's point m-1
	T tb[28];
	const T *Ab = A + ab0, *rb = r + xb0, *xb = x + xb0;
	T s[PPL], rv[PPL];
	T sx = (T)0, rx = (T)0;
#pragma unroll
	for (int t = 0; t < PPL; t++) s[t] = (T)0;
	ld_vec_stream<T, PPL>(rb, ox, rv);
	if (ext) rx = ld_row<T>(rb, ox + oe);
#pragma unroll
	for (int rq = 0; rq < 9; rq++) {
		const int dj = rq % 3 - 1, dk = rq / 3 - 1;
		const T *xq = xb + dj * px + dk * pxy;
		T a[3][PPL], xv[PPL];
#pragma unroll
		for (int d = 0; d < 3; d++) ld_vec_coef<T, PPL>(Ab + (3 * rq + d) * pa, oa, a[d], 3 * rq + d >= split);
		ld_vec<T, PPL>(xq, ox, xv);
		const T xm = ld_row<T>(xq - 1, ox);
		const T xp = needxp ? ld_row<T>(xq + PPL, ox) : (T)0;
#pragma unroll
		for (int t = 0; t < PPL; t++) {
			const T vm = t == 0 ? xm : xv[t - 1 < 0 ? 0 : t - 1];
			const T vp = t == PPL - 1 ? xp : xv[t + 1 < PPL ? t + 1 : PPL - 1];
			s[t] = s[t] + a[0][t] * vm;
			s[t] = s[t] + a[1][t] * xv[t];
			s[t] = s[t] + a[2][t] * vp;
		}
		if constexpr (PO == PO_TAIL) {
			T b[3];
			point_out_tail<T>(Ab, ot, rq, tb, b);
			// every lane sums; only the last lane's sx is used
			sx = sx + b[0] * xv[PPL - 1];
			sx = sx + b[1] * xp;
			sx = sx + b[2] * (T)0;  // the halo point i = m
		} else if constexpr (PO == PO_ROWS) {
			// every lane loads its row's point-out entries (the lanes of a row
			// share the addresses) and sums, with no branch; only the last
			// lane's sx is used
			if (W.ext) {
				sx = sx + ld_row<T>(Ab + (3 * rq) * pa, oz) * xv[PPL - 1];
				sx = sx + ld_row<T>(Ab + (3 * rq + 1) * pa, oz) * xp;
				sx = sx + ld_row<T>(Ab + (3 * rq + 2) * pa, oz) * (T)0;  // the halo point i = m
			}
		}
	}
	TR o[PPL], z[PPL];
#pragma unroll
	for (int t = 0; t < PPL; t++) {
		const T v = rv[t] - s[t];
		o[t] = (TR)v;
		z[t] = (TR)0;
		if (i0 + t < L.m) v2 = v2 + (double)v * (double)v;
	}
	const T vx = rx - sx;
	if (ext) v2 = v2 + (double)vx * (double)vx;
	const uint32_t orr = (uint32_t)(dx * (long long)sizeof(TR)), oer = (uint32_t)(PPL * sizeof(TR));
	TR *resb = res ? res + xb0 : nullptr;
	TR *zxb = zx ? zx + xb0 : nullptr;
	if (i0 + PPL <= L.m) {
		if (resb) st_vec<TR, PPL>(resb, orr, o);
		if (zxb) st_vec<TR, PPL>(zxb, orr, z);
	} else {
#pragma unroll
		for (int t = 0; t < PPL; t++)
			if (i0 + t < L.m) {
				if (resb) st_row<TR>(resb, orr + (uint32_t)(t * sizeof(TR)), o[t]);
				if (zxb) st_row<TR>(zxb, orr + (uint32_t)(t * sizeof(TR)), (TR)0);
			}
	}
	if (ext) {
		if (resb) st_row<TR>(resb, orr + oer, (TR)vx);
		if (zxb) st_row<TR>(zxb, orr + oer, (TR)0);
	}
	return v2;
}

// Forward colour-pair update of one row: four points per lane in FP32,
// two in FP64 (whose 16-byte pair loads already fill the registers).
// The first-sweep-from-zero form (Z) exists in FP32 only.  ALL: as in
// gs_pair_row.
template <typename T, bool Z, bool ALL = true>
__device__ __forceinline__ void gs_row(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
				       const T *__restrict__ r, int j, int k) {
	if constexpr (sizeof(T) == 4) {
		gs_quad_row_f32<Z>(L, A, x, r, j, k);
	} else {
		static_assert(!Z, "first sweep from zero: FP32 only");
		gs_pair_row<T, ALL>(L, A, x, r, j, k);
	}
}

template <typename T> struct Ppl { static const int v = sizeof(T) == 4 ? 4 : 2; };

// Wavefront w of a colour-pair launch: rows w*rpw .. of the pair when they
// fit the lanes, else the single long row w, streamed in segments.  w must be
// wavefront-uniform.
template <typename T, bool Z = false, bool ALL = true>
__device__ __forceinline__ void gs_wave(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
					const T *__restrict__ r, int cj, int ck, int w) {
	const int nb = (L.n - cj + 1) >> 1, nd = (L.o - ck + 1) >> 1, rows = nb * nd;
	const RowLanes W = row_lanes(L.m, (int)sizeof(T));
	if (W.lpr <= 64) {
		const int q0 = w * W.rpw;
		if (q0 < rows) {
			// FP32 point outs come with a tail copy (m = 4q+1)
			if constexpr (sizeof(T) == 8) gs_lanes<T, Ppl<T>::v, Z, PO_ROWS>(L, A, x, r, cj, ck, q0, rows);
			else if (W.ext) gs_lanes<T, Ppl<T>::v, Z, PO_TAIL>(L, A, x, r, cj, ck, q0, rows);
			else gs_lanes<T, Ppl<T>::v, Z, PO_NONE>(L, A, x, r, cj, ck, q0, rows);
		}
	} else if (w < rows) {
		gs_row<T, Z, ALL>(L, A, x, r, 2 * (w % nb) + cj, 2 * (w / nb) + ck);
	}
}

template <typename T, bool Z = false>
__global__ __launch_bounds__(256) void k_gs_pair(LevelView L, const T *__restrict__ A, T *__restrict__ x,
						 const T *__restrict__ r, int cj, int ck) {
	const int w = __builtin_amdgcn_readfirstlane(xcd_block((int)blockIdx.x, (int)gridDim.x) * (int)blockDim.y + (int)threadIdx.y);
	gs_wave<T, Z>(L, A, x, r, cj, ck, w);
}

// Wavefronts of one colour-pair launch.
template <typename T>
__host__ __device__ inline int gs_pair_waves(const LevelView &L, int pair) {
	const int cj = pair & 1, ck = (pair >> 1) & 1;
	const int rows = ((L.n - cj + 1) >> 1) * ((L.o - ck + 1) >> 1);
	return row_lanes(L.m, (int)sizeof(T)).lpr <= 64 ? row_waves(rows, L.m, (int)sizeof(T)) : rows;
}

// Wavefronts per workgroup of the row smoother: one, so the XCD-contiguous
// row mapping works at the finest grain.
constexpr int kGsRows = 1;

template <typename T>
void launch_gs_pair(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int pair, bool from_zero) {
	const int cj = pair & 1, ck = (pair >> 1) & 1;
	const int nb = (L.n - cj + 1) >> 1, nd = (L.o - ck + 1) >> 1;
	if (nb <= 0 || nd <= 0) return;
	const int waves = gs_pair_waves<T>(L, pair);
	const dim3 grid((waves + kGsRows - 1) / kGsRows), block(64, kGsRows, 1);
	if constexpr (sizeof(T) == 4) {
		if (from_zero) {
			hipLaunchKernelGGL((k_gs_pair<T, true>), grid, block, 0, s, L, A, x, r, cj, ck);
			HIPCHECK_LAUNCH();
			return;
		}
	}
	hipLaunchKernelGGL((k_gs_pair<T, false>), grid, block, 0, s, L, A, x, r, cj, ck);
	HIPCHECK_LAUNCH();
}

// Hyperplane i + 2j + 4k = t.  Every stencil offset (di,dj,dk) != 0 has
// di + 2dj + 4dk != 0, negative exactly for the offsets that precede the
// centre in lexicographic order; sweeping t upwards therefore gives each
// point the same mix of new and old neighbours as the lexicographic sweep of
// region.cc:367-446, and t downwards the reverse sweep of region.cc:463-537.
// Thread (jj, kk) takes row k = klo + kk, j = jlo(k) + jj, i = t - 2j - 4k.
template <typename T>
__global__ __launch_bounds__(256) void k_gs_plane(LevelView L, const T *__restrict__ A, T *__restrict__ x,
						  const T *__restrict__ r, int t, int klo) {
	const int k = klo + (int)blockIdx.y;
	const int sres = t - 4 * k;  // i + 2j
	const int jlo = max(0, (sres - (L.m - 1) + 1) >> 1);
	const int jhi = min(L.n - 1, sres >> 1);
	const int j = jlo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j > jhi) return;
	const int i = sres - 2 * j;
	gs_point<T>(L, A, x, r, i, j, k);
}

template <typename T>
void launch_gs_plane(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int t) {
	const int maxi = L.m - 1, maxj = L.n - 1;
	int klo = t - maxi - 2 * maxj;
	klo = klo <= 0 ? 0 : (klo + 3) / 4;
	const int khi = min(L.o - 1, t / 4);
	if (khi < klo) return;
	const int jspan = min(L.n, (L.m + 1) / 2 + 1);
	const int bs = jspan >= 128 ? 128 : 64;
	dim3 grid((jspan + bs - 1) / bs, khi - klo + 1, 1), block(bs, 1, 1);
	hipLaunchKernelGGL(k_gs_plane<T>, grid, block, 0, s, L, A, x, r, t, klo);
	HIPCHECK_LAUNCH();
}

// Several forward sweeps in flight: launch tau updates hyperplane tau - 8q
// of sweep q (blockIdx.z = q).  A stencil offset moves t by at most 7.  When
// sweep q updates plane t, its own planes below t are done, and sweep q-1
// has finished every plane up to t+7, whose values are the ones sweep q
// must read there.  Sweep q-1 works on plane t+8 in the same launch, which
// shares no stencil with plane t.  Every point therefore sees exactly the
// neighbour values of the sweeps run one after another: the result is
// bit-identical, with tmax+1+8(S-1) launches for S sweeps instead of
// S(tmax+1).
template <typename T>
__global__ __launch_bounds__(128) void k_gs_planes(LevelView L, const T *__restrict__ A, T *__restrict__ x,
						   const T *__restrict__ r, int tau, int tmax) {
	const int t = tau - 8 * (int)blockIdx.z;
	if (t < 0 || t > tmax) return;
	const int k = (int)blockIdx.y;
	const int sres = t - 4 * k;  // i + 2j
	if (sres < 0) return;
	const int jlo = max(0, (sres - (L.m - 1) + 1) >> 1);
	const int jhi = min(L.n - 1, sres >> 1);
	const int j = jlo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j > jhi) return;
	gs_point<T>(L, A, x, r, sres - 2 * j, j, k);
}

template <typename T>
void launch_gs_planes(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int nsweeps) {
	const int tmax = (L.m - 1) + 2 * (L.n - 1) + 4 * (L.o - 1);
	const int jspan = min(L.n, (L.m + 1) / 2 + 1);
	const dim3 grid((jspan + 127) / 128, L.o, nsweeps), block(128, 1, 1);
	for (int tau = 0; tau <= tmax + 8 * (nsweeps - 1); tau++) {
		hipLaunchKernelGGL(k_gs_planes<T>, grid, block, 0, s, L, A, x, r, tau, tmax);
		HIPCHECK_LAUNCH();
	}
}

// The same pipelined sweeps on the plane-skewed copies (SkewView): the
// threads of a block take consecutive j of one (t, k), whose points, their
// coefficients and each of their neighbours are consecutive in memory, so
// every load of a hyperplane is a coalesced row segment.  Per-point
// arithmetic and order are gs_point's.
template <typename T>
__global__ __launch_bounds__(256) void k_gs_planes_skew(SkewView S, const T *__restrict__ As, T *__restrict__ xs,
							const T *__restrict__ rs, int tau, int tmax) {
	const int t = tau - 8 * (int)blockIdx.z;
	if (t < 0 || t > tmax) return;
	const int k = (int)blockIdx.y;
	const int sres = t - 4 * k;  // i + 2j
	if (sres < 0) return;
	const int jlo = max(0, (sres - (S.m - 1) + 1) >> 1);
	const int jhi = min(S.n - 1, sres >> 1);
	const int j = jlo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j > jhi) return;
	const int i = sres - 2 * j;
	const long long p = sidx3(S, i, j, k);
	T ax = (T)0;
#pragma unroll
	for (int e = 0; e < 27; e++) {
		if (e == 13) continue;
		const int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
		ax = ax + As[saidx3(S, e, i, j, k)] * xs[sidx3(S, i + di, j + dj, k + dk)];
	}
	xs[p] = (rs[p] - ax) * ((T)1 / As[saidx3(S, 13, i, j, k)]);
}

template <typename T>
void launch_gs_planes_skew(hipStream_t s, const SkewView &S, const T *As, T *xs, const T *rs, int nsweeps) {
	const int tmax = (S.m - 1) + 2 * (S.n - 1) + 4 * (S.o - 1);
	const int jspan = min(S.n, (S.m + 1) / 2 + 1);
	// one block per (k, sweep) row segment where it fits
	const int bs = jspan <= 64 ? 64 : (jspan <= 128 ? 128 : 256);
	const dim3 grid((jspan + bs - 1) / bs, S.o, nsweeps), block(bs, 1, 1);
	for (int tau = 0; tau <= tmax + 8 * (nsweeps - 1); tau++) {
		hipLaunchKernelGGL(k_gs_planes_skew<T>, grid, block, 0, s, S, As, xs, rs, tau, tmax);
		HIPCHECK_LAUNCH();
	}
}

// ---- exact sweeps in one launch: planes signal their progress ----------
//
// Workgroup b owns plane k (b renumbered so that each XCD holds a run of
// consecutive planes) of the plane-skewed copies and walks its rows
// u = 0 .. umax in order; row u of plane k is the part of hyperplane
// t = u + 4k in that plane.  The nsweeps forward sweeps of a call run
// together, sweep s kFlowLag rows behind sweep s-1: step tau updates row
// tau - kFlowLag*s of every sweep s that has one.  Thread (s, jj) updates
// point j = jlo(u) + jj of sweep s's row.
//
// The stencil of a point in row (k, u) reads
//   plane k-1, rows u-3 .. u+3: hyperplanes t-7 .. t-1, which must hold
//     their new values (this sweep's);
//   plane k, rows u-3 .. u-1 (this sweep, earlier steps here) and
//     u+1 .. u+3 (the previous sweep's values, written kFlowLag-3 or more
//     steps ago);
//   plane k+1, rows u-3 .. u+3: hyperplanes t+1 .. t+7, which must still
//     hold the previous sweep's values (the start values for sweep 0).
// So step tau waits until plane k-1 has finished row min(u+3, umax) of
// sweep s, and plane k+1 row min(u+3, umax) of sweep s-1.  The converse
// holds as well: plane k+1 overwrites a row u' in sweep s only after this
// plane has finished row u'+3 of sweep s, and plane k-1 overwrites row u'
// in sweep s+1 only after this plane has finished row u'+3 of sweep s, so
// no value is overwritten before all its readers have read it.  Every point
// therefore reads the values it reads in the sweeps run one after another in
// lexicographic order (the hyperplane argument of k_gs_plane): bit-identical
// results, with the per-point arithmetic of gs_point.
//
// A step: wait (from counts read one step earlier where they suffice, else
// by polling), load every x value the step's points read, wait for all of
// the wave's memory accesses, publish that the previous step is done, store
// the new values, and load the coefficients and r of the next step's points
// (they depend on no other plane).  One wavefront of the workgroup does the
// waiting and publishing and nothing else, so its polls never queue behind
// those loads.  prog[k] counts the finished steps.  The count is published
// one step late because that costs no wait of its own: the previous step's
// stores were issued before this step's loads, so once those loads are back
// and vmcnt is 0 the stores are complete at agent scope, in every wave of
// the workgroup after the barrier.  x stores and the loads of the
// neighbouring planes are agent-scope (sc1), so no XCD's L2 hands out a line
// that another XCD has since written.
//
// The waits form no cycle: with the late count, with F(k, tau) = tau + 5k
// every wait of a step is for a step of strictly smaller F when
// kFlowLag >= 15, so the launch completes whenever all its workgroups are
// resident together, which the launcher checks.  Every wait is bounded:
// after `limit` polls it sets *err and its workgroup leaves; the other
// workgroups leave when they see *err.  The last workgroup out (count in
// prog[o]) clears the counters for the next launch.
const int kFlowLag = 16;

template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt vmcnt(0): every vector memory access of the wave has completed.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Loads and stores of x in k_gs_flow: raw buffer accesses (32-bit byte
// offsets; the launcher keeps each skewed array under 2 GiB), with the
// agent-scope cache policy (sc1) where another workgroup writes or reads the
// value.  Unlike atomic loads, the compiler issues these together, so a
// point's 26 neighbours cost one memory round trip, not 26.
constexpr int kLoadSc1 = 16;
template <typename T>
__device__ __forceinline__ T ld_x(const __amdgpu_buffer_rsrc_t &rs, int off, bool agent) {
	if constexpr (sizeof(T) == 8) {
		typedef unsigned u2v __attribute__((ext_vector_type(2)));
		const u2v w = agent ? __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, kLoadSc1)
				    : __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
		T v;
		__builtin_memcpy(&v, &w, 8);
		return v;
	} else {
		const unsigned w = agent ? __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, kLoadSc1)
					 : __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
		T v;
		__builtin_memcpy(&v, &w, 4);
		return v;
	}
}
template <typename T>
__device__ __forceinline__ void st_x(const __amdgpu_buffer_rsrc_t &rs, int off, T v) {
	if constexpr (sizeof(T) == 8) {
		typedef unsigned u2v __attribute__((ext_vector_type(2)));
		u2v w;
		__builtin_memcpy(&w, &v, 8);
		__builtin_amdgcn_raw_buffer_store_b64(w, rs, off, 0, kStoreSc1);
	} else {
		unsigned w;
		__builtin_memcpy(&w, &v, 4);
		__builtin_amdgcn_raw_buffer_store_b32(w, rs, off, 0, kStoreSc1);
	}
}

// Points of the longest row (k, u) of a level: rows of the plane-skewed
// layout hold the j with i = u - 2j in [0, m).
__host__ __device__ inline int flow_row_points(const SkewView &S) { return min(S.n, (S.m + 1) / 2); }

// The counts plane k-1 (*need_b) and plane k+1 (*need_a) must have reached
// before step tau.
__device__ __forceinline__ void flow_needs(int tau, int nsw, int umax, int *need_b, int *need_a) {
	int nb = 0, na = 0;
	for (int s = 0; s < nsw; s++) {
		const int u = tau - kFlowLag * s;
		if (u < 0 || u > umax) continue;
		const int top = min(u + 3, umax);
		nb = max(nb, top + kFlowLag * s + 1);
		if (s > 0) na = max(na, top + kFlowLag * (s - 1) + 1);
	}
	*need_b = nb;
	*need_a = na;
}

// Polls counter *c until it reaches need (known: last value seen).
// Returns false when the wait gives up (limit polls, or *err set).
__device__ __forceinline__ bool flow_wait(const int *c, int need, int &known, int *err, int limit) {
	int spins = 0;
	while (known < need) {
		known = ld_agent(c);
		if (known >= need) break;
		if (ld_agent(err) != 0) return false;
		if (++spins > limit) {
			st_agent(err, 1);
			return false;
		}
		__builtin_amdgcn_s_sleep(1);
	}
	return true;
}

// Block size of k_gs_flow: rpad = the longest row rounded up to whole
// wavefronts, per sweep in flight, so that each wavefront holds a segment of
// one sweep's row (its row bases are wave-uniform), and one wavefront more.
__host__ __device__ inline int flow_row_pad(const SkewView &S) { return ((flow_row_points(S) + 63) / 64) * 64; }
const int kFlowMaxThreads = 512;
const int kFlowMaxPlanes = 160;

// First j of row u in the plane-skewed layout's row storage (sidx3).
__device__ __forceinline__ int skew_jf(const SkewView &S, int u) { return S.cmp ? max(-1, (u - S.m + 1) >> 1) : -1; }

// PRD: the level has periodic axes, whose halo holds copies of x, not zeros.
template <typename T, bool PRD>
__global__ __launch_bounds__(kFlowMaxThreads) void k_gs_flow(SkewView S, const T *__restrict__ As, T *__restrict__ xs,
							     const T *__restrict__ rs, int nsw, int rpad,
							     int *__restrict__ prog, int *__restrict__ err, int limit) {
	// the last four rows of each sweep in flight, [sweep][u & 3][sidx3's
	// position in the row]: the in-plane neighbours a point takes from the
	// rows just before its own, which this workgroup stored in the last
	// three steps (a load from memory would wait for those stores)
	extern __shared__ char flow_lds_[];
	T *const ring = reinterpret_cast<T *>(flow_lds_);
	__shared__ int stop;
	const int k = xcd_block((int)blockIdx.x, (int)gridDim.x);
	const int umax = (S.m - 1) + 2 * (S.n - 1);
	const int nsteps = umax + 1 + kFlowLag * (nsw - 1);
	const int tid = (int)threadIdx.x;
	// waves 0 .. nsw*rpad/64 - 1 update points; the last wave (the watcher)
	// waits for and publishes the counts, with no loads of its own in flight
	const int s = __builtin_amdgcn_readfirstlane(tid / rpad), jj = tid - s * rpad;
	const bool watcher = s >= nsw;
	const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(xs, 0, 0x7fffffff, kBufferWord3);
	int known_b = 0, known_a = 0;
	// this thread's point of a step (if any), and its coefficients and r,
	// loaded a step ahead
	T a[27], rv = (T)0;
	auto point = [&](int t, int &i, int &j) -> bool {
		const int u = t - kFlowLag * s;
		if (watcher || u < 0 || u > umax) return false;
		j = max(0, (u - (S.m - 1) + 1) >> 1) + jj;
		i = u - 2 * j;
		return j <= min(S.n - 1, u >> 1);
	};
	auto fetch = [&](int t) {
		int i, j;
		if (!point(t, i, j)) return;
		const T *Ar = As + saidx3(S, 0, i, j, k);
#pragma unroll
		for (int e = 0; e < 27; e++) a[e] = Ar[e * S.R];
		rv = rs[sidx3(S, i, j, k)];
	};
	fetch(0);
	if (tid == 0) stop = 0;
	__syncthreads();
	int tau = 0;
	for (; tau < nsteps; tau++) {
		if (watcher && jj == 0) {
			// polls only when the counts last seen are short
			int nb, na;
			flow_needs(tau, nsw, umax, &nb, &na);
			bool ok = true;
			if (k > 0) ok = flow_wait(prog + k - 1, nb, known_b, err, limit);
			if (ok && k + 1 < S.o) ok = flow_wait(prog + k + 1, na, known_a, err, limit);
			if (!ok) stop = 1;
		}
		__syncthreads();
		if (stop) break;
		int i = 0, j = 0;
		const bool live = point(tau, i, j);
		long long p = 0;
		T val = (T)0;
		if (live) {
			p = sidx3(S, i, j, k);
			T v[27];
#pragma unroll
			for (int e = 0; e < 27; e++) {
				if (e == 13) continue;
				const int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
				if (dk == 0 && di + 2 * dj < 0) {
					// rows u-1 .. u-3 of this plane and sweep: from the ring,
					// or the halo outside the grid
					const int ii = i + di, jn = j + dj, un = ii + 2 * jn;
					if (ii >= 0 && ii < S.m && jn >= 0)
						v[e] = ring[(s * 4 + (un & 3)) * S.R + (jn - skew_jf(S, un))];
					else if (PRD)
						v[e] = ld_x<T>(xr, (int)(sidx3(S, ii, jn, k) * (long long)sizeof(T)), false);
					else
						v[e] = (T)0;
					continue;
				}
				v[e] = ld_x<T>(xr, (int)(sidx3(S, i + di, j + dj, k + dk) * (long long)sizeof(T)), dk != 0);
			}
			T ax = (T)0;
#pragma unroll
			for (int e = 0; e < 27; e++) {
				if (e == 13) continue;
				ax = ax + a[e] * v[e];
			}
			val = (rv - ax) * ((T)1 / a[13]);
		}
		// this step's loads are back, so the previous step's stores (issued
		// before them) are complete, in every wave after the barrier
		wait_vm_all();
		__syncthreads();
		if (watcher && jj == 0 && tau > 0) st_agent(prog + k, tau);
		if (live) {
			st_x<T>(xr, (int)(p * (long long)sizeof(T)), val);
			const int u = i + 2 * j;
			ring[(s * 4 + (u & 3)) * S.R + (j - skew_jf(S, u))] = val;
		}
		fetch(tau + 1);
	}
	wait_vm_all();
	__syncthreads();
	if (watcher && jj == 0) {
		if (tau == nsteps) st_agent(prog + k, nsteps);
		// this plane's last count must be in place before it counts itself
		// out, or it could land after the last workgroup's reset
		wait_vm_all();
		const int old = __hip_atomic_fetch_add(prog + S.o, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (old == (int)gridDim.x - 1)
			for (int q = 0; q <= S.o; q++) st_agent(prog + q, 0);
	}
}

struct FlowOccupancy {
	int cus;
	long long lds;   // LDS bytes per CU
	int per_cu[2][17];  // resident workgroups per CU of 64*w threads (no dynamic LDS), [PRD]
};
template <typename T>
FlowOccupancy flow_occupancy() {
	FlowOccupancy f;
	int dev = 0;
	hipDeviceProp_t pr;
	const bool ok = hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&pr, dev) == hipSuccess;
	f.cus = ok ? pr.multiProcessorCount : 0;
	f.lds = ok ? (long long)pr.maxSharedMemoryPerMultiProcessor : 0;
	for (int q = 0; q < 2; q++) {
		f.per_cu[q][0] = 0;
		for (int w = 1; w <= 16; w++) {
			const hipError_t e = 64 * w > kFlowMaxThreads ? hipErrorInvalidValue
				: q ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&f.per_cu[q][w], k_gs_flow<T, true>, 64 * w, 0)
				    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&f.per_cu[q][w], k_gs_flow<T, false>, 64 * w, 0);
			if (e != hipSuccess) {
				f.per_cu[q][w] = 0;
				(void)hipGetLastError();
			}
		}
	}
	return f;
}

// Dynamic LDS of a k_gs_flow launch: the row ring.
inline size_t flow_lds_bytes(const SkewView &S, int nsw, size_t elem) { return (size_t)nsw * 4 * S.R * elem; }

// Most sweeps per launch of k_gs_flow for a level (0: the level cannot
// take it): the block must fit kFlowMaxThreads, and every plane's workgroup
// must be resident at once (see k_gs_flow).
template <typename T>
int flow_max_sweeps(const SkewView &S, bool prd) {
	const int rpad = flow_row_pad(S);
	if (rpad + 64 > kFlowMaxThreads || S.o < 1) return 0;
	// x is addressed with 32-bit byte offsets
	if (S.size * (long long)sizeof(T) >= (1LL << 31)) return 0;
	// Each plane trails the one below by about five steps (four rows of the
	// hyperplane offset t = u + 4k, one for the count published a step
	// late), so a level of o planes takes about umax + 5(o - 1) steps of a
	// few microseconds each.  Up to kFlowMaxPlanes planes that beats one
	// launch per hyperplane; beyond it (level 0 of 257^3: 11.4 ms per
	// two-sweep call against 10.1 ms) the hyperplane launches are kept.
	if (S.o > kFlowMaxPlanes) return 0;
	// resident workgroups of each block size, queried once (the queries
	// take long against a sweep); initialised once even with several
	// host threads
	static const FlowOccupancy occ = flow_occupancy<T>();
	auto resident = [&](int nsw) {
		// workgroups per CU by threads and by LDS (rounded up to 1 KiB)
		long long per = occ.per_cu[prd ? 1 : 0][nsw * rpad / 64 + 1];
		const long long lds = ((long long)flow_lds_bytes(S, nsw, sizeof(T)) + 64 + 1023) / 1024 * 1024;
		if (occ.lds > 0) per = min(per, occ.lds / lds);
		return per * occ.cus;
	};
	int nsw = min(16, (kFlowMaxThreads - 64) / rpad);
	while (nsw > 0 && resident(nsw) < S.o) nsw--;
	return nsw;
}

template <typename T>
bool launch_gs_flow(hipStream_t s, const SkewView &S, bool prd, const T *As, T *xs, const T *rs, int nsweeps,
		    int *prog, int *err, int limit) {
	const int smax = flow_max_sweeps<T>(S, prd);
	if (smax < 1) return false;
	const int rpad = flow_row_pad(S);
	for (int done = 0; done < nsweeps; done += smax) {
		const int nsw = min(smax, nsweeps - done);
		const size_t lds = flow_lds_bytes(S, nsw, sizeof(T));
		if (prd)
			hipLaunchKernelGGL((k_gs_flow<T, true>), dim3(S.o), dim3(nsw * rpad + 64), lds, s, S, As, xs, rs, nsw,
					   rpad, prog, err, limit);
		else
			hipLaunchKernelGGL((k_gs_flow<T, false>), dim3(S.o), dim3(nsw * rpad + 64), lds, s, S, As, xs, rs, nsw,
					   rpad, prog, err, limit);
		HIPCHECK_LAUNCH();
	}
	return true;
}

template <typename T>
__global__ __launch_bounds__(256) void k_skew_pack_op(LevelView L, SkewView S, const T *__restrict__ A,
						      T *__restrict__ As) {
	const long long N = (long long)L.m * L.n * L.o;
	for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < N * 27; q += (long long)gridDim.x * 256) {
		const int e = (int)(q / N);
		const long long p = q - (long long)e * N;
		const int i = (int)(p % L.m), j = (int)((p / L.m) % L.n), k = (int)(p / ((long long)L.m * L.n));
		As[saidx3(S, e, i, j, k)] = A[aidx3(L, e, i, j, k)];
	}
}

template <typename T>
void launch_skew_pack_op(hipStream_t s, const LevelView &L, const SkewView &S, const T *A, T *As) {
	hipLaunchKernelGGL(k_skew_pack_op<T>, dim3(4096), dim3(256), 0, s, L, S, A, As);
	HIPCHECK_LAUNCH();
}

// Ghost copies of a periodic level's halo on its skewed copy (as
// k_halo_wrap on the level layout).
template <typename T>
__global__ __launch_bounds__(256) void k_skew_halo_wrap(LevelView L, SkewView S, T *__restrict__ xs) {
	const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
	int i, j, k, si, sj, sk;
	if (q >= halo_points(L) || !halo_src(L, q, i, j, k, si, sj, sk)) return;
	xs[sidx3(S, i, j, k)] = xs[sidx3(S, si, sj, sk)];
}

template <typename T>
void launch_skew_halo_wrap(hipStream_t s, const LevelView &L, const SkewView &S, T *xs) {
	if (!L.prd) return;
	hipLaunchKernelGGL(k_skew_halo_wrap<T>, dim3((unsigned)((halo_points(L) + 255) / 256)), dim3(256), 0, s, L, S, xs);
	HIPCHECK_LAUNCH();
}

// A vector between the level layout and the plane-skewed one, through LDS:
// a tile of 64 i by 16 j of one k-plane per workgroup.  Level rows are read
// (or written) 64 points at a time, and the tile's part of each skewed row
// u = i + 2j, up to 16 consecutive j, is written (or read) as one run.
// (Copied point by point, consecutive i are a skewed row apart, so every
// access on that side touched a line of its own.)
constexpr int kSkewTileI = 64, kSkewTileJ = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_skew_tile(LevelView L, SkewView S, const T *__restrict__ v, T *__restrict__ vs,
						   int to_skew) {
	__shared__ T tile[kSkewTileJ][kSkewTileI + 1];
	const int i0 = (int)blockIdx.x * kSkewTileI, j0 = (int)blockIdx.y * kSkewTileJ, k = (int)blockIdx.z;
	const int tid = (int)threadIdx.x;
	// skewed rows the tile touches, and the (u, jj) slots of the skewed side
	const int ulo = i0 + 2 * j0;
	const int slots = (kSkewTileI - 1 + 2 * (kSkewTileJ - 1) + 1) * kSkewTileJ;
	if (to_skew) {
		for (int q = tid; q < kSkewTileI * kSkewTileJ; q += 256) {
			const int jj = q / kSkewTileI, ii = q % kSkewTileI;
			const int i = i0 + ii, j = j0 + jj;
			if (i < L.m && j < L.n) tile[jj][ii] = v[idx3(L, i, j, k)];
		}
		__syncthreads();
		for (int q = tid; q < slots; q += 256) {
			const int u = ulo + q / kSkewTileJ, jj = q % kSkewTileJ;
			const int j = j0 + jj, i = u - 2 * j;
			if (i >= i0 && i < i0 + kSkewTileI && i < L.m && j < L.n) vs[sidx3(S, i, j, k)] = tile[jj][i - i0];
		}
	} else {
		for (int q = tid; q < slots; q += 256) {
			const int u = ulo + q / kSkewTileJ, jj = q % kSkewTileJ;
			const int j = j0 + jj, i = u - 2 * j;
			if (i >= i0 && i < i0 + kSkewTileI && i < L.m && j < L.n) tile[jj][i - i0] = v[sidx3(S, i, j, k)];
		}
		__syncthreads();
		for (int q = tid; q < kSkewTileI * kSkewTileJ; q += 256) {
			const int jj = q / kSkewTileI, ii = q % kSkewTileI;
			const int i = i0 + ii, j = j0 + jj;
			if (i < L.m && j < L.n) vs[idx3(L, i, j, k)] = tile[jj][ii];
		}
	}
}

// to_skew: vs (skewed) = v (level layout); else vs (level layout) = v (skewed).
template <typename T>
void launch_skew_vec(hipStream_t s, const LevelView &L, const SkewView &S, const T *v, T *vs, bool to_skew) {
	const dim3 grid((L.m + kSkewTileI - 1) / kSkewTileI, (L.n + kSkewTileJ - 1) / kSkewTileJ, L.o);
	hipLaunchKernelGGL(k_skew_tile<T>, grid, dim3(256), 0, s, L, S, v, vs, to_skew ? 1 : 0);
	HIPCHECK_LAUNCH();
}

// A small level smoothed by one workgroup: all sweeps of one call in one
// launch, ordered by workgroup barriers.  Exact mode walks the hyperplanes;
// colour mode the eight classes.
template <typename T>
__device__ __forceinline__ void gs_small_body(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
					      const T *__restrict__ r, int nsweeps, int mode, int exact) {
	const int nt = blockDim.x, tid = threadIdx.x;
	const int tmax = (L.m - 1) + 2 * (L.n - 1) + 4 * (L.o - 1);
	const int npair = L.n * L.o;
	for (int sw = 0; sw < nsweeps; sw++) {
		for (int pass = 0; pass < (mode ? 2 : 1); pass++) {
			// periodic axes: ghost copies before each pass (region.cc:361, 457)
			halo_wrap_block<T>(L, x, L.base, L.px, L.pxy);
			if (exact) {
				for (int tt = 0; tt <= tmax; tt++) {
					const int t = pass ? tmax - tt : tt;
					for (int q = tid; q < npair; q += nt) {
						const int j = q % L.n, k = q / L.n;
						const int i = t - 2 * j - 4 * k;
						if (i >= 0 && i < L.m) gs_point<T>(L, A, x, r, i, j, k);
					}
					__syncthreads();
				}
			} else {
				for (int cc = 0; cc < 8; cc++) {
					const int c = pass ? 7 - cc : cc;
					const int ci = c & 1, cj = (c >> 1) & 1, ck = (c >> 2) & 1;
					const int na = (L.m - ci + 1) / 2, nb = (L.n - cj + 1) / 2, nd = (L.o - ck + 1) / 2;
					const int tot = na * nb * nd;
					for (int q = tid; q < tot; q += nt) {
						const int a = q % na, b = (q / na) % nb, d = q / (na * nb);
						gs_point<T>(L, A, x, r, 2 * a + ci, 2 * b + cj, 2 * d + ck);
					}
					__syncthreads();
				}
			}
		}
	}
}

// The same sweeps with x staged in LDS (small_lds_fits): x, halo included,
// is copied into LDS in a compact layout (pitches m+2 and (m+2)(n+2)), every
// sweep of the call reads and writes it there, and the interior is written
// back at the end, so a colour class (or hyperplane) costs one coefficient
// fetch from L2 and a barrier, not a round trip to memory for x.  A level of
// at most blockDim points keeps each thread's own point's 27 coefficients
// and r in registers for the whole call (points numbered colour class by
// colour class in colour mode, lexicographically in exact mode).  Per-point
// arithmetic is gs_point's, so the result is bit-identical.
template <typename T>
__device__ __forceinline__ void gs_lds_update(const T *__restrict__ c, T rv, T *__restrict__ xs, int lq, int hx, int hxy) {
	T ax = (T)0;
#pragma unroll
	for (int e = 0; e < 27; e++) {
		if (e == 13) continue;
		const int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
		ax = ax + c[e] * xs[lq + di + dj * hx + dk * hxy];
	}
	xs[lq] = (rv - ax) * ((T)1 / c[13]);
}

template <typename T>
__device__ __forceinline__ void gs_lds_global(const LevelView &L, const T *__restrict__ A, const T *__restrict__ r,
					      T *__restrict__ xs, int i, int j, int k, int hx, int hxy) {
	const long long p = idx3(L, i, j, k);
	const T *Ap = A + aidx3(L, 0, i, j, k);
	T c[27];
#pragma unroll
	for (int e = 0; e < 27; e++) c[e] = Ap[e * L.pa];
	gs_lds_update<T>(c, r[p], xs, (i + 1) + hx * (j + 1) + hxy * (k + 1), hx, hxy);
}

// One colour pair (0,cj,ck), (1,cj,ck) of a small level with x in LDS, one
// point pair (i0, i0+1), i0 even, per thread (at most blockDim pairs).
// Forward (rev = 0) updates the even point first, reverse the odd one; the
// second point's coefficients, loaded with the first's, are kept in
// registers over the barrier between the two classes.  The caller's barrier
// follows.
// One colour pair (0,cj,ck), (1,cj,ck) of a small level with x in LDS, one
// point pair (i0, i0+1), i0 even, per thread (at most blockDim pairs): the
// even point is updated from vector loads of both points' coefficients, the
// odd point's are kept in registers over the barrier between the two
// classes, then it is updated.  The caller's barrier follows.
template <typename T>
__device__ __forceinline__ void gs_lds_pair_pass(const LevelView &L, const T *__restrict__ A, const T *__restrict__ r,
						 T *__restrict__ xs, int pq, int hx, int hxy) {
	typedef typename Vec2<T>::type V2;
	const int cj = pq & 1, ck = pq >> 1;
	const int na = (L.m + 1) / 2, nb = (L.n - cj + 1) / 2, nd = (L.o - ck + 1) / 2;
	const int q = threadIdx.x;
	const bool live = q < na * nb * nd;
	T c2[27], r2 = (T)0;
	int lq = 0;
	bool two = false;
	if (live) {
		const int i0 = 2 * (q % na), j = 2 * ((q / na) % nb) + cj, k = 2 * (q / (na * nb)) + ck;
		two = i0 + 1 < L.m;
		const long long p = idx3(L, i0, j, k);
		const T *Ap = A + aidx3(L, 0, i0, j, k);
		lq = (i0 + 1) + hx * (j + 1) + hxy * (k + 1);
		const V2 rv = *(const V2 *)(r + p);
		r2 = rv.y;
		// the even point's sum, in box order, as the planes arrive
		T ax = (T)0, a13 = (T)1;
#pragma unroll
		for (int e = 0; e < 27; e++) {
			const V2 a = *(const V2 *)(Ap + e * L.pa);
			c2[e] = a.y;
			if (e == 13) { a13 = a.x; continue; }
			const int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
			ax = ax + a.x * xs[lq + di + dj * hx + dk * hxy];
		}
		xs[lq] = (rv.x - ax) * ((T)1 / a13);
	}
	__syncthreads();
	if (live && two) gs_lds_update<T>(c2, r2, xs, lq + 1, hx, hxy);
}

template <typename T, bool resident>
__device__ __forceinline__ void gs_lds_run(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
					   const T *__restrict__ r, int nsweeps, int mode, int exact, T *__restrict__ xs) {
	const int nt = blockDim.x, tid = threadIdx.x;
	const int hx = L.m + 2, hxy = hx * (L.n + 2), nh = hxy * (L.o + 2);
	const int npts = L.m * L.n * L.o;
	for (int q = tid; q < nh; q += nt) {
		const int i = q % hx - 1, j = (q / hx) % (L.n + 2) - 1, k = q / hxy - 1;
		xs[q] = x[idx3(L, i, j, k)];
	}
	// resident: this thread's point, its colour or hyperplane, coefficients, r
	bool own = false;
	int oc = -1, ot = -1, olq = 0;
	T ca[27], cr = (T)0;
	if (resident && tid < npts) {
		int i = 0, j = 0, k = 0;
		if (exact) {
			i = tid % L.m; j = (tid / L.m) % L.n; k = tid / (L.m * L.n);
		} else {
			int q = tid, c = 0;
			for (; c < 8; c++) {
				const int na = (L.m - (c & 1) + 1) / 2, nb = (L.n - ((c >> 1) & 1) + 1) / 2, nd = (L.o - (c >> 2) + 1) / 2;
				const int cnt = na * nb * nd;
				if (q < cnt) {
					i = 2 * (q % na) + (c & 1); j = 2 * ((q / na) % nb) + ((c >> 1) & 1); k = 2 * (q / (na * nb)) + (c >> 2);
					break;
				}
				q -= cnt;
			}
			oc = c;
		}
		own = true;
		ot = i + 2 * j + 4 * k;
		olq = (i + 1) + hx * (j + 1) + hxy * (k + 1);
		const long long p = idx3(L, i, j, k);
		const T *Ap = A + aidx3(L, 0, i, j, k);
#pragma unroll
		for (int e = 0; e < 27; e++) ca[e] = Ap[e * L.pa];
		cr = r[p];
	}
	__syncthreads();
	const int tmax = (L.m - 1) + 2 * (L.n - 1) + 4 * (L.o - 1);
	const int npair = L.n * L.o;
	// forward colour sweeps: classes (0,cj,ck) and (1,cj,ck) as one pass
	const bool staged = !resident && !exact && !mode && small_pairs_per_pass(L) <= nt;
	if (resident && exact && !mode && !L.prd) {
		// Forward sweeps pipelined eight hyperplanes apart: sweep s updates
		// hyperplane t at step t + 8s.  A neighbour moves t by 1 to 7, so
		// when a point updates, its earlier neighbours hold this sweep's
		// values and its later ones the previous sweep's, as in separate
		// sweeps.  (The ghosts of a periodic level must hold each sweep's
		// starting values, so periodic levels sweep one at a time.)
		const int steps = tmax + 1 + 8 * (nsweeps - 1);
#pragma unroll 1
		for (int st = 0; st < steps; st++) {
			const int d = st - ot;
			if (own && d >= 0 && (d & 7) == 0 && (d >> 3) < nsweeps) gs_lds_update<T>(ca, cr, xs, olq, hx, hxy);
			__syncthreads();
		}
	} else if (staged) {
#pragma unroll 1
		for (int sw = 0; sw < nsweeps; sw++) {
			halo_wrap_block<T>(L, xs, 1 + hx + hxy, hx, hxy);
#pragma unroll 1
			for (int pp = 0; pp < 4; pp++) {
				gs_lds_pair_pass<T>(L, A, r, xs, pp, hx, hxy);
				__syncthreads();
			}
		}
	} else {
#pragma unroll 1
	for (int sw = 0; sw < nsweeps; sw++) {
#pragma unroll 1
		for (int pass = 0; pass < (mode ? 2 : 1); pass++) {
			halo_wrap_block<T>(L, xs, 1 + hx + hxy, hx, hxy);
			if (exact) {
#pragma unroll 1
				for (int tt = 0; tt <= tmax; tt++) {
					const int t = pass ? tmax - tt : tt;
					if (resident) {
						if (own && ot == t) gs_lds_update<T>(ca, cr, xs, olq, hx, hxy);
					} else {
						for (int q = tid; q < npair; q += nt) {
							const int j = q % L.n, k = q / L.n;
							const int i = t - 2 * j - 4 * k;
							if (i >= 0 && i < L.m) gs_lds_global<T>(L, A, r, xs, i, j, k, hx, hxy);
						}
					}
					__syncthreads();
				}
			} else {
#pragma unroll 1
				for (int cc = 0; cc < 8; cc++) {
					const int c = pass ? 7 - cc : cc;
					if (resident) {
						if (own && oc == c) gs_lds_update<T>(ca, cr, xs, olq, hx, hxy);
					} else {
						const int ci = c & 1, cj = (c >> 1) & 1, ck = (c >> 2) & 1;
						const int na = (L.m - ci + 1) / 2, nb = (L.n - cj + 1) / 2, nd = (L.o - ck + 1) / 2;
						const int tot = na * nb * nd;
						for (int q = tid; q < tot; q += nt) {
							const int a = q % na, b = (q / na) % nb, d = q / (na * nb);
							gs_lds_global<T>(L, A, r, xs, 2 * a + ci, 2 * b + cj, 2 * d + ck, hx, hxy);
						}
					}
					__syncthreads();
				}
			}
		}
	}
	}
	for (int q = tid; q < npts; q += nt) {
		const int i = q % L.m, j = (q / L.m) % L.n, k = q / (L.m * L.n);
		x[idx3(L, i, j, k)] = xs[(i + 1) + hx * (j + 1) + hxy * (k + 1)];
	}
}

template <typename T>
__device__ __forceinline__ void gs_lds_body(const LevelView &L, const T *__restrict__ A, T *__restrict__ x,
					    const T *__restrict__ r, int nsweeps, int mode, int exact, T *__restrict__ xs) {
	if ((long long)L.m * L.n * L.o <= (long long)blockDim.x)
		gs_lds_run<T, true>(L, A, x, r, nsweeps, mode, exact, xs);
	else
		gs_lds_run<T, false>(L, A, x, r, nsweeps, mode, exact, xs);
}

template <typename T>
__global__ __launch_bounds__(1024) void k_gs_small_lds(LevelView L, const T *__restrict__ A, T *__restrict__ x,
						       const T *__restrict__ r, int nsweeps, int mode, int exact) {
	extern __shared__ char smem_[];
	gs_lds_body<T>(L, A, x, r, nsweeps, mode, exact, (T *)smem_);
}

template <typename T>
__global__ __launch_bounds__(1024) void k_gs_small(LevelView L, const T *__restrict__ A, T *__restrict__ x,
						   const T *__restrict__ r, int nsweeps, int mode, int exact) {
	gs_small_body<T>(L, A, x, r, nsweeps, mode, exact);
}

template <typename T>
void launch_gs_small(hipStream_t s, const LevelView &L, const T *A, T *x, const T *r, int nsweeps, int mode, int exact) {
	const long long lds = small_lds_bytes(L, (int)sizeof(T));
	if (lds <= kSmallLdsBytes)
		hipLaunchKernelGGL(k_gs_small_lds<T>, dim3(1), dim3(1024), (unsigned)lds, s, L, A, x, r, nsweeps, mode, exact);
	else
		hipLaunchKernelGGL(k_gs_small<T>, dim3(1), dim3(1024), 0, s, L, A, x, r, nsweeps, mode, exact);
	HIPCHECK_LAUNCH();
}

// ------------------------------------------------------ residual, norms

// Deterministic block sum of 256 values (4 wavefronts of 64): a fixed
// shuffle tree inside each wavefront (lane l adds lane l + w, w = 32 .. 1),
// then the four wavefront sums as (s0 + s1) + (s2 + s3): two barriers
// instead of the nine of a 256-slot LDS tree.  Every thread gets the sum,
// and the LDS slots may be reused as soon as it returns.
__device__ __forceinline__ double block_sum_256(double v) {
	__shared__ double red[4];
	const int tid = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
	for (int w = 32; w > 0; w >>= 1) v = v + __shfl_down(v, w, 64);
	if ((tid & 63) == 0) red[tid >> 6] = v;
	__syncthreads();
	const double s = (red[0] + red[1]) + (red[2] + red[3]);
	__syncthreads();
	return s;
}

// Residual r - A x and per-block sums of its square, on rows longer than
// one wavefront's lanes: one wavefront per segment seg of grid row q (see
// residual_segments), each lane a pair of points (FP64) or four (FP32) read
// with vector loads (as in k_gs_pair, with the same uniform row bases);
// blocks of 4 wavefronts, renumbered per XCD.
// res (may be null) receives the residual in type TR, partial (may be null)
// one sum per block.  q and seg must be wavefront-uniform.
template <typename T, typename TR>
__device__ __forceinline__ double residual_row_pair(const LevelView &L, const T *__restrict__ A, const T *__restrict__ x,
						    const T *__restrict__ r, TR *__restrict__ res, int q, int seg,
						    TR *__restrict__ zx = nullptr) {
	typedef typename Vec2<T>::type V2;
	typedef typename Vec2<TR>::type R2;
	const int lane = threadIdx.x;
	const long long px = L.px, pxy = L.pxy, pa = L.pa;
	double v2 = 0.0;
	if (q < L.n * L.o) {
		const int j = q % L.n, k = q / L.n;
		const long long row = idx3(L, 0, j, k);
		const T *Ar = A + arow0(L, j, k), *xr = x + row, *rr_ = r + row;
		TR *resr = res ? res + row : nullptr;
		TR *zxr = zx ? zx + row : nullptr;
		// segments of 128 points; the last one takes up to 129, lane 63
		// doing the odd point out.  The wavefront does segment seg only; its
		// stores go out after its sum, through the same flush as before.
		bool pend = false, ptwo = false, pext = false;
		uint32_t pobr = 0;
		T p0 = (T)0, p1 = (T)0, p2 = (T)0;
		auto flush = [&]() {
			if (!pend) return;
			if (resr) {
				if (ptwo) { R2 o; o.x = (TR)p0; o.y = (TR)p1; st_row<R2>(resr, pobr, o); }
				else st_row<TR>(resr, pobr, (TR)p0);
				if (pext) st_row<TR>(resr, pobr + 2u * (uint32_t)sizeof(TR), (TR)p2);
			}
			if (zxr) {
				if (ptwo) { R2 o; o.x = (TR)0; o.y = (TR)0; st_row<R2>(zxr, pobr, o); }
				else st_row<TR>(zxr, pobr, (TR)0);
				if (pext) st_row<TR>(zxr, pobr + 2u * (uint32_t)sizeof(TR), (TR)0);
			}
		};
		for (int s0 = 128 * seg;; s0 += 128) {
			const bool last = L.m - s0 <= 129;
			const int i0 = s0 + 2 * lane;
			const bool zseg = last && L.m - s0 == 129;  // the segment has a point out, m-1
			const bool ext = zseg && lane == 63;         // lane 63 does it
			const bool live = i0 < L.m;
			const bool two = i0 + 1 < L.m;
			T ax = (T)0, ay = (T)0, az = (T)0, r0 = (T)0, r1 = (T)0, rz = (T)0;
			const uint32_t ob = (uint32_t)i0 * (uint32_t)sizeof(T);
			// lanes 0..26: coefficient `lane` of the point out, read by every
			// lane with v_readlane (its sum is formed with no branch; inside
			// lane 63's branch the products made the wavefront wait for each
			// row step's loads)
			T zco = (T)0;
			if (zseg && lane < 27) zco = ld_row<T>(Ar + lane * pa, (uint32_t)(L.m - 1) * (uint32_t)sizeof(T));
			if (live) {
				const V2 rr = ld_stream<V2>(rr_, ob);
				r0 = rr.x; r1 = rr.y;
				if (ext) rz = ld_row<T>(rr_, ob + 2u * (uint32_t)sizeof(T));
#pragma unroll
				for (int rq = 0; rq < 9; rq++) {
					const int dj = rq % 3 - 1, dk = rq / 3 - 1;
					const T *xq = xr + dj * px + dk * pxy;
					const V2 am = ld_stream<V2>(Ar + (3 * rq) * pa, ob);
					const V2 a0 = ld_stream<V2>(Ar + (3 * rq + 1) * pa, ob);
					const V2 ap = ld_stream<V2>(Ar + (3 * rq + 2) * pa, ob);
					const V2 xv = ld_row<V2>(xq, ob);
					const T xm = ld_row<T>(xq - 1, ob);
					const T xp = two ? ld_row<T>(xq + 2, ob) : (T)0;
					ax = ax + am.x * xm;
					ax = ax + a0.x * xv.x;
					ax = ax + ap.x * xv.y;
					ay = ay + am.y * xv.x;
					ay = ay + a0.y * xv.y;
					ay = ay + ap.y * xp;
					// the point out (lane 63's az is the one used)
					az = az + read_lane(zco, 3 * rq) * xv.y;
					az = az + read_lane(zco, 3 * rq + 1) * xp;
					az = az + read_lane(zco, 3 * rq + 2) * (T)0;  // the halo point i = m
				}
			}
			flush();
			if (live) {
				const T rv0 = r0 - ax, rv1 = r1 - ay, rv2 = rz - az;
				v2 = v2 + (double)rv0 * (double)rv0;
				if (two) v2 = v2 + (double)rv1 * (double)rv1;
				if (ext) v2 = v2 + (double)rv2 * (double)rv2;
				p0 = rv0; p1 = rv1; p2 = rv2;
			}
			pend = live; ptwo = two; pext = ext;
			pobr = (uint32_t)i0 * (uint32_t)sizeof(TR);
			break;  // one segment per wavefront
		}
		flush();
	}
	return v2;
}

// FP32 residual with four points per lane (16-byte loads): segments of 256
// points, the last one up to 257 with lane 63 doing the point out; the
// wavefront does segment seg of row q.
__device__ __forceinline__ double residual_row_quad_f32(const LevelView &L, const float *__restrict__ A,
							const float *__restrict__ x, const float *__restrict__ r,
							float *__restrict__ res, int q, int seg) {
	const int lane = threadIdx.x;
	const long long px = L.px, pxy = L.pxy, pa = L.pa;
	double v2 = 0.0;
	if (q < L.n * L.o) {
		const int j = q % L.n, k = q / L.n;
		const long long row = idx3(L, 0, j, k);
		const float *Ar = A + arow0(L, j, k), *xr = x + row, *rr_ = r + row;
		float *resr = res ? res + row : nullptr;
		// the segment's stores go out after its sum
		int pnv = 0;
		bool pext = false;
		uint32_t pob = 0;
		float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f, q3 = 0.0f, qz = 0.0f;
		auto flush = [&]() {
			if (!resr || pnv == 0) return;
			if (pnv == 4) st_row<float4>(resr, pob, make_float4(q0, q1, q2, q3));
			else {
				st_row<float>(resr, pob, q0);
				if (pnv > 1) st_row<float>(resr, pob + 4u, q1);
				if (pnv > 2) st_row<float>(resr, pob + 8u, q2);
			}
			if (pext) st_row<float>(resr, pob + 16u, qz);
		};
		for (int s0 = 256 * seg;; s0 += 256) {
			const bool last = L.m - s0 <= 257;
			const int i0 = s0 + 4 * lane;
			const bool zseg = last && L.m - s0 == 257;  // the segment has a point out, m-1
			const bool ext = zseg && lane == 63;         // lane 63 does it
			const int nv = i0 < L.m ? min(4, L.m - i0) : 0;
			const uint32_t ob = (uint32_t)i0 * 4u;
			// lanes 0..26: coefficient `lane` of the point out, read by every
			// lane with v_readlane (its sum is formed with no branch, as in
			// residual_row_pair)
			float zco = 0.0f;
			if (zseg && lane < 27) zco = ld_row<float>(Ar + lane * pa, (uint32_t)(L.m - 1) * 4u);
			float a0s = 0.0f, a1s = 0.0f, a2s = 0.0f, a3s = 0.0f, azs = 0.0f;
			float4 rr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			float rz = 0.0f;
			if (nv > 0) {
				rr = ld_stream<float4>(rr_, ob);
				rz = ext ? ld_row<float>(rr_, ob + 16u) : 0.0f;
#pragma unroll
				for (int rq = 0; rq < 9; rq++) {
					const int dj = rq % 3 - 1, dk = rq / 3 - 1;
					const float *xq = xr + dj * px + dk * pxy;
					const float4 am = ld_stream<float4>(Ar + (3 * rq) * pa, ob);
					const float4 ac = ld_stream<float4>(Ar + (3 * rq + 1) * pa, ob);
					const float4 ap = ld_stream<float4>(Ar + (3 * rq + 2) * pa, ob);
					const float4 xv = ld_row<float4>(xq, ob);
					const float xm = ld_row<float>(xq - 1, ob);
					const float xp = nv == 4 ? ld_row<float>(xq + 4, ob) : 0.0f;
					a0s = a0s + am.x * xm;
					a0s = a0s + ac.x * xv.x;
					a0s = a0s + ap.x * xv.y;
					a1s = a1s + am.y * xv.x;
					a1s = a1s + ac.y * xv.y;
					a1s = a1s + ap.y * xv.z;
					a2s = a2s + am.z * xv.y;
					a2s = a2s + ac.z * xv.z;
					a2s = a2s + ap.z * xv.w;
					a3s = a3s + am.w * xv.z;
					a3s = a3s + ac.w * xv.w;
					a3s = a3s + ap.w * xp;
					// the point out (lane 63's azs is the one used)
					azs = azs + read_lane(zco, 3 * rq) * xv.w;
					azs = azs + read_lane(zco, 3 * rq + 1) * xp;
					azs = azs + read_lane(zco, 3 * rq + 2) * 0.0f;  // halo
				}
			}
			flush();
			if (nv > 0) {
				const float v0 = rr.x - a0s, v1 = rr.y - a1s, v2f = rr.z - a2s, v3 = rr.w - a3s, vz = rz - azs;
				v2 = v2 + (double)v0 * (double)v0;
				if (nv > 1) v2 = v2 + (double)v1 * (double)v1;
				if (nv > 2) v2 = v2 + (double)v2f * (double)v2f;
				if (nv > 3) v2 = v2 + (double)v3 * (double)v3;
				if (ext) v2 = v2 + (double)vz * (double)vz;
				q0 = v0; q1 = v1; q2 = v2f; q3 = v3; qz = vz;
			}
			pnv = nv; pext = ext; pob = ob;
			break;  // one segment per wavefront
		}
		flush();
	}
	return v2;
}

// Residual rows, dispatched on the arithmetic type: four points per lane in
// FP32 (residual into FP32), two in FP64.
template <typename T, typename TR>
__device__ __forceinline__ double residual_row(const LevelView &L, const T *__restrict__ A, const T *__restrict__ x,
					       const T *__restrict__ r, TR *__restrict__ res, int q, int seg,
					       TR *__restrict__ zx = nullptr) {
	return residual_row_pair<T, TR>(L, A, x, r, res, q, seg, zx);
}
template <>
__device__ __forceinline__ double residual_row<float, float>(const LevelView &L, const float *__restrict__ A,
							     const float *__restrict__ x, const float *__restrict__ r,
							     float *__restrict__ res, int q, int seg, float *__restrict__ zx) {
	return residual_row_quad_f32(L, A, x, r, res, q, seg);  // zx: not used in FP32
}

// Wavefront w of a residual launch: rows w*rpw .. when they fit the lanes,
// else one segment of a long row (residual_segments per row).  w must be
// wavefront-uniform.
template <typename T, typename TR>
__device__ __forceinline__ double residual_wave(const LevelView &L, const T *__restrict__ A, const T *__restrict__ x,
						const T *__restrict__ r, TR *__restrict__ res, TR *__restrict__ zx, int w) {
	const int rows = L.n * L.o;
	const RowLanes W = row_lanes(L.m, (int)sizeof(T));
	if (W.lpr <= 64) {
		const int q0 = w * W.rpw;
		if (q0 >= rows) return 0.0;
		if constexpr (sizeof(T) == 8) return residual_lanes<T, TR, Ppl<T>::v, PO_ROWS>(L, A, x, r, res, zx, q0, rows);
		// FP32 point outs come with a tail copy (m = 4q+1)
		if (W.ext) return residual_lanes<T, TR, Ppl<T>::v, PO_TAIL>(L, A, x, r, res, zx, q0, rows);
		return residual_lanes<T, TR, Ppl<T>::v, PO_NONE>(L, A, x, r, res, zx, q0, rows);
	}
	const int nseg = residual_segments(L.m, (int)sizeof(T));
	return residual_row<T, TR>(L, A, x, r, res, w / nseg, w % nseg, zx);
}

template <typename T, typename TR>
__global__ __launch_bounds__(256) void k_residual(LevelView L, const T *__restrict__ A, const T *__restrict__ x,
						  const T *__restrict__ r, TR *__restrict__ res, double *__restrict__ partial,
						  TR *__restrict__ zx) {
	const int qb = xcd_block((int)blockIdx.x, (int)gridDim.x);
	const double v2 = residual_wave<T, TR>(L, A, x, r, res, zx, __builtin_amdgcn_readfirstlane(qb * 4 + (int)threadIdx.y));
	if (partial) {
		// indexed by the block's logical number, so the sum order is that
		// of a batched launch
		const double sblk = block_sum_256(v2);
		if (threadIdx.x == 0 && threadIdx.y == 0) partial[qb] = sblk;
	}
}

// Residual of a level with periodic axes, one point per thread, after the
// halo holds the ghost copies: the 27-entry sum in box order from zero,
// centre included (region.cc:1833-1842), as the row kernels form it.
template <typename T, typename TR>
__global__ __launch_bounds__(256) void k_residual_pt(LevelView L, const T *__restrict__ A, const T *__restrict__ x,
						     const T *__restrict__ r, TR *__restrict__ res,
						     double *__restrict__ partial, TR *__restrict__ zx) {
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y), k = (int)blockIdx.z;
	double v2 = 0.0;
	if (i < L.m && j < L.n) {
		const long long p = idx3(L, i, j, k), px = L.px, pxy = L.pxy;
		const T *Ap = A + aidx3(L, 0, i, j, k);
		T ax = (T)0;
#pragma unroll
		for (int e = 0; e < 27; e++) {
			const int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
			ax = ax + Ap[e * L.pa] * x[p + di + dj * px + dk * pxy];
		}
		const T v = r[p] - ax;
		if (res) res[p] = (TR)v;
		if (zx) zx[p] = (TR)0;
		v2 = (double)v * (double)v;
	}
	if (partial) {
		const double sblk = block_sum_256(v2);
		if (threadIdx.x == 0 && threadIdx.y == 0)
			partial[blockIdx.x + gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = sblk;
	}
}

int residual_partial_count(const LevelView &L) {
	if (L.prd) return ((L.m + 63) / 64) * ((L.n + 3) / 4) * L.o;
	return residual_blocks(L, 8);
}

template <typename T, typename TR>
int launch_residual(hipStream_t s, const LevelView &L, const T *A, const T *x, const T *r, TR *res, double *partial,
		    TR *zero_out) {
	if (L.prd) {
		// the ghost copies are not part of the vector: refreshing them
		// leaves its value unchanged
		launch_halo_wrap<T>(s, L, const_cast<T *>(x));
		dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o), block(64, 4, 1);
		hipLaunchKernelGGL((k_residual_pt<T, TR>), grid, block, 0, s, L, A, x, r, res, partial, zero_out);
		HIPCHECK_LAUNCH();
		return residual_partial_count(L);
	}
	const int G = residual_blocks(L, (int)sizeof(T));
	hipLaunchKernelGGL((k_residual<T, TR>), dim3(G), dim3(64, 4, 1), 0, s, L, A, x, r, res, partial, zero_out);
	HIPCHECK_LAUNCH();
	return G;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sumsq(LevelView L, const T *__restrict__ r, double *__restrict__ partial) {
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y), k = (int)blockIdx.z;
	double v2 = 0.0;
	if (i < L.m && j < L.n) {
		const double v = (double)r[idx3(L, i, j, k)];
		v2 = v * v;
	}
	const double sblk = block_sum_256(v2);
	if (threadIdx.x == 0 && threadIdx.y == 0)
		partial[blockIdx.x + gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = sblk;
}

template <typename T>
int launch_sumsq(hipStream_t s, const LevelView &L, const T *r, double *partial) {
	dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o), block(64, 4, 1);
	hipLaunchKernelGGL(k_sumsq<T>, grid, block, 0, s, L, r, partial);
	HIPCHECK_LAUNCH();
	return ((L.m + 63) / 64) * ((L.n + 3) / 4) * L.o;
}

// out = A in at the interior points of L (region::mul_A with an outside
// array, region.cc:1861-1869, used by region::Ax): the 27-entry sum in box
// order from zero.  in must have a zero halo on non-periodic axes and the
// ghost copies on periodic ones.
__global__ __launch_bounds__(256) void k_apply(LevelView L, const double *__restrict__ A, const double *__restrict__ in,
					       double *__restrict__ out) {
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y), k = (int)blockIdx.z;
	if (i >= L.m || j >= L.n) return;
	const long long p = idx3(L, i, j, k), px = L.px, pxy = L.pxy;
	const double *Ap = A + aidx3(L, 0, i, j, k);
	double ax = 0.0;
#pragma unroll
	for (int e = 0; e < 27; e++) {
		const int di = e % 3 - 1, dj = (e / 3) % 3 - 1, dk = e / 9 - 1;
		ax = ax + Ap[e * L.pa] * in[p + di + dj * px + dk * pxy];
	}
	out[p] = ax;
}

void launch_apply(hipStream_t s, const LevelView &L, const double *A, const double *in, double *out) {
	dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o), block(64, 4, 1);
	hipLaunchKernelGGL(k_apply, grid, block, 0, s, L, A, in, out);
	HIPCHECK_LAUNCH();
}

// Block sums over the interior points of L of a*b (diff = 0) or (a-b)^2
// (diff = 1).
__global__ __launch_bounds__(256) void k_pair_sum(LevelView L, const double *__restrict__ a, const double *__restrict__ b,
						  int diff, double *__restrict__ partial) {
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y), k = (int)blockIdx.z;
	double v = 0.0;
	if (i < L.m && j < L.n) {
		const long long p = idx3(L, i, j, k);
		if (diff) { const double d = a[p] - b[p]; v = d * d; }
		else v = a[p] * b[p];
	}
	const double sblk = block_sum_256(v);
	if (threadIdx.x == 0 && threadIdx.y == 0)
		partial[blockIdx.x + gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = sblk;
}

int launch_pair_sum(hipStream_t s, const LevelView &L, const double *a, const double *b, int diff, double *partial) {
	dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o), block(64, 4, 1);
	hipLaunchKernelGGL(k_pair_sum, grid, block, 0, s, L, a, b, diff, partial);
	HIPCHECK_LAUNCH();
	return ((L.m + 63) / 64) * ((L.n + 3) / 4) * L.o;
}

// The same block sums of (ext - x)^2 with ext given point-major
// (i + m*(j + n*k)) instead of in the level layout: the same blocks, the same
// per-thread values, so the same partial sums as k_pair_sum with diff = 1.
__global__ __launch_bounds__(256) void k_diff2_compact(LevelView L, const double *__restrict__ ext,
						       const double *__restrict__ x, double *__restrict__ partial) {
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y), k = (int)blockIdx.z;
	double v = 0.0;
	if (i < L.m && j < L.n) {
		const double d = ext[i + (long long)L.m * (j + (long long)L.n * k)] - x[idx3(L, i, j, k)];
		v = d * d;
	}
	const double sblk = block_sum_256(v);
	if (threadIdx.x == 0 && threadIdx.y == 0)
		partial[blockIdx.x + gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = sblk;
}

int launch_diff2_compact(hipStream_t s, const LevelView &L, const double *ext, const double *x, double *partial) {
	dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o), block(64, 4, 1);
	hipLaunchKernelGGL(k_diff2_compact, grid, block, 0, s, L, ext, x, partial);
	HIPCHECK_LAUNCH();
	return ((L.m + 63) / 64) * ((L.n + 3) / 4) * L.o;
}

// Fixed-order reduction: thread t sums partial[t], partial[t+1024], ... in
// order, then a fixed binary tree.
__global__ __launch_bounds__(1024) void k_reduce(const double *__restrict__ partial, int n, double *__restrict__ out) {
	__shared__ double red[1024];
	double s = 0.0;
	for (int q = threadIdx.x; q < n; q += 1024) s = s + partial[q];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 512; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) out[0] = red[0];
}

void launch_reduce(hipStream_t s, const double *partial, int n, double *out) {
	hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, partial, n, out);
	HIPCHECK_LAUNCH();
}

// ------------------------------------------------------------ transfers

// Restriction, gather form.  The reference scatters: it scans the fine grid
// in 2x2x2 cubes in (k,j,i) order and adds each fine residual, scaled by
// 0.125*wx*wy*wz, into up to 8 coarse points (region.cc:665-746, 818-912).
// Per axis of fine size M, cube c holds fine 2c (local 0) and 2c+1 (local 1);
// the last cube is c = (M-1)>>1, and for even M its local 1 is injected into
// coarse c+1 alone.  Coarse point I therefore receives, per axis, from
// (cube I-1, local 1), (cube I, local 0) and (cube I, local 1), and this
// kernel visits them in the reference's order: cube (k,j,i) first, then the
// local position (k,j,i), so every coarse sum is formed in the same order
// and rounds identically.  The reference's restrict_partial tests the x
// injection flag for the y split of local (0,1,1) in boundary cubes
// (region.cc:888-895); that is reproduced.  Also clears coarse x
// (multigrid.cc:219, region.cc:1679-1682).
// Periodic axes: no injection flag (erow is false on a periodic axis,
// region.cc:666-690); on an even periodic axis the last cube's local 1
// (fine M-1) also belongs to coarse 0, which the reference reaches through
// its ghost strip.  In cube-scan order that cube comes after cube 0, so for
// coarse 0 the axis visits cube 0 first and the wrapped last cube second.
// A periodic axis of size 2 has one cube, which is also the last: its local 1
// reaches coarse 0 both as that point's own share and through the wrap, and
// the scatter adds the share twice in a row as it visits the cube (twice per
// such axis, so four times with two of them), not in a second pass.
template <typename T>
__device__ __forceinline__ void restrict_point(const LevelView &F, const LevelView &C, const T *__restrict__ res,
					       T *__restrict__ rc, T *__restrict__ xc, int I, int J, int K) {
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	const bool px = F.prd & 1, py = (F.prd >> 1) & 1, pz = (F.prd >> 2) & 1;
	const bool evx = !(F.m & 1) && !px, evy = !(F.n & 1) && !py, evz = !(F.o & 1) && !pz;
	const bool dx = px && F.m == 2, dy = py && F.n == 2, dz = pz && F.o == 2;
	const bool wx = px && !(F.m & 1) && I == 0 && !dx, wy = py && !(F.n & 1) && J == 0 && !dy,
		   wz = pz && !(F.o & 1) && K == 0 && !dz;
	T acc = (T)0;
	for (int oz = 0; oz < 2; oz++) {
		const int cz = wz ? 1 - oz : oz;
		const int kc = (wz && cz == 0) ? lastz : K - 1 + cz;
		if (kc < 0 || kc > lastz) continue;
		const bool sz = evz && kc == lastz;
		for (int oy = 0; oy < 2; oy++) {
			const int cy = wy ? 1 - oy : oy;
			const int jc = (wy && cy == 0) ? lasty : J - 1 + cy;
			if (jc < 0 || jc > lasty) continue;
			const bool sy = evy && jc == lasty;
			for (int ox = 0; ox < 2; ox++) {
				const int cx = wx ? 1 - ox : ox;
				const int ic = (wx && cx == 0) ? lastx : I - 1 + cx;
				if (ic < 0 || ic > lastx) continue;
				const bool sx = evx && ic == lastx;
				const bool partial = ic == lastx || jc == lasty || kc == lastz;
				for (int lz = cz ? 0 : 1; lz < 2; lz++) {
					const int fk = 2 * kc + lz;
					if (fk >= F.o) continue;
					T wz;
					if (lz == 0) wz = (T)1;
					else if (cz == 0) wz = sz ? (T)1 : (T)0.5;
					else { if (sz) continue; wz = (T)0.5; }
					for (int ly = cy ? 0 : 1; ly < 2; ly++) {
						const int fj = 2 * jc + ly;
						if (fj >= F.n) continue;
						for (int lx = cx ? 0 : 1; lx < 2; lx++) {
							const int fi = 2 * ic + lx;
							if (fi >= F.m) continue;
							const bool syq = (partial && lx == 0 && ly == 1 && lz == 1) ? sx : sy;
							T wy, wx;
							if (ly == 0) wy = (T)1;
							else if (cy == 0) wy = syq ? (T)1 : (T)0.5;
							else { if (syq) continue; wy = (T)0.5; }
							if (lx == 0) wx = (T)1;
							else if (cx == 0) wx = sx ? (T)1 : (T)0.5;
							else { if (sx) continue; wx = (T)0.5; }
							const T w = (T)0.125 * wx * wy * wz;
							const T t = w * res[idx3(F, fi, fj, fk)];
							acc = acc + t;
							const int times = (dx && lx ? 2 : 1) * (dy && ly ? 2 : 1) * (dz && lz ? 2 : 1);
							for (int q = 1; q < times; q++) acc = acc + t;
						}
					}
				}
			}
		}
	}
	const long long pc = idx3(C, I, J, K);
	rc[pc] = acc;
	xc[pc] = (T)0;
}

// Restriction at a coarse point whose eight surrounding cubes are all
// interior (1 <= I < lastx, likewise J, K): the same 27 terms, weights and
// order as restrict_point, with the fine values of each of the 9 fine rows
// read as one pair (2I, 2I+1) and one scalar (2I-1).
template <typename T>
__device__ __forceinline__ void restrict_interior(const LevelView &F, const LevelView &C, const T *__restrict__ res,
						  T *__restrict__ rc, T *__restrict__ xc, int I, int J, int K) {
	typedef typename Vec2<T>::type V2;
	T v[3][3][3];  // [dk+1][dj+1][di+1], fine point (2I+di, 2J+dj, 2K+dk)
#pragma unroll
	for (int dk = -1; dk <= 1; dk++) {
#pragma unroll
		for (int dj = -1; dj <= 1; dj++) {
			const long long p = idx3(F, 2 * I, 2 * J + dj, 2 * K + dk);
			const V2 pr = *(const V2 *)(res + p);
			v[dk + 1][dj + 1][0] = res[p - 1];
			v[dk + 1][dj + 1][1] = pr.x;
			v[dk + 1][dj + 1][2] = pr.y;
		}
	}
	T acc = (T)0;
#pragma unroll
	for (int cz = 0; cz < 2; cz++)
#pragma unroll
		for (int cy = 0; cy < 2; cy++)
#pragma unroll
			for (int cx = 0; cx < 2; cx++)
#pragma unroll
				for (int lz = cz ? 0 : 1; lz < 2; lz++)
#pragma unroll
					for (int ly = cy ? 0 : 1; ly < 2; ly++)
#pragma unroll
						for (int lx = cx ? 0 : 1; lx < 2; lx++) {
							const T wz = lz == 0 ? (T)1 : (T)0.5;
							const T wy = ly == 0 ? (T)1 : (T)0.5;
							const T wx = lx == 0 ? (T)1 : (T)0.5;
							const T w = (T)0.125 * wx * wy * wz;
							acc = acc + w * v[2 * cz + lz - 1][2 * cy + ly - 1][2 * cx + lx - 1];
						}
	const long long pc = idx3(C, I, J, K);
	rc[pc] = acc;
	xc[pc] = (T)0;
}

template <typename T>
__device__ __forceinline__ void restrict_any(const LevelView &F, const LevelView &C, const T *__restrict__ res,
					     T *__restrict__ rc, T *__restrict__ xc, int I, int J, int K) {
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	if (I >= 1 && I < lastx && J >= 1 && J < lasty && K >= 1 && K < lastz)
		restrict_interior<T>(F, C, res, rc, xc, I, J, K);
	else
		restrict_point<T>(F, C, res, rc, xc, I, J, K);
}

// The lone hierarchy restricts onto a coarse level of at least
// kRestrictSplitPoints points in two launches: the coarse points whose eight
// surrounding cubes are interior (restrict_interior, every lane on the same
// path), and the shell of the others (restrict_point).  One launch over all
// points put a boundary point into a third of the wavefronts (I = 0 and the
// last I), which then ran both paths: 257^3 -> 129^3 in FP32 87 us, split
// 37 us; onto 65^3 30 -> 22 us; onto 33^3 the second launch cost more (18 ->
// 21 us).
const long long kRestrictSplitPoints = 100000;
//
// Coarse indices on an axis of coarse size n whose last full cube is `last`:
// interior [1, last), shell {0} and [max(last, 1), n).
struct ShellAxis {
	int n, nin, nsh;  // points, interior points, shell points
	int hi;           // first index of the upper shell
	__host__ __device__ ShellAxis(int n_, int last) : n(n_) {
		nin = last > 1 ? min(last, n_) - 1 : 0;
		nsh = n_ - nin;
		hi = max(last, 1);
	}
	__host__ __device__ int shell(int p) const { return p < 1 ? 0 : hi + (p - 1); }
};

// Shell point q (0 <= q < shell_count) of the coarse level: the planes of
// the shell in K whole, then in every interior plane the shell rows in J
// whole and the shell points of the interior rows.
__host__ __device__ inline long long shell_count(const ShellAxis &ax, const ShellAxis &ay, const ShellAxis &az) {
	return (long long)az.nsh * ax.n * ay.n + (long long)az.nin * ((long long)ay.nsh * ax.n + (long long)ay.nin * ax.nsh);
}
__device__ __forceinline__ void shell_point(const ShellAxis &ax, const ShellAxis &ay, const ShellAxis &az, long long q,
					   int &I, int &J, int &K) {
	const long long plane = (long long)ax.n * ay.n;
	const long long faces = (long long)az.nsh * plane;
	const long long ring = (long long)ay.nsh * ax.n + (long long)ay.nin * ax.nsh;
	if (q < faces) {
		K = az.shell((int)(q / plane));
		const long long rem = q % plane;
		J = (int)(rem / ax.n);
		I = (int)(rem % ax.n);
	} else {
		const long long q2 = q - faces;
		K = 1 + (int)(q2 / ring);
		const long long r = q2 % ring;
		if (r < (long long)ay.nsh * ax.n) {
			J = ay.shell((int)(r / ax.n));
			I = (int)(r % ax.n);
		} else {
			const long long r2 = r - (long long)ay.nsh * ax.n;
			J = 1 + (int)(r2 / ax.nsh);
			I = ax.shell((int)(r2 % ax.nsh));
		}
	}
}

// Work units of a batched restriction onto C from F: 64 x 4 tiles of the
// interior points first, then blocks of 256 shell points, so that no
// workgroup mixes the two gathers.
struct RestrictUnits {
	ShellAxis ax, ay, az;
	int gx, gy;
	long long tiles, shell;
	__host__ __device__ RestrictUnits(const LevelView &F, const LevelView &C)
	    : ax(C.m, (F.m - 1) >> 1), ay(C.n, (F.n - 1) >> 1), az(C.o, (F.o - 1) >> 1) {
		gx = (ax.nin + 63) / 64;
		gy = (ay.nin + 3) / 4;
		tiles = (ax.nin > 0 && ay.nin > 0 && az.nin > 0) ? (long long)gx * gy * az.nin : 0;
		shell = shell_count(ax, ay, az);
	}
	__host__ __device__ long long units() const { return tiles + (shell + 255) / 256; }
};

int batch_restrict_units(const LevelView &F, const LevelView &C) { return (int)RestrictUnits(F, C).units(); }

template <typename T>
__global__ __launch_bounds__(256) void k_restrict_interior(LevelView F, LevelView C, const T *__restrict__ res,
							   T *__restrict__ rc, T *__restrict__ xc) {
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1;
	const int I = 1 + (int)(blockIdx.x * 64 + threadIdx.x), J = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
	const int K = 1 + (int)blockIdx.z;
	if (I >= lastx || J >= lasty) return;
	restrict_interior<T>(F, C, res, rc, xc, I, J, K);
}

template <typename T>
__global__ __launch_bounds__(256) void k_restrict_shell(LevelView F, LevelView C, const T *__restrict__ res,
							T *__restrict__ rc, T *__restrict__ xc, long long total) {
	const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
	if (q >= total) return;
	const ShellAxis ax(C.m, (F.m - 1) >> 1), ay(C.n, (F.n - 1) >> 1), az(C.o, (F.o - 1) >> 1);
	int I, J, K;
	shell_point(ax, ay, az, q, I, J, K);
	restrict_point<T>(F, C, res, rc, xc, I, J, K);
}

// One launch over all points, for small coarse levels, where a second launch
// costs more than the divergence it saves.
template <typename T>
__global__ __launch_bounds__(256) void k_restrict(LevelView F, LevelView C, const T *__restrict__ res,
						  T *__restrict__ rc, T *__restrict__ xc) {
	const int I = (int)(blockIdx.x * 64 + threadIdx.x), J = (int)(blockIdx.y * 4 + threadIdx.y), K = (int)blockIdx.z;
	if (I >= C.m || J >= C.n) return;
	restrict_any<T>(F, C, res, rc, xc, I, J, K);
}

template <typename T>
void launch_restrict(hipStream_t s, const LevelView &F, const LevelView &C, const T *res, T *rc, T *xc) {
	if ((long long)C.m * C.n * C.o < kRestrictSplitPoints) {
		dim3 grid((C.m + 63) / 64, (C.n + 3) / 4, C.o), block(64, 4, 1);
		hipLaunchKernelGGL(k_restrict<T>, grid, block, 0, s, F, C, res, rc, xc);
		HIPCHECK_LAUNCH();
		return;
	}
	const ShellAxis ax(C.m, (F.m - 1) >> 1), ay(C.n, (F.n - 1) >> 1), az(C.o, (F.o - 1) >> 1);
	if (ax.nin > 0 && ay.nin > 0 && az.nin > 0) {
		dim3 grid((ax.nin + 63) / 64, (ay.nin + 3) / 4, az.nin), block(64, 4, 1);
		hipLaunchKernelGGL(k_restrict_interior<T>, grid, block, 0, s, F, C, res, rc, xc);
		HIPCHECK_LAUNCH();
	}
	const long long total = shell_count(ax, ay, az);
	if (total > 0) {
		hipLaunchKernelGGL(k_restrict_shell<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, F, C, res, rc, xc,
				   total);
		HIPCHECK_LAUNCH();
	}
}

// Interpolation x_f += P x_c.  Fine point (2ic+li, 2jc+lj, 2kc+lk) takes its
// value from the eight coarse points around cube (ic,jc,kc); interior cubes
// use interpolate_full's association order (region.cc:764-771), cubes at the
// last position of any axis interpolate_partial's (region.cc:795-809).
template <typename T>
__device__ __forceinline__ T interp_value(int pos, bool full, bool xs, bool ys, bool zs, T v0, T v1, T v2, T v3, T v4,
					  T v5, T v6, T v7) {
	const T h = (T)0.5, q = (T)0.25, e = (T)0.125;
	if (full) {
		const T t1 = v0 + v1;
		switch (pos) {
		case 0: return v0;
		case 1: return h * t1;
		case 2: return h * (v0 + v2);
		case 3: return q * ((t1 + v2) + v3);
		case 4: return h * (v0 + v4);
		case 5: return q * (t1 + (v4 + v5));
		case 6: return q * (((v0 + v2) + v4) + v6);
		default: return e * (((((t1 + v2) + v3) + (v4 + v5)) + v6) + v7);
		}
	}
	switch (pos) {
	case 0: return v0;
	case 1: return xs ? v1 : h * (v0 + v1);
	case 2: return ys ? v2 : h * (v0 + v2);
	case 4: return zs ? v4 : h * (v0 + v4);
	case 5: return zs ? (xs ? v5 : h * (v4 + v5)) : (xs ? h * (v1 + v5) : q * (((v0 + v1) + v4) + v5));
	case 6: return zs ? (ys ? v6 : h * (v4 + v6)) : (ys ? h * (v2 + v6) : q * (((v0 + v2) + v4) + v6));
	default: {
		const T t2 = ys ? (xs ? v3 : h * (v2 + v3)) : (xs ? h * (v1 + v3) : q * (((v0 + v1) + v2) + v3));
		if (pos == 3) return t2;
		const T t3 = ys ? (xs ? v7 : h * (v6 + v7)) : (xs ? h * (v5 + v7) : q * (((v4 + v5) + v6) + v7));
		return zs ? t3 : h * (t3 + t2);
	}
	}
}

// One fine cube per thread: its eight coarse values are read once and the
// (up to) eight fine points updated as four x pairs (2ic, 2ic+1).
template <typename T>
__device__ __forceinline__ void interp_cube(const LevelView &C, const LevelView &F, const T *__restrict__ xc,
					    T *__restrict__ xf, int ic, int jc, int kc) {
	typedef typename Vec2<T>::type V2;
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	// on a periodic axis the upper neighbour of the last cube is coarse 0
	// (the reference's ghost strip, region.cc:1369-1388), and no injection
	// flag is set
	const bool px = F.prd & 1, py = (F.prd >> 1) & 1, pz = (F.prd >> 2) & 1;
	const int i1 = ic + 1 < C.m ? ic + 1 : (px ? 0 : ic);
	const int j1 = jc + 1 < C.n ? jc + 1 : (py ? 0 : jc);
	const int k1 = kc + 1 < C.o ? kc + 1 : (pz ? 0 : kc);
	const T v0 = xc[idx3(C, ic, jc, kc)], v1 = xc[idx3(C, i1, jc, kc)];
	const T v2 = xc[idx3(C, ic, j1, kc)], v3 = xc[idx3(C, i1, j1, kc)];
	const T v4 = xc[idx3(C, ic, jc, k1)], v5 = xc[idx3(C, i1, jc, k1)];
	const T v6 = xc[idx3(C, ic, j1, k1)], v7 = xc[idx3(C, i1, j1, k1)];
	const bool full = ic < lastx && jc < lasty && kc < lastz;
	const bool xs = !px && (ic == lastx) && !(F.m & 1), ys = !py && (jc == lasty) && !(F.n & 1);
	const bool zs = !pz && (kc == lastz) && !(F.o & 1);
	const bool two = 2 * ic + 1 < F.m;
#pragma unroll
	for (int lk = 0; lk < 2; lk++) {
		if (2 * kc + lk >= F.o) break;
#pragma unroll
		for (int lj = 0; lj < 2; lj++) {
			if (2 * jc + lj >= F.n) break;
			const int pos = 2 * lj + 4 * lk;
			const long long p = idx3(F, 2 * ic, 2 * jc + lj, 2 * kc + lk);
			const T a = interp_value<T>(pos, full, xs, ys, zs, v0, v1, v2, v3, v4, v5, v6, v7);
			if (two) {
				const T b = interp_value<T>(pos + 1, full, xs, ys, zs, v0, v1, v2, v3, v4, v5, v6, v7);
				V2 f = *(const V2 *)(xf + p);
				f.x = f.x + a;
				f.y = f.y + b;
				*(V2 *)(xf + p) = f;
			} else {
				xf[p] = xf[p] + a;
			}
		}
	}
}

template <typename T>
__global__ __launch_bounds__(256) void k_interp(LevelView C, LevelView F, const T *__restrict__ xc, T *__restrict__ xf) {
	const int ic = (int)(blockIdx.x * 64 + threadIdx.x), jc = (int)(blockIdx.y * 4 + threadIdx.y), kc = (int)blockIdx.z;
	if (ic >= (F.m + 1) / 2 || jc >= (F.n + 1) / 2) return;
	interp_cube<T>(C, F, xc, xf, ic, jc, kc);
}

template <typename T>
void launch_interp(hipStream_t s, const LevelView &C, const LevelView &F, const T *xc, T *xf) {
	const int ncx = (F.m + 1) / 2, ncy = (F.n + 1) / 2, ncz = (F.o + 1) / 2;
	dim3 grid((ncx + 63) / 64, (ncy + 3) / 4, ncz), block(64, 4, 1);
	hipLaunchKernelGGL(k_interp<T>, grid, block, 0, s, C, F, xc, xf);
	HIPCHECK_LAUNCH();
}

// ------------------------------------------------------------------ RAT

// Support of interpolation column J on an axis of fine size M: fine range
// [lo,hi) and the fine index c that coarse J sits on (igrid_map,
// region.cc:1243-1255).
__device__ __forceinline__ void psupport(int J, int M, int &lo, int &hi, int &c) {
	const int j2 = 2 * J;
	if (M & 1) { lo = j2 == 0 ? 0 : j2 - 1; hi = j2 == M - 1 ? M : j2 + 2; c = j2; }
	else if (j2 == M) { lo = M - 1; hi = M; c = M - 1; }
	else { lo = j2 == 0 ? 0 : j2 - 1; hi = j2 == M - 2 ? M - 1 : j2 + 2; c = j2; }
}

// Galerkin coefficient Ac(I, I+d) of one coarse point I and one entry d.
// The reference (region.cc:1008-1192) takes each fine point f in cube-scan
// order, forms At[J] = sum_g 0.125*Af(f,g)*P(g,J) over g in f's box in
// (k,j,i) order (A_red), scales At by f's restriction weight (A_mul/A_sca)
// and adds it to the coarse rows of f (A_add).  Here the sum for one (I,J)
// visits the same f in the same order, and forms each inner sum the same way.
__device__ __forceinline__ double rat_entry(const LevelView &F, const LevelView &C, const double *__restrict__ Af,
					    int I, int Jy, int K, int eJ) {
	const int dI = eJ % 3 - 1, dJ = (eJ / 3) % 3 - 1, dK = eJ / 9 - 1;
	const int JI = I + dI, JJ = Jy + dJ, JK = K + dK;
	double acc = 0.0;
	if (JI < 0 || JI >= C.m || JJ < 0 || JJ >= C.n || JK < 0 || JK >= C.o) return acc;
	int plx, phx, pcx, ply, phy, pcy, plz, phz, pcz;
	psupport(JI, F.m, plx, phx, pcx);
	psupport(JJ, F.n, ply, phy, pcy);
	psupport(JK, F.o, plz, phz, pcz);
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	const bool evx = !(F.m & 1), evy = !(F.n & 1), evz = !(F.o & 1);
	// The loops are unrolled with the fine point's validity as a predicate
	// instead of a branch, and its row block is loaded from a clamped (always
	// valid) address, so the loads of several fine points can be in flight
	// at once; the terms are still added in cube-scan order.
#pragma unroll
	for (int cz = 0; cz < 2; cz++) {
		const int kc = K - 1 + cz;
		const bool vkc = kc >= 0 && kc <= lastz;
		const bool sz = evz && kc == lastz;
#pragma unroll
		for (int cy = 0; cy < 2; cy++) {
			const int jc = Jy - 1 + cy;
			const bool vjc = vkc && jc >= 0 && jc <= lasty;
			const bool sy = evy && jc == lasty;
#pragma unroll
			for (int cx = 0; cx < 2; cx++) {
				const int ic = I - 1 + cx;
				const bool vic = vjc && ic >= 0 && ic <= lastx;
				const bool sx = evx && ic == lastx;
#pragma unroll
				for (int lz = cz ? 0 : 1; lz < 2; lz++) {
					const int fk = 2 * kc + lz;
					// weight of f in the cube's restriction; a second-cube local
					// 1 past an even-size last point does not exist
					const double wz = lz == 0 ? 1.0 : (cz == 0 ? (sz ? 1.0 : 0.5) : 0.5);
					const int gzl = max(plz, fk - 1), gzh = min(phz, min(fk + 2, F.o));
					const bool vz = vic && fk < F.o && !(cz == 1 && lz == 1 && sz) && gzl < gzh;
#pragma unroll
					for (int ly = cy ? 0 : 1; ly < 2; ly++) {
						const int fj = 2 * jc + ly;
						const double wy = ly == 0 ? 1.0 : (cy == 0 ? (sy ? 1.0 : 0.5) : 0.5);
						const int gyl = max(ply, fj - 1), gyh = min(phy, min(fj + 2, F.n));
						const bool vy = vz && fj < F.n && !(cy == 1 && ly == 1 && sy) && gyl < gyh;
#pragma unroll
						for (int lx = cx ? 0 : 1; lx < 2; lx++) {
							const int fi = 2 * ic + lx;
							const double wx = lx == 0 ? 1.0 : (cx == 0 ? (sx ? 1.0 : 0.5) : 0.5);
							const int gxl = max(plx, fi - 1), gxh = min(phx, min(fi + 2, F.m));
							const bool valid = vy && fi < F.m && !(cx == 1 && lx == 1 && sx) && gxl < gxh;
							// all 27 planes of a fine point's row block are valid
							// memory; an invalid f reads a clamped in-grid one
							const int ci = min(max(fi, 0), F.m - 1), cj2 = min(max(fj, 0), F.n - 1),
								  ck2 = min(max(fk, 0), F.o - 1);
							const double *Afr = Af + aidx3(F, 0, ci, cj2, ck2);
							double a[27];
#pragma unroll
							for (int ef = 0; ef < 27; ef++) a[ef] = Afr[ef * F.pa];
							if (valid) {
								double inner = 0.0;
#pragma unroll
								for (int dk = 0; dk < 3; dk++) {
									const int kk = fk - 1 + dk;
									const bool okk = kk >= gzl && kk < gzh;
									const double kfac = kk == pcz ? 1.0 : 0.5;
#pragma unroll
									for (int dj = 0; dj < 3; dj++) {
										const int jj = fj - 1 + dj;
										const bool okj = okk && jj >= gyl && jj < gyh;
										const double jkfac = jj == pcy ? kfac : 0.5 * kfac;
#pragma unroll
										for (int di = 0; di < 3; di++) {
											const int ii = fi - 1 + di;
											if (okj && ii >= gxl && ii < gxh)
												inner = inner + (0.125 * a[dk * 9 + dj * 3 + di]) *
														    (ii == pcx ? jkfac : 0.5 * jkfac);
										}
									}
								}
								const double wr = wx * wy * wz;
								acc = acc + inner * wr;
							}
						}
					}
				}
			}
		}
	}
	return acc;
}

// Interior coarse points (1 <= I < lastx, likewise J, K: every cube of I's
// restriction full, every fine box and interpolation support unclipped): the
// terms rat_entry visits are fixed by the offsets alone, so they are
// expanded at compile time.  For fine point f = 2I + df (df per axis -1, 0,
// 1 in the cube-scan order) and entry J = I + d, the term of box entry ef
// (offset eg) enters with P weight pw(df+eg, 2d): 1 at J's fine point, 1/2
// next to it, else absent; the products and sums are those of rat_entry in
// the same order.
namespace rat {
__host__ __device__ constexpr double pw(int dg, int c) {
	return dg == c ? 1.0 : ((dg - c == 1 || c - dg == 1) ? 0.5 : 0.0);
}
template <int DFX, int DFY, int DFZ, int J, int EF>
__device__ __forceinline__ void term(const double (&a8)[27], double &inner) {
	constexpr int dx = J % 3 - 1, dy = (J / 3) % 3 - 1, dz = J / 9 - 1;
	constexpr int ex = EF % 3 - 1, ey = (EF / 3) % 3 - 1, ez = EF / 9 - 1;
	constexpr double wx = pw(DFX + ex, 2 * dx), wy = pw(DFY + ey, 2 * dy), wz = pw(DFZ + ez, 2 * dz);
	if constexpr (wx != 0.0 && wy != 0.0 && wz != 0.0) {
		constexpr double kfac = wz;
		constexpr double jkfac = wy == 1.0 ? kfac : 0.5 * kfac;
		constexpr double w = wx == 1.0 ? jkfac : 0.5 * jkfac;
		inner = inner + a8[EF] * w;
	}
}
template <int DFX, int DFY, int DFZ, int J>
__host__ __device__ constexpr bool touches() {
	constexpr int dx = J % 3 - 1, dy = (J / 3) % 3 - 1, dz = J / 9 - 1;
	bool any = false;
	for (int EF = 0; EF < 27; EF++)
		if (pw(DFX + EF % 3 - 1, 2 * dx) != 0.0 && pw(DFY + (EF / 3) % 3 - 1, 2 * dy) != 0.0 &&
		    pw(DFZ + EF / 9 - 1, 2 * dz) != 0.0)
			any = true;
	return any;
}
template <int DFX, int DFY, int DFZ, int J, int... EF>
__device__ __forceinline__ void entry(const double (&a8)[27], double wr, double (&acc)[27],
				      std::integer_sequence<int, EF...>) {
	if constexpr (touches<DFX, DFY, DFZ, J>()) {
		double inner = 0.0;
		(term<DFX, DFY, DFZ, J, EF>(a8, inner), ...);
		acc[J] = acc[J] + inner * wr;
	}
}
template <int DFX, int DFY, int DFZ, int... J>
__device__ __forceinline__ void fine_point(const LevelView &F, const double *__restrict__ Af, int fi, int fj, int fk,
					   double (&acc)[27], std::integer_sequence<int, J...>) {
	// restriction weight of f: 1 at the cube's local 0, 1/2 at local 1
	constexpr double wx = DFX == 0 ? 1.0 : 0.5, wy = DFY == 0 ? 1.0 : 0.5, wz = DFZ == 0 ? 1.0 : 0.5;
	const double wr = wx * wy * wz;
	const double *Afr = Af + aidx3(F, 0, fi + DFX, fj + DFY, fk + DFZ);
	double a8[27];
#pragma unroll
	for (int e = 0; e < 27; e++) a8[e] = 0.125 * Afr[e * F.pa];
	(entry<DFX, DFY, DFZ, J>(a8, wr, acc, std::make_integer_sequence<int, 27>()), ...);
}
// Cube q (cz = q>>2, cy, cx) and its local point l (lz = l>>2, ly, lx), in
// the scan order cube first, then local; local 0 of a cube at offset 0 on
// an axis exists only in the cube at offset 1 (restrict_point's loops).
// JS: the entries J formed (all 27, or one).
template <int Q, int L, typename JS>
__device__ __forceinline__ void cube_local(const LevelView &F, const double *__restrict__ Af, int fi, int fj, int fk,
					   double (&acc)[27]) {
	constexpr int cz = Q >> 2, cy = (Q >> 1) & 1, cx = Q & 1;
	constexpr int lz = L >> 2, ly = (L >> 1) & 1, lx = L & 1;
	if constexpr ((cz || lz) && (cy || ly) && (cx || lx))
		fine_point<2 * (cx - 1) + lx, 2 * (cy - 1) + ly, 2 * (cz - 1) + lz>(F, Af, fi, fj, fk, acc, JS{});
}
template <int Q, typename JS, int... L>
__device__ __forceinline__ void cube(const LevelView &F, const double *__restrict__ Af, int fi, int fj, int fk,
				     double (&acc)[27], std::integer_sequence<int, L...>) {
	(cube_local<Q, L, JS>(F, Af, fi, fj, fk, acc), ...);
}
template <typename JS, int... Q>
__device__ __forceinline__ void all(const LevelView &F, const double *__restrict__ Af, int fi, int fj, int fk,
				    double (&acc)[27], std::integer_sequence<int, Q...>) {
	(cube<Q, JS>(F, Af, fi, fj, fk, acc, std::make_integer_sequence<int, 8>()), ...);
}
}  // namespace rat

__global__ __launch_bounds__(256) void k_rat_interior(LevelView F, LevelView C, const double *__restrict__ Af,
						      double *__restrict__ Ac) {
	const int I = 1 + (int)(blockIdx.x * 64 + threadIdx.x), Jy = 1 + (int)(blockIdx.y * 4 + threadIdx.y), K = 1 + (int)blockIdx.z;
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1;
	if (I >= lastx || Jy >= lasty) return;
	double acc[27];
#pragma unroll
	for (int e = 0; e < 27; e++) acc[e] = 0.0;
	rat::all<std::make_integer_sequence<int, 27>>(F, Af, 2 * I, 2 * Jy, 2 * K, acc, std::make_integer_sequence<int, 8>());
	double *Acr = Ac + aidx3(C, 0, I, Jy, K);
#pragma unroll
	for (int e = 0; e < 27; e++) Acr[e * C.pa] = acc[e];
}

// The same interior points one entry per thread (blockIdx.y = entry J, so a
// wavefront shares it): for levels with few interior points, where one
// thread per point leaves most of the GPU idle and each thread long.  The
// entry's sum takes the same fine points in the same order, so its bits are
// k_rat_interior's.
template <int J>
__device__ __forceinline__ void rat_interior_entry(const LevelView &F, const LevelView &C, const double *__restrict__ Af,
						   double *__restrict__ Ac, int I, int Jy, int K) {
	double acc[27];
#pragma unroll
	for (int e = 0; e < 27; e++) acc[e] = 0.0;
	rat::all<std::integer_sequence<int, J>>(F, Af, 2 * I, 2 * Jy, 2 * K, acc, std::make_integer_sequence<int, 8>());
	Ac[aidx3(C, J, I, Jy, K)] = acc[J];
}
template <int... J>
__device__ __forceinline__ void rat_interior_dispatch(int e, const LevelView &F, const LevelView &C,
						      const double *__restrict__ Af, double *__restrict__ Ac, int I, int Jy,
						      int K, std::integer_sequence<int, J...>) {
	((e == J ? rat_interior_entry<J>(F, C, Af, Ac, I, Jy, K) : void()), ...);
}
__global__ __launch_bounds__(256) void k_rat_interior_entry(LevelView F, LevelView C, const double *__restrict__ Af,
							    double *__restrict__ Ac) {
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	const int ix = lastx - 1, iy = lasty - 1, iz = lastz - 1;
	const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= (long long)ix * iy * iz) return;
	const int I = 1 + (int)(q % ix), Jy = 1 + (int)((q / ix) % iy), K = 1 + (int)(q / ((long long)ix * iy));
	rat_interior_dispatch((int)blockIdx.y, F, C, Af, Ac, I, Jy, K, std::make_integer_sequence<int, 27>());
}

// The coarse points k_rat_interior leaves out, enumerated densely: those
// with K in the z boundary set {0} u [lastz, C.o), then (K interior) those
// with J in the y set, then (K, J interior) those with I in the x set.
struct RatEdges {
	int nbx, nby, nbz;     // sizes of the boundary index sets
	long long na, nb, nc;  // points in the three groups
};
__host__ __device__ inline int edge_index(int q, int last) { return q == 0 ? 0 : last + q - 1; }
__host__ __device__ inline RatEdges rat_edges(const LevelView &F, const LevelView &C) {
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	RatEdges r;
	r.nbx = lastx > 1 ? 1 + C.m - lastx : C.m;
	r.nby = lasty > 1 ? 1 + C.n - lasty : C.n;
	r.nbz = lastz > 1 ? 1 + C.o - lastz : C.o;
	const int ix = C.m - r.nbx, iy = C.n - r.nby, iz = C.o - r.nbz;  // interior counts
	r.na = (long long)r.nbz * C.m * C.n;
	r.nb = (long long)iz * r.nby * C.m;
	r.nc = (long long)iz * iy * r.nbx;
	(void)ix;
	return r;
}

// One thread per (boundary point, entry): small levels have few interior
// points, so the boundary work is spread as widely as possible.
__global__ __launch_bounds__(256) void k_rat_boundary(LevelView F, LevelView C, const double *__restrict__ Af,
						      double *__restrict__ Ac, RatEdges E) {
	const long long tt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int eJ = (int)(tt % 27);
	long long t = tt / 27;
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	const bool allx = E.nbx == C.m, ally = E.nby == C.n, allz = E.nbz == C.o;
	int I, Jy, K;
	if (t < E.na) {
		I = (int)(t % C.m); Jy = (int)((t / C.m) % C.n);
		const int q = (int)(t / ((long long)C.m * C.n));
		K = allz ? q : edge_index(q, lastz);
	} else if ((t -= E.na) < E.nb) {
		I = (int)(t % C.m);
		const int q = (int)((t / C.m) % E.nby);
		Jy = ally ? q : edge_index(q, lasty);
		K = 1 + (int)(t / ((long long)C.m * E.nby));
	} else if ((t -= E.nb) < E.nc) {
		const int q = (int)(t % E.nbx);
		I = allx ? q : edge_index(q, lastx);
		Jy = 1 + (int)((t / E.nbx) % (C.n - E.nby));
		K = 1 + (int)(t / ((long long)E.nbx * (C.n - E.nby)));
	} else {
		return;
	}
	Ac[aidx3(C, eJ, I, Jy, K)] = rat_entry(F, C, Af, I, Jy, K, eJ);
}

// Interior coarse points below which a level's Galerkin entries are formed
// one per thread.
constexpr long long kRatEntrySplit = 1LL << 17;

// Galerkin entries with periodic axes, one (coarse point, entry) per thread,
// on the unwrapped index line of each periodic axis (the reference's
// periodic grid_map/igrid_map, region.cc:1201-1242).  One axis of P: coarse
// J sits on fine pfpos(J); a fine point on a coarse position takes that
// value alone, any other half of each coarse neighbour.  Entry d of coarse
// row I sums, over fine f with P(f,I) != 0 in (k,j,i) order,
// P(f,I) * [sum over f's box g in (k,j,i) order of 0.125*Af(f,g)*P(g,I+d)],
// the bracket formed as A_red forms At (region.cc:1117-1166).
__device__ __forceinline__ int pfpos(int J, int M, int Mc, bool prd) {
	if (prd) {
		const int q = J >= 0 ? J / Mc : -((-J + Mc - 1) / Mc);
		return 2 * (J - q * Mc) + M * q;
	}
	return (!(M & 1) && J == Mc - 1) ? M - 1 : 2 * J;
}
__device__ __forceinline__ double pweight(int g, int J, int M, int Mc, bool prd) {
	const int f = pfpos(J, M, Mc, prd);
	const bool c = prd ? ((((g % M) + M) % M) & 1) == 0 : ((g & 1) == 0 || (!(M & 1) && g == M - 1));
	if (c) return g == f ? 1.0 : 0.0;
	return (g - f == 1 || f - g == 1) ? 0.5 : 0.0;
}

// The one-axis factors of an entry take few values: for each axis a thread
// tabulates P(f,I) over its three fine positions f and P(g,J) over their
// three box neighbours g, so the 27 x 27 terms need no integer division.
// The products are formed in the same order (axis x, y, z, from 1) and are
// exact (powers of two or zero), so the sums are those of the untabulated
// form term by term.
__global__ __launch_bounds__(256) void k_rat_prd(LevelView F, LevelView C, const double *__restrict__ Af,
						 double *__restrict__ Ac) {
	const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
	const long long npts = (long long)C.m * C.n * C.o;
	if (t >= npts * 27) return;
	// entry-major: the lanes of a wavefront take one entry d at neighbouring
	// coarse points, so their zero factors (and skipped loads) agree
	const int d = (int)(t / npts);
	const long long pc = t - (long long)d * npts;
	const int I[3] = {(int)(pc % C.m), (int)((pc / C.m) % C.n), (int)(pc / ((long long)C.m * C.n))};
	const bool P[3] = {(F.prd & 1) != 0, (F.prd & 2) != 0, (F.prd & 4) != 0};
	const int Mf[3] = {F.m, F.n, F.o}, Mc[3] = {C.m, C.n, C.o};
	const int J[3] = {I[0] + d % 3 - 1, I[1] + (d / 3) % 3 - 1, I[2] + d / 9 - 1};
	bool ok = true;
	for (int a = 0; a < 3; a++)
		if (!P[a] && (J[a] < 0 || J[a] >= Mc[a])) ok = false;
	double acc = 0.0;
	if (ok) {
		double wf[3][3], pg[3][3][3];
		int fw[3][3];
#pragma unroll
		for (int a = 0; a < 3; a++) {
			const int fc = pfpos(I[a], Mf[a], Mc[a], P[a]);
#pragma unroll
			for (int df = 0; df < 3; df++) {
				const int f = fc + df - 1;
				const bool vf = P[a] || (f >= 0 && f < Mf[a]);
				wf[a][df] = vf ? pweight(f, I[a], Mf[a], Mc[a], P[a]) : 0.0;
				fw[a][df] = ((f % Mf[a]) + Mf[a]) % Mf[a];
#pragma unroll
				for (int de = 0; de < 3; de++) {
					const int g = f + de - 1;
					const bool vg = P[a] || (g >= 0 && g < Mf[a]);
					pg[a][df][de] = (vf && vg) ? pweight(g, J[a], Mf[a], Mc[a], P[a]) : 0.0;
				}
			}
		}
#pragma unroll
		for (int fz = 0; fz < 3; fz++)
#pragma unroll
		for (int fy = 0; fy < 3; fy++)
#pragma unroll
		for (int fx = 0; fx < 3; fx++) {
			const double w = ((1.0 * wf[0][fx]) * wf[1][fy]) * wf[2][fz];
			if (w == 0.0) continue;
			const double *Arow = Af + aidx3(F, 0, fw[0][fx], fw[1][fy], fw[2][fz]);
			double sum = 0.0;
#pragma unroll
			for (int e = 0; e < 27; e++) {
				const double pw = ((1.0 * pg[0][fx][e % 3]) * pg[1][fy][(e / 3) % 3]) * pg[2][fz][e / 9];
				if (pw != 0.0) sum = sum + 0.125 * Arow[(long long)e * F.pa] * pw;
			}
			acc = acc + sum * w;
		}
	}
	Ac[aidx3(C, d, I[0], I[1], I[2])] = acc;
}

// Copies the 27 coefficients of the last point of every grid row into the
// row block's tail (levels with tail; see arow0).  Run by every launcher that
// writes coefficients, after its writes.
__global__ void k_coef_tail(LevelView L, double *__restrict__ A) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long long)L.n * L.o * 27) return;
	const int e = (int)(t % 27);
	const long long q = t / 27;
	const int j = (int)(q % L.n), k = (int)(q / L.n);
	A[arow0(L, j, k) + 27LL * L.pa + e] = A[aidx3(L, e, L.m - 1, j, k)];
}
static void launch_coef_tail(hipStream_t s, const LevelView &L, double *A) {
	if (!L.tail) return;
	const long long tot = (long long)L.n * L.o * 27;
	hipLaunchKernelGGL(k_coef_tail, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, L, A);
	HIPCHECK_LAUNCH();
}

void launch_rat(hipStream_t s, const LevelView &F, const LevelView &C, const double *Af, double *Ac) {
	if (F.prd) {
		const long long tot = (long long)C.m * C.n * C.o * 27;
		hipLaunchKernelGGL(k_rat_prd, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, F, C, Af, Ac);
		HIPCHECK_LAUNCH();
		launch_coef_tail(s, C, Ac);
		return;
	}
	const int lastx = (F.m - 1) >> 1, lasty = (F.n - 1) >> 1, lastz = (F.o - 1) >> 1;
	if (lastx > 1 && lasty > 1 && lastz > 1) {
		const long long ni = (long long)(lastx - 1) * (lasty - 1) * (lastz - 1);
		if (ni < kRatEntrySplit) {
			hipLaunchKernelGGL(k_rat_interior_entry, dim3((unsigned)((ni + 255) / 256), 27), dim3(256), 0, s, F, C, Af, Ac);
		} else {
			dim3 gi((lastx - 1 + 63) / 64, (lasty - 1 + 3) / 4, lastz - 1), block(64, 4, 1);
			hipLaunchKernelGGL(k_rat_interior, gi, block, 0, s, F, C, Af, Ac);
		}
		HIPCHECK_LAUNCH();
	}
	const RatEdges E = rat_edges(F, C);
	const long long nbnd = (E.na + E.nb + E.nc) * 27;
	hipLaunchKernelGGL(k_rat_boundary, dim3((unsigned)((nbnd + 255) / 256)), dim3(256), 0, s, F, C, Af, Ac, E);
	HIPCHECK_LAUNCH();
	launch_coef_tail(s, C, Ac);
}

// ---------------------------------------------------------- bottom LU

// Dense column-major matrix of the bottom level, row = point, column =
// neighbour (region_tr.cc:486-504).  The caller zero-fills LU first.
__global__ void k_lu_assemble(LevelView L, const double *__restrict__ A, double *__restrict__ LU) {
	const int N = L.m * L.n * L.o;
	const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (tid >= (long long)N * 27) return;
	const int e = (int)(tid % 27), p = (int)(tid / 27);
	const int i = p % L.m, j = (p / L.m) % L.n, k = p / (L.m * L.n);
	const int qi = i + e % 3 - 1, qj = j + (e / 3) % 3 - 1, qk = k + e / 9 - 1;
	if (qi < 0 || qi >= L.m || qj < 0 || qj >= L.n || qk < 0 || qk >= L.o) return;
	const int q = qi + L.m * (qj + L.n * qk);
	LU[p + (long long)q * N] = A[aidx3(L, e, i, j, k)];
}

// With periodic axes, one thread per row: a wrapped neighbour's column is its
// wrapped index (step_mod, region.hh:466; region_tr.cc:493-501), and entries
// are stored in box order, so where two offsets reach one column (an axis
// of size 1 or 2) the later one is kept, as in the reference.
__global__ void k_lu_assemble_prd(LevelView L, const double *__restrict__ A, double *__restrict__ LU) {
	const int N = L.m * L.n * L.o;
	const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if (p >= N) return;
	const int i = p % L.m, j = (p / L.m) % L.n, k = p / (L.m * L.n);
	for (int e = 0; e < 27; e++) {
		int qi = i + e % 3 - 1, qj = j + (e / 3) % 3 - 1, qk = k + e / 9 - 1;
		if (qi < 0 || qi >= L.m) { if (!(L.prd & 1)) continue; qi = (qi + L.m) % L.m; }
		if (qj < 0 || qj >= L.n) { if (!(L.prd & 2)) continue; qj = (qj + L.n) % L.n; }
		if (qk < 0 || qk >= L.o) { if (!(L.prd & 4)) continue; qk = (qk + L.o) % L.o; }
		const int q = qi + L.m * (qj + L.n * qk);
		LU[p + (long long)q * N] = A[aidx3(L, e, i, j, k)];
	}
}

void launch_lu_assemble(hipStream_t s, const LevelView &L, const double *A, double *LU) {
	if (L.prd) {
		const int N = L.m * L.n * L.o;
		hipLaunchKernelGGL(k_lu_assemble_prd, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, L, A, LU);
		HIPCHECK_LAUNCH();
		return;
	}
	const long long total = (long long)L.m * L.n * L.o * 27;
	hipLaunchKernelGGL(k_lu_assemble, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, L, A, LU);
	HIPCHECK_LAUNCH();
}

// Right-looking LU with partial pivoting by one workgroup: pivot = first
// row of largest modulus (dgetrf's rule, region_tr.cc:518), column scaled
// by the pivot's reciprocal, rank-1 update.  info = first zero pivot (1-based)
// and info[1] = 1 when the factor fails the test of region_tr.cc:526-533.
__global__ __launch_bounds__(1024) void k_lu_factor(int N, double *__restrict__ LU, int *__restrict__ ipiv,
						    int *__restrict__ info) {
	__shared__ double bv[1024];
	__shared__ int bi[1024];
	__shared__ int s_piv, s_info, s_nz;
	const int tid = threadIdx.x, nt = blockDim.x;
	if (tid == 0) s_info = 0;
	for (int c = 0; c < N; c++) {
		double best = -1.0;
		int bidx = N;
		for (int r = c + tid; r < N; r += nt) {
			const double v = fabs(LU[r + (long long)c * N]);
			if (v > best) { best = v; bidx = r; }
		}
		bv[tid] = best; bi[tid] = bidx;
		__syncthreads();
		for (int w = nt / 2; w > 0; w >>= 1) {
			if (tid < w) {
				const double ov = bv[tid + w];
				const int oi = bi[tid + w];
				if (ov > bv[tid] || (ov == bv[tid] && oi < bi[tid])) { bv[tid] = ov; bi[tid] = oi; }
			}
			__syncthreads();
		}
		if (tid == 0) {
			// The sequential rule keeps the diagonal unless another entry
			// is strictly larger; a NaN column keeps the diagonal.
			int piv = bi[0];
			if (!(bv[0] > fabs(LU[c + (long long)c * N])) || piv >= N) piv = c;
			s_piv = piv;
			s_nz = LU[piv + (long long)c * N] != 0.0;
			ipiv[c] = piv;
		}
		__syncthreads();
		const int piv = s_piv;
		// Read once by thread 0 before any swap, so every thread takes the
		// same branch and reaches the same barriers.
		const bool nz = s_nz != 0;
		if (nz) {
			if (piv != c) {
				for (int jcol = tid; jcol < N; jcol += nt) {
					const double t = LU[c + (long long)jcol * N];
					LU[c + (long long)jcol * N] = LU[piv + (long long)jcol * N];
					LU[piv + (long long)jcol * N] = t;
				}
			}
			__syncthreads();
			const double rp = 1.0 / LU[c + (long long)c * N];
			for (int r = c + 1 + tid; r < N; r += nt) LU[r + (long long)c * N] *= rp;
		} else if (tid == 0 && s_info == 0) {
			s_info = c + 1;
		}
		__syncthreads();
		const int rem = N - c - 1;
		const long long tot = (long long)rem * rem;
		for (long long q = tid; q < tot; q += nt) {
			const int r = c + 1 + (int)(q % rem), jcol = c + 1 + (int)(q / rem);
			const double u = LU[c + (long long)jcol * N];
			if (u != 0.0) LU[r + (long long)jcol * N] -= LU[r + (long long)c * N] * u;
		}
		__syncthreads();
	}
	if (tid == 0) {
		info[0] = s_info;
		const double hi = fabs(LU[0]), lo = fabs(LU[(long long)N * N - 1]);
		info[1] = (s_info != 0 || hi == 0.0 || lo / hi < 2.220446049250313e-16 * 1e4) ? 1 : 0;
	}
}

void launch_lu_factor(hipStream_t s, int N, double *LU, int *ipiv, int *info) {
	hipLaunchKernelGGL(k_lu_factor, dim3(1), dim3(1024), 0, s, N, LU, ipiv, info);
	HIPCHECK_LAUNCH();
}

// dgetrs('N') on one workgroup (region_tr.cc:544-555): interchanges, then
// column-oriented forward and back substitution in FP64, one barrier per
// column.  r and x are in the level layout and type T.
template <typename T>
__device__ __forceinline__ void lu_solve_body(const LevelView &L, int N, const double *__restrict__ LU,
					      const int *__restrict__ ipiv, const T *__restrict__ r, T *__restrict__ x) {
	__shared__ double xs[1024];
	const int tid = threadIdx.x, nt = blockDim.x;
	for (int q = tid; q < N; q += nt) {
		const int i = q % L.m, j = (q / L.m) % L.n, k = q / (L.m * L.n);
		xs[q] = (double)r[idx3(L, i, j, k)];
	}
	__syncthreads();
	if (tid == 0) {
		for (int c = 0; c < N; c++) {
			const int p = ipiv[c];
			if (p != c) { const double t = xs[c]; xs[c] = xs[p]; xs[p] = t; }
		}
	}
	__syncthreads();
	for (int c = 0; c < N; c++) {
		const double v = xs[c];
		if (v != 0.0)
			for (int rr = c + 1 + tid; rr < N; rr += nt) xs[rr] -= v * LU[rr + (long long)c * N];
		__syncthreads();
	}
	for (int c = N - 1; c >= 0; c--) {
		const double v = xs[c] / LU[c + (long long)c * N];
		__syncthreads();
		if (tid == 0) xs[c] = v;
		if (v != 0.0)
			for (int rr = tid; rr < c; rr += nt) xs[rr] -= v * LU[rr + (long long)c * N];
		__syncthreads();
	}
	for (int q = tid; q < N; q += nt) {
		const int i = q % L.m, j = (q / L.m) % L.n, k = q / (L.m * L.n);
		x[idx3(L, i, j, k)] = (T)xs[q];
	}
}

template <typename T>
__global__ __launch_bounds__(1024) void k_lu_solve(LevelView L, int N, const double *__restrict__ LU,
						   const int *__restrict__ ipiv, const T *__restrict__ r, T *__restrict__ x) {
	lu_solve_body<T>(L, N, LU, ipiv, r, x);
}

template <typename T>
void launch_lu_solve(hipStream_t s, const LevelView &L, int N, const double *LU, const int *ipiv, const T *r, T *x) {
	hipLaunchKernelGGL(k_lu_solve<T>, dim3(1), dim3(1024), 0, s, L, N, LU, ipiv, r, x);
	HIPCHECK_LAUNCH();
}

// ------------------------------------------------- generation, layout

// One point per thread (the 7-point families, and Q1 on a grid with periodic
// axes); the family is a template constant so the row stays in registers.
template <int FAM>
__global__ __launch_bounds__(256) void k_generate(LevelView L, uint64_t seed, double *__restrict__ A,
						  double *__restrict__ r) {
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y), k = (int)blockIdx.z;
	if (i >= L.m || j >= L.n) return;
	double row[27];
	pgmg_row27_prd(FAM, seed, L.m, L.n, L.o, L.prd, i, j, k, row);
	const long long p = idx3(L, i, j, k);
	double *Ap = A + aidx3(L, 0, i, j, k);
#pragma unroll
	for (int e = 0; e < 27; e++) Ap[e * L.pa] = row[e];
	r[p] = pgmg_rhs(seed, L.m, L.n, L.o, i, j, k);
}

// Q1 family: a point's row sums element matrices of its 8 cells, and each
// cell tensor (two 10^u series, a rotation) is shared by 8 points.  A block
// owns a 64 x 4 column of points over kz planes and keeps the two cell planes
// under the current point plane in LDS (struct of arrays, 6 tensor entries),
// so each tensor is formed about once instead of 8 times.  The row is summed
// exactly as pgmg_row27 sums it: cells in (k,j,i) order, local node lb
// ascending, each entry from pgmg_q1_entry with the same operation order, so
// the rows are bit-identical.
constexpr int kGenCx = 65, kGenCy = 5, kGenPlane = kGenCx * kGenCy;

__device__ __forceinline__ void gen_cell_plane(double *Kp, uint64_t seed, int m, int n, int o, int i0, int j0, int ck) {
	const int t0 = threadIdx.y * 64 + threadIdx.x;
	if (ck > o) return;
	for (int t = t0; t < kGenPlane; t += 256) {
		const int ci = i0 + t % kGenCx, cj = j0 + t / kGenCx;
		if (ci > m || cj > n) continue;
		double K[6];
		pgmg_cell_tensor(seed, m, n, ci, cj, ck, K);
#pragma unroll
		for (int e = 0; e < 6; e++) Kp[e * kGenPlane + t] = K[e];
	}
}

template <int LA>
__device__ __forceinline__ void gen_q1_cell(const double *Kp, int c, int i, int j, int k, int m, int n, int o,
					    double (&row)[27]) {
	PGMG_NO_CONTRACT
	constexpr int dc0 = 1 - (LA & 1), dc1 = 1 - ((LA >> 1) & 1), dc2 = 1 - ((LA >> 2) & 1);
	double K[6];
#pragma unroll
	for (int e = 0; e < 6; e++) K[e] = Kp[e * kGenPlane + c];
#pragma unroll
	for (int lb = 0; lb < 8; lb++) {
		const int bi = i + dc0 - 1 + (lb & 1), bj = j + dc1 - 1 + ((lb >> 1) & 1), bk = k + dc2 - 1 + ((lb >> 2) & 1);
		if (bi < 0 || bi >= m || bj < 0 || bj >= n || bk < 0 || bk >= o) continue;
		const int idx = (dc2 + ((lb >> 2) & 1)) * 9 + (dc1 + ((lb >> 1) & 1)) * 3 + (dc0 + (lb & 1));
		row[idx] = row[idx] + pgmg_q1_entry(K, LA, lb);
	}
}

__global__ __launch_bounds__(256) void k_generate_q1(LevelView L, uint64_t seed, int kz, double *__restrict__ A,
						     double *__restrict__ r) {
	__shared__ double Ks[2][6 * kGenPlane];
	const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 4;
	const int kb = blockIdx.z * kz, ke = min(L.o, kb + kz);
	const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
	const bool in = i < L.m && j < L.n;
	gen_cell_plane(Ks[kb & 1], seed, L.m, L.n, L.o, i0, j0, kb);
	for (int k = kb; k < ke; k++) {
		gen_cell_plane(Ks[(k + 1) & 1], seed, L.m, L.n, L.o, i0, j0, k + 1);
		__syncthreads();
		if (in) {
			double row[27];
#pragma unroll
			for (int e = 0; e < 27; e++) row[e] = 0.0;
			const double *lo = Ks[k & 1], *hi = Ks[(k + 1) & 1];
			const int c = threadIdx.y * kGenCx + threadIdx.x;
			// cells (ci, cj, ck) = (i + dc0, j + dc1, k + dc2) in (k,j,i) order
			gen_q1_cell<7>(lo, c, i, j, k, L.m, L.n, L.o, row);
			gen_q1_cell<6>(lo, c + 1, i, j, k, L.m, L.n, L.o, row);
			gen_q1_cell<5>(lo, c + kGenCx, i, j, k, L.m, L.n, L.o, row);
			gen_q1_cell<4>(lo, c + kGenCx + 1, i, j, k, L.m, L.n, L.o, row);
			gen_q1_cell<3>(hi, c, i, j, k, L.m, L.n, L.o, row);
			gen_q1_cell<2>(hi, c + 1, i, j, k, L.m, L.n, L.o, row);
			gen_q1_cell<1>(hi, c + kGenCx, i, j, k, L.m, L.n, L.o, row);
			gen_q1_cell<0>(hi, c + kGenCx + 1, i, j, k, L.m, L.n, L.o, row);
			double *Ap = A + aidx3(L, 0, i, j, k);
#pragma unroll
			for (int e = 0; e < 27; e++) Ap[e * L.pa] = row[e];
			r[idx3(L, i, j, k)] = pgmg_rhs(seed, L.m, L.n, L.o, i, j, k);
		}
		__syncthreads();
	}
}

void launch_generate(hipStream_t s, const LevelView &L, int family, uint64_t seed, double *A, double *r) {
	dim3 block(64, 4, 1);
	if (family == PGMG_FAM_Q1_ANISO && L.prd) {
		// periodic cells: the per-point form, which wraps the cell indices
		dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o);
		hipLaunchKernelGGL(k_generate<PGMG_FAM_Q1_ANISO>, grid, block, 0, s, L, seed, A, r);
	} else if (family == PGMG_FAM_Q1_ANISO) {
		// planes per block: as many as keep >= 2048 blocks, at most 8
		const long long cols = (long long)((L.m + 63) / 64) * ((L.n + 3) / 4);
		int kz = 8;
		while (kz > 1 && cols * ((L.o + kz - 1) / kz) < 2048) kz >>= 1;
		dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, (L.o + kz - 1) / kz);
		hipLaunchKernelGGL(k_generate_q1, grid, block, 0, s, L, seed, kz, A, r);
	} else {
		dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o);
		if (family == PGMG_FAM_FACE7_LOGUNI)
			hipLaunchKernelGGL(k_generate<PGMG_FAM_FACE7_LOGUNI>, grid, block, 0, s, L, seed, A, r);
		else
			hipLaunchKernelGGL(k_generate<PGMG_FAM_FACE7_UNIFORM>, grid, block, 0, s, L, seed, A, r);
	}
	HIPCHECK_LAUNCH();
	launch_coef_tail(s, L, A);
}

__device__ __forceinline__ long long compact_to_padded(const LevelView &L, long long p) {
	const int i = (int)(p % L.m), j = (int)((p / L.m) % L.n), k = (int)(p / ((long long)L.m * L.n));
	return idx3(L, i, j, k);
}

__device__ __forceinline__ long long compact_to_coef(const LevelView &L, int e, long long p) {
	const int i = (int)(p % L.m), j = (int)((p / L.m) % L.n), k = (int)(p / ((long long)L.m * L.n));
	return aidx3(L, e, i, j, k);
}

__global__ void k_pack_vec(LevelView L, const double *__restrict__ c, double *__restrict__ d, long long p0, long long n) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) d[compact_to_padded(L, p0 + t)] = c[t];
}
__global__ void k_unpack_vec(LevelView L, const double *__restrict__ d, double *__restrict__ c, long long p0, long long n) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) c[t] = d[compact_to_padded(L, p0 + t)];
}
__global__ void k_pack_op(LevelView L, const double *__restrict__ c, double *__restrict__ A, long long p0, long long n) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n * 27) return;
	const int e = (int)(t % 27);
	const long long p = t / 27;
	A[compact_to_coef(L, e, p0 + p)] = c[t];
}
__global__ void k_unpack_op(LevelView L, const double *__restrict__ A, double *__restrict__ c, long long p0, long long n) {
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n * 27) return;
	const int e = (int)(t % 27);
	const long long p = t / 27;
	c[t] = A[compact_to_coef(L, e, p0 + p)];
}

static unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_pack_vec(hipStream_t s, const LevelView &L, const double *c, double *d, long long p0, long long n) {
	if (n > 0) hipLaunchKernelGGL(k_pack_vec, dim3(nblk(n, 256)), dim3(256), 0, s, L, c, d, p0, n);
	HIPCHECK_LAUNCH();
}
void launch_unpack_vec(hipStream_t s, const LevelView &L, const double *d, double *c, long long p0, long long n) {
	if (n > 0) hipLaunchKernelGGL(k_unpack_vec, dim3(nblk(n, 256)), dim3(256), 0, s, L, d, c, p0, n);
	HIPCHECK_LAUNCH();
}
void launch_pack_op(hipStream_t s, const LevelView &L, const double *c, double *A, long long p0, long long n) {
	if (n > 0) hipLaunchKernelGGL(k_pack_op, dim3(nblk(n * 27, 256)), dim3(256), 0, s, L, c, A, p0, n);
	HIPCHECK_LAUNCH();
	launch_coef_tail(s, L, A);
}
void launch_unpack_op(hipStream_t s, const LevelView &L, const double *A, double *c, long long p0, long long n) {
	if (n > 0) hipLaunchKernelGGL(k_unpack_op, dim3(nblk(n * 27, 256)), dim3(256), 0, s, L, A, c, p0, n);
	HIPCHECK_LAUNCH();
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, long long n, T v) {
	for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) p[t] = v;
}
template <typename T>
void launch_fill(hipStream_t s, T *p, long long n, T v) {
	if (n <= 0) return;
	long long b = (n + 255) / 256;
	if (b > 65536) b = 65536;
	hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)b), dim3(256), 0, s, p, n, v);
	HIPCHECK_LAUNCH();
}

__global__ void k_round_f32(const double *__restrict__ a, float *__restrict__ b, long long n) {
	for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) b[t] = (float)a[t];
}
void launch_round_f32(hipStream_t s, const double *a, float *b, long long n) {
	long long bl = (n + 255) / 256;
	if (bl > 65536) bl = 65536;
	hipLaunchKernelGGL(k_round_f32, dim3((unsigned)bl), dim3(256), 0, s, a, b, n);
	HIPCHECK_LAUNCH();
}

__global__ __launch_bounds__(256) void k_add_f32(LevelView L, const float *__restrict__ e, double *__restrict__ x) {
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y), k = (int)blockIdx.z;
	if (i >= L.m || j >= L.n) return;
	const long long p = idx3(L, i, j, k);
	x[p] = x[p] + (double)e[p];
}
void launch_add_f32_to_f64(hipStream_t s, const LevelView &L, const float *e, double *x) {
	dim3 grid((L.m + 63) / 64, (L.n + 3) / 4, L.o), block(64, 4, 1);
	hipLaunchKernelGGL(k_add_f32, grid, block, 0, s, L, e, x);
	HIPCHECK_LAUNCH();
}

// ------------------------------------------------------ batched kernels
//
// A batch is a set of independent hierarchies (slots) of any sizes.  One
// launch applies one operation to one level of every slot in a work list;
// blocks walk the list's units grid-stride and find their slot by binary
// search over the cumulative unit counts.  The per-row and per-point bodies
// are those of the single-hierarchy kernels, so a slot's arithmetic does not
// depend on which other slots share its launches.

int batch_units(int kind, const LevelView &L, int elem) {
	switch (kind) {
	case BK_PAIR0: case BK_PAIR1: case BK_PAIR2: case BK_PAIR3: {
		const int waves = elem == 4 ? gs_pair_waves<float>(L, kind - BK_PAIR0) : gs_pair_waves<double>(L, kind - BK_PAIR0);
		return (waves + 3) / 4;
	}
	case BK_ONE: return 1;
	case BK_ROWS: return residual_blocks(L, elem);
	case BK_CUBES: return (((L.m + 1) / 2 + 63) / 64) * (((L.n + 1) / 2 + 3) / 4) * ((L.o + 1) / 2);
	case BK_FLAT: return (int)(((long long)L.m * L.n * L.o + 255) / 256);
	default: return ((L.m + 63) / 64) * ((L.n + 3) / 4) * L.o;
	}
}

__device__ __forceinline__ int wl_find(const WorkList &wl, int u) {
	int lo = 0, hi = wl.n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (wl.prefix[mid] <= u) lo = mid; else hi = mid - 1;
	}
	return lo;
}

template <typename T> __device__ __forceinline__ T *sA(const SlotDev *S, int l);
template <typename T> __device__ __forceinline__ T *sx(const SlotDev *S, int l);
template <typename T> __device__ __forceinline__ T *sr(const SlotDev *S, int l);
template <typename T> __device__ __forceinline__ T *sres(const SlotDev *S);
template <> __device__ __forceinline__ double *sA<double>(const SlotDev *S, int l) { return S->A[l]; }
template <> __device__ __forceinline__ double *sx<double>(const SlotDev *S, int l) { return S->x[l]; }
template <> __device__ __forceinline__ double *sr<double>(const SlotDev *S, int l) { return S->r[l]; }
template <> __device__ __forceinline__ double *sres<double>(const SlotDev *S) { return S->res64; }
template <> __device__ __forceinline__ float *sA<float>(const SlotDev *S, int l) { return S->A32[l]; }
template <> __device__ __forceinline__ float *sx<float>(const SlotDev *S, int l) { return S->x32[l]; }
template <> __device__ __forceinline__ float *sr<float>(const SlotDev *S, int l) { return S->r32[l]; }
template <> __device__ __forceinline__ float *sres<float>(const SlotDev *S) { return S->res32; }

// Decodes a 64x4-point block unit of level view L into the thread's point.
__device__ __forceinline__ bool pts_unit(const LevelView &L, int lu, int &i, int &j, int &k) {
	const int gx = (L.m + 63) / 64, gy = (L.n + 3) / 4;
	i = (lu % gx) * 64 + (int)threadIdx.x;
	j = ((lu / gx) % gy) * 4 + (int)threadIdx.y;
	k = lu / (gx * gy);
	return i < L.m && j < L.n;
}

// Unit lu of kind BK_FLAT: 256 consecutive grid points in (i, j, k) order,
// every lane busy whatever the row length.
__device__ __forceinline__ bool flat_unit(const LevelView &L, int lu, int &i, int &j, int &k) {
	const long long p = (long long)lu * 256 + (int)threadIdx.y * 64 + (int)threadIdx.x;
	const long long mn = (long long)L.m * L.n;
	if (p >= mn * L.o) return false;
	k = (int)(p / mn);
	const int q = (int)(p - (long long)k * mn);
	j = q / L.m;
	i = q - j * L.m;
	return true;
}

// Unit lu of kind BK_CUBES: the fine cubes (2ic.., 2jc.., 2kc..) of level L.
__device__ __forceinline__ bool cubes_unit(const LevelView &L, int lu, int &ic, int &jc, int &kc) {
	const int ncx = (L.m + 1) / 2, ncy = (L.n + 1) / 2;
	const int gx = (ncx + 63) / 64, gy = (ncy + 3) / 4;
	ic = (lu % gx) * 64 + (int)threadIdx.x;
	jc = ((lu / gx) % gy) * 4 + (int)threadIdx.y;
	kc = lu / (gx * gy);
	return ic < ncx && jc < ncy;
}

// Each workgroup takes one contiguous chunk of the list's units (chunks
// numbered per XCD, so neighbouring units share an L2), finds the first
// unit's slot by bisection and then walks forward through the slots.
#define BATCH_LOOP                                                                         \
	const int total_ = wl.prefix[wl.n];                                                \
	const int per_ = (total_ + (int)gridDim.x - 1) / (int)gridDim.x;                   \
	const int u0_ = xcd_block((int)blockIdx.x, (int)gridDim.x) * per_;                 \
	const int u1_ = min(total_, u0_ + per_);                                           \
	int q_ = u0_ < u1_ ? __builtin_amdgcn_readfirstlane(wl_find(wl, u0_)) : 0;         \
	for (int u = u0_; u < u1_; u++) {                                                  \
		while (u >= wl.prefix[q_ + 1]) q_++;                                       \
		const SlotDev *S = slots + wl.slot[q_];                                    \
		const int lu = u - wl.prefix[q_];

template <typename T, bool Z = false>
__global__ __launch_bounds__(256) void kb_gs_pair(const SlotDev *__restrict__ slots, WorkList wl,
								   int level, int cj, int ck) {
	BATCH_LOOP
		const LevelView L = S->lv[level];
		gs_wave<T, Z, false>(L, sA<T>(S, level), sx<T>(S, level), sr<T>(S, level), cj, ck,
			      __builtin_amdgcn_readfirstlane(lu * 4 + (int)threadIdx.y));
	}
}

template <typename T>
__global__ __launch_bounds__(1024) void kb_gs_small(const SlotDev *__restrict__ slots, WorkList wl, int level, int nsweeps) {
	extern __shared__ char smem_[];
	BATCH_LOOP
		(void)lu;
		const LevelView L = S->lv[level];
		if (small_lds_bytes(L, (int)sizeof(T)) <= kSmallLdsBytes)
			gs_lds_body<T>(L, sA<T>(S, level), sx<T>(S, level), sr<T>(S, level), nsweeps, 0, 0, (T *)smem_);
		else
			gs_small_body<T>(L, sA<T>(S, level), sx<T>(S, level), sr<T>(S, level), nsweeps, 0, 0);
		__syncthreads();
	}
}

template <typename T>
__global__ __launch_bounds__(256) void kb_residual(const SlotDev *__restrict__ slots, WorkList wl, int level) {
	BATCH_LOOP
		const LevelView L = S->lv[level];
		(void)residual_wave<T, T>(L, sA<T>(S, level), sx<T>(S, level), sr<T>(S, level), sres<T>(S), (T *)nullptr,
					  __builtin_amdgcn_readfirstlane(lu * 4 + (int)threadIdx.y));
	}
}

// Top-level FP64 residual with per-slot block sums; optionally the residual
// rounded to FP32 into the slot's level-0 r32 (the defect-correction source).
__global__ __launch_bounds__(256) void kb_residual_top(const SlotDev *__restrict__ slots, WorkList wl, int to_f32) {
	BATCH_LOOP
		const LevelView L = S->lv[0];
		const int w = __builtin_amdgcn_readfirstlane(lu * 4 + (int)threadIdx.y);
		const double v2 = to_f32 ? residual_wave<double, float>(L, S->A[0], S->x[0], S->r[0], S->r32[0],
									 S->top_zero ? (float *)nullptr : S->x32[0], w)
					 : residual_wave<double, double>(L, S->A[0], S->x[0], S->r[0], (double *)nullptr, (double *)nullptr, w);
		const double sblk = block_sum_256(v2);
		if (threadIdx.x == 0 && threadIdx.y == 0) S->partial[lu] = sblk;
		__syncthreads();
	}
}

__global__ __launch_bounds__(1024) void kb_reduce(const SlotDev *__restrict__ slots, WorkList wl, double *__restrict__ norms) {
	__shared__ double red[1024];
	BATCH_LOOP
		(void)lu;
		const LevelView L = S->lv[0];
		const int n = residual_blocks(L, 8);
		double s = 0.0;
		for (int t = threadIdx.x; t < n; t += 1024) s = s + S->partial[t];
		red[threadIdx.x] = s;
		__syncthreads();
		for (int w = 512; w > 0; w >>= 1) {
			if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
			__syncthreads();
		}
		if (threadIdx.x == 0) norms[wl.slot[q_]] = red[0];
		__syncthreads();
	}
}

template <typename T>
__global__ __launch_bounds__(256) void kb_restrict(const SlotDev *__restrict__ slots, WorkList wl, int level) {
	BATCH_LOOP
		const LevelView F = S->lv[level - 1], C = S->lv[level];
		const RestrictUnits R(F, C);
		if (lu < R.tiles) {
			const int I = 1 + (int)(lu % R.gx) * 64 + (int)threadIdx.x;
			const int J = 1 + (int)((lu / R.gx) % R.gy) * 4 + (int)threadIdx.y;
			const int K = 1 + (int)(lu / ((long long)R.gx * R.gy));
			if (I <= R.ax.nin && J <= R.ay.nin)
				restrict_interior<T>(F, C, sres<T>(S), sr<T>(S, level), sx<T>(S, level), I, J, K);
		} else {
			const long long q = (lu - R.tiles) * 256 + (int)threadIdx.y * 64 + (int)threadIdx.x;
			if (q < R.shell) {
				int I, J, K;
				shell_point(R.ax, R.ay, R.az, q, I, J, K);
				restrict_point<T>(F, C, sres<T>(S), sr<T>(S, level), sx<T>(S, level), I, J, K);
			}
		}
	}
}

template <typename T>
__global__ __launch_bounds__(256) void kb_interp(const SlotDev *__restrict__ slots, WorkList wl, int level) {
	BATCH_LOOP
		const LevelView F = S->lv[level - 1];
		int ic, jc, kc;
		if (cubes_unit(F, lu, ic, jc, kc)) interp_cube<T>(S->lv[level], F, sx<T>(S, level), sx<T>(S, level - 1), ic, jc, kc);
	}
}

template <typename T>
__global__ __launch_bounds__(1024) void kb_lu_solve(const SlotDev *__restrict__ slots, WorkList wl, int level) {
	BATCH_LOOP
		(void)lu;
		lu_solve_body<T>(S->lv[level], S->lu_n, S->LU, S->ipiv, sr<T>(S, level), sx<T>(S, level));
		__syncthreads();
	}
}

__global__ __launch_bounds__(256) void kb_add_x(const SlotDev *__restrict__ slots, WorkList wl) {
	BATCH_LOOP
		const LevelView L = S->lv[0];
		int i, j, k;
		if (flat_unit(L, lu, i, j, k)) {
			const long long p = idx3(L, i, j, k);
			S->x[0][p] = S->x[0][p] + (double)S->x32[0][p];
		}
	}
}
#undef BATCH_LOOP

// chunk work units per workgroup (1: the hardware schedules the units of
// all slots as it would a single hierarchy's launch).
static unsigned batch_grid(const WorkList &wl, int chunk = 1) {
	const int c = chunk > 0 ? chunk : 1;
	return (unsigned)(wl.total > 0 ? (wl.total + c - 1) / c : 1);
}

template <typename T>
void launch_b_gs_pair(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int pair, bool from_zero, int chunk) {
	if (wl.n <= 0) return;
	if constexpr (sizeof(T) == 4) {
		if (from_zero) {
			hipLaunchKernelGGL((kb_gs_pair<T, true>), dim3(batch_grid(wl, chunk)), dim3(64, 4), 0, s, slots, wl, level, pair & 1,
					   (pair >> 1) & 1);
			HIPCHECK_LAUNCH();
			return;
		}
	}
	hipLaunchKernelGGL((kb_gs_pair<T, false>), dim3(batch_grid(wl, chunk)), dim3(64, 4), 0, s, slots, wl, level, pair & 1,
			   (pair >> 1) & 1);
	HIPCHECK_LAUNCH();
}
template <typename T>
void launch_b_gs_small(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int nsweeps) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_gs_small<T>, dim3(batch_grid(wl)), dim3(1024), kSmallLdsBytes, s, slots, wl, level, nsweeps);
	HIPCHECK_LAUNCH();
}
template <typename T>
void launch_b_residual(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int chunk) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_residual<T>, dim3(batch_grid(wl, chunk)), dim3(64, 4), 0, s, slots, wl, level);
	HIPCHECK_LAUNCH();
}
void launch_b_residual_top(hipStream_t s, const SlotDev *slots, WorkList wl, bool to_f32, int chunk) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_residual_top, dim3(batch_grid(wl, chunk)), dim3(64, 4), 0, s, slots, wl, to_f32 ? 1 : 0);
	HIPCHECK_LAUNCH();
}
void launch_b_reduce(hipStream_t s, const SlotDev *slots, WorkList wl, double *norms) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_reduce, dim3(batch_grid(wl)), dim3(1024), 0, s, slots, wl, norms);
	HIPCHECK_LAUNCH();
}
template <typename T>
void launch_b_restrict(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int chunk) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_restrict<T>, dim3(batch_grid(wl, chunk)), dim3(64, 4), 0, s, slots, wl, level);
	HIPCHECK_LAUNCH();
}
template <typename T>
void launch_b_interp(hipStream_t s, const SlotDev *slots, WorkList wl, int level, int chunk) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_interp<T>, dim3(batch_grid(wl, chunk)), dim3(64, 4), 0, s, slots, wl, level);
	HIPCHECK_LAUNCH();
}
template <typename T>
void launch_b_lu_solve(hipStream_t s, const SlotDev *slots, WorkList wl, int level) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_lu_solve<T>, dim3(batch_grid(wl)), dim3(1024), 0, s, slots, wl, level);
	HIPCHECK_LAUNCH();
}
void launch_b_add_x(hipStream_t s, const SlotDev *slots, WorkList wl, int chunk) {
	if (wl.n <= 0) return;
	hipLaunchKernelGGL(kb_add_x, dim3(batch_grid(wl, chunk)), dim3(64, 4), 0, s, slots, wl);
	HIPCHECK_LAUNCH();
}

// ------------------------------------------------------ instantiations
#define INST(T) \
	template void launch_gs_colour<T>(hipStream_t, const LevelView &, const T *, T *, const T *, int); \
	template void launch_gs_pair<T>(hipStream_t, const LevelView &, const T *, T *, const T *, int, bool); \
	template void launch_gs_plane<T>(hipStream_t, const LevelView &, const T *, T *, const T *, int); \
	template void launch_gs_planes<T>(hipStream_t, const LevelView &, const T *, T *, const T *, int); \
	template void launch_gs_planes_skew<T>(hipStream_t, const SkewView &, const T *, T *, const T *, int); \
	template bool launch_gs_flow<T>(hipStream_t, const SkewView &, bool, const T *, T *, const T *, int, int *, int *, int); \
	template int flow_max_sweeps<T>(const SkewView &, bool); \
	template void launch_skew_pack_op<T>(hipStream_t, const LevelView &, const SkewView &, const T *, T *); \
	template void launch_skew_vec<T>(hipStream_t, const LevelView &, const SkewView &, const T *, T *, bool); \
	template void launch_skew_halo_wrap<T>(hipStream_t, const LevelView &, const SkewView &, T *); \
	template void launch_gs_small<T>(hipStream_t, const LevelView &, const T *, T *, const T *, int, int, int); \
	template int launch_sumsq<T>(hipStream_t, const LevelView &, const T *, double *); \
	template void launch_restrict<T>(hipStream_t, const LevelView &, const LevelView &, const T *, T *, T *); \
	template void launch_interp<T>(hipStream_t, const LevelView &, const LevelView &, const T *, T *); \
	template void launch_lu_solve<T>(hipStream_t, const LevelView &, int, const double *, const int *, const T *, T *); \
	template void launch_fill<T>(hipStream_t, T *, long long, T); \
	template void launch_b_gs_pair<T>(hipStream_t, const SlotDev *, WorkList, int, int, bool, int); \
	template void launch_b_gs_small<T>(hipStream_t, const SlotDev *, WorkList, int, int); \
	template void launch_b_residual<T>(hipStream_t, const SlotDev *, WorkList, int, int); \
	template void launch_b_restrict<T>(hipStream_t, const SlotDev *, WorkList, int, int); \
	template void launch_b_interp<T>(hipStream_t, const SlotDev *, WorkList, int, int); \
	template void launch_b_lu_solve<T>(hipStream_t, const SlotDev *, WorkList, int);
INST(double)
INST(float)
template void launch_halo_wrap<double>(hipStream_t, const LevelView &, double *);
template void launch_halo_wrap<float>(hipStream_t, const LevelView &, float *);
template int launch_residual<double, double>(hipStream_t, const LevelView &, const double *, const double *, const double *, double *, double *, double *);
template int launch_residual<double, float>(hipStream_t, const LevelView &, const double *, const double *, const double *, float *, double *, float *);
template int launch_residual<float, float>(hipStream_t, const LevelView &, const float *, const float *, const float *, float *, double *, float *);
template void launch_fill<int>(hipStream_t, int *, long long, int);

}  // namespace pgmg
