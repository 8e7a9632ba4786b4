// xsb_device.h -- device-side building blocks of the lookup path (gfx950).
//
// Everything here is compiled with -ffp-contract=off: each FP64 operation is
// the single IEEE operation the reference performs, in the reference's order,
// so the FP64 path reproduces openmp-threading's results bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xsb {

constexpr uint64_t kLcgA = 2806196910506780709ULL;
constexpr uint64_t kLcgMask = 0x7FFFFFFFFFFFFFFFULL;  // mod 2^63
constexpr int kMaterials = 12;
constexpr int kPointDoubles = 6;  // energy + 5 XS channels (NuclideGridPoint)
constexpr int kIvlDoubles = 12;   // interval record: upper point, then upper - lower
constexpr int kHwinDoubles = 12;  // hash window record (Tables::hwin)

enum GridType : int { kUnionized = 0, kNuclide = 1, kHash = 2 };

// Lookup-path choices (all give identical results).
constexpr int kVariantWave = 1;  // wave-uniform path for material-uniform waves (unionized FP64)
constexpr int kVariantSortedRows = 2;  // every nuclide's energy row is non-decreasing
constexpr int kVariantNonNeg = 4;  // no negative XS value or concentration in the tables
constexpr int kVariantDefault = kVariantWave;

// Device view of the tables.  Passed by value as a kernel argument.
struct Tables {
	const double *__restrict__ ueg;         // unionized energies [n_union]
	const int32_t *__restrict__ index_grid; // [n_union or hash_bins][n_iso]
	const double *__restrict__ grid;        // FP64 AoS, 6 doubles per point
	// unionized grid: ueg_acc[b] = #{u : ueg[u] <= b / 2^16}, b = 0..2^16 --
	// the bisection of an energy in [b, b+1) / 2^16 starts from
	// [ueg_acc[b], ueg_acc[b+1]] instead of the whole array
	const int32_t *__restrict__ ueg_acc;
	// unionized grid: FP64 interval records for the wave path, record
	// nuc * n_gp + k = point k+1 (E, 5 XS), then point k+1 - point k
	const double *__restrict__ ivl;
	// unionized grid with FP32 tables: FP32 interval records, 64 B each:
	// {E_{k+1}, E_{k+1} - E_k, float xs_{k+1}[5], float xs_{k+1} - xs_k [5]}
	const double *__restrict__ ivl32;
	const float *__restrict__ xs32;         // FP32 XS, 8 floats per point (5 used)
	const uint16_t *__restrict__ xs16;      // BF16 XS, 8 halves per point (5 used)
	const float *__restrict__ xs16_scale;   // per nuclide x channel scale (5 per nuclide, 8 stride)
	const int32_t *__restrict__ num_nucs;   // [12]
	const int32_t *__restrict__ mats;       // [12][max_num_nucs]
	const double *__restrict__ concs;       // [12][max_num_nucs]
	const float *__restrict__ concs32;      // FP32 copy of concs
	const uint64_t *__restrict__ ff_tab;    // skip-ahead table: 8 digits x 256 x {A, C}
	int64_t n_union;
	int32_t n_iso;
	int32_t idx_pitch;  // row pitch of the device index grid (ints, multiple of 4)
	int32_t n_gp;
	int32_t hash_bins;
	int32_t max_num_nucs;
	int32_t grid_type;
	int32_t variant;  // kVariant* bits: lookup-path choices (tuning, same results)
	double thr[kMaterials];  // pick_mat thresholds, reverse-order partial sums
	double hash_du;          // 1.0 / hash_bins, as computed by the reference
	// unionized grid, per material m: its own copy of the index grid holding
	// only its nuclides' columns, in material-list order, rows padded to 16
	// bytes -- row e starts at element midx_off[m] + e * midx_pitch[m].
	// Elements are 2 << (midx_shift - 1) bytes: uint16 (midx_shift 1) when
	// every bracket fits (n_gp <= 65536), else int32 (midx_shift 2).
	const uint8_t *__restrict__ midx;
	int64_t midx_off[kMaterials];
	int32_t midx_pitch[kMaterials];
	int32_t midx_shift;
	// hash grid, for lanes whose bins differ (the history rounds): each
	// nuclide's energies on their own, row nuc at esoa + nuc * e_pitch
	// (e_pitch >= n_gp + 16, a multiple of 8; +inf past n_gp), and per bin b
	// and nuclide a 96-byte window record at hwin + (b * idx_pitch + nuc) *
	// kHwinDoubles: {u_low, u_high} (int32 x 2, the two index grid entries
	// the reference reads, Simulation.c:282-288), 8 bytes of padding, then
	// E[u_low & ~1 ...] ten energies of the nuclide's row
	const double *__restrict__ hwin;
	const double *__restrict__ esoa;
	int32_t e_pitch;
};

// ---------------------------------------------------------------------------
// Sampling (openmp-threading/Simulation.c:426-499)
// ---------------------------------------------------------------------------

// LCG_random_double: s = (a s + 1) mod 2^63, value s / 2^63.  The u64 -> f64
// conversion is correctly rounded on both sides (hi*2^32 + lo in one rounded
// add here, cvtsi2sd on x86), and the scaling by 2^-63 is exact.
__device__ __forceinline__ double lcg_double(uint64_t &s)
{
	s = (kLcgA * s + 1ULL) & kLcgMask;
	return (double)s * 0x1p-63;
}

// fast_forward_LCG: O(log n) skip-ahead by square-and-multiply.
__device__ __forceinline__ uint64_t ff_lcg(uint64_t seed, uint64_t n)
{
	uint64_t a = kLcgA, c = 1ULL, an = 1ULL, cn = 0ULL;
	n &= kLcgMask;
	while (n) {
		if (n & 1ULL) {
			an *= a;
			cn = cn * a + c;
		}
		c *= (a + 1ULL);
		a *= a;
		n >>= 1;
	}
	return (an * seed + cn) & kLcgMask;
}

// fast_forward_LCG through a radix-256 table: tab[2 (256 k + d)] = {A, C}
// of the skip by d * 256^k.  Skips of one LCG commute, so applying the
// digits of n in any order gives the reference's FF(seed, n) exactly.
__device__ __forceinline__ uint64_t ff_jump(uint64_t seed, uint64_t n,
                                            const uint64_t *__restrict__ tab)
{
	n &= kLcgMask;
	uint64_t s = seed;
	for (int k = 0; n; k++, n >>= 8) {
		const uint32_t d = (uint32_t)(n & 0xFFu);
		if (d) {
			const ulonglong2 ac = *reinterpret_cast<const ulonglong2 *>(tab + 2 * (256 * k + d));
			s = ac.x * s + ac.y;
		}
	}
	return s & kLcgMask;
}

// Starting skip of history particle p (openmp-threading/Simulation.c:167):
// p * lookups * 2 * 5 in the reference's 32-bit int arithmetic, sign-extended
// into the unsigned skip count.
__device__ __forceinline__ uint64_t history_offset(int64_t p, int lookups)
{
	const uint32_t prod = (uint32_t)p * (uint32_t)lookups * 10u;
	return (uint64_t)(int64_t)(int32_t)prod;
}

// pick_mat: the reference returns the first i with roll < thr[i], where
// thr[i] is its reverse-order partial sum (thr[0] = 0), else 0.  The host
// checks that thr is strictly increasing, so that first i equals the number
// of thresholds <= roll.
__device__ __forceinline__ int pick_mat(uint64_t &s, const double *thr)
{
	double roll = lcg_double(s);
	int cnt = 0;
#pragma unroll
	for (int i = 0; i < kMaterials; i++)
		cnt += (thr[i] <= roll) ? 1 : 0;
	return cnt < kMaterials ? cnt : 0;
}

// ---------------------------------------------------------------------------
// Searches (openmp-threading/Simulation.c:380-423).  Both return the unique
// clamp(upper_bound(q) - 1, low, high - 1); any exact bisection gives it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t grid_search(int32_t n, double q, const double *__restrict__ A)
{
	int32_t lo = 0, hi = n - 1;
	while (hi - lo > 1) {
		int32_t mid = lo + ((hi - lo) >> 1);
		if (A[mid] > q)
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

// grid_search with the bracket of upper_bound(q) narrowed to [L, H] (the
// count of A <= q lies in it): same result, fewer probes.  Reads A[L..H-1].
__device__ __forceinline__ int32_t grid_search_in(int32_t n, double q,
                                                  const double *__restrict__ A, int32_t L,
                                                  int32_t H)
{
	int32_t lo = L - 1, hi = H;  // A[lo] <= q (or lo = -1), A[hi] > q (or hi = n)
	while (hi - lo > 1) {
		const int32_t mid = lo + ((hi - lo) >> 1);
		if (A[mid] > q)
			hi = mid;
		else
			lo = mid;
	}
	return min(max(hi - 1, 0), n - 2);
}

constexpr int kUegAccBits = 16;

// bisection on energies stored with a stride (in doubles)
template <int STRIDE>
__device__ __forceinline__ int32_t grid_search_nuclide(double q, const double *__restrict__ E,
                                                       int32_t low, int32_t high)
{
	int32_t lo = low, hi = high;
	while (hi - lo > 1) {
		int32_t mid = lo + ((hi - lo) >> 1);
		if (E[(int64_t)mid * STRIDE] > q)
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

// ---------------------------------------------------------------------------
// Per-nuclide bracket (openmp-threading/Simulation.c:241-302): the lower
// point index `lower` such that points lower and lower+1 are interpolated.
// E_row points at the nuclide's first energy word, stored with stride STRIDE.
// ---------------------------------------------------------------------------
// Unionized grid: row idx of material m's index table (entry j = the
// bracket of the material's j-th nuclide).
__device__ __forceinline__ const uint8_t *ueg_row(const Tables &T, int m, int64_t idx)
{
	return T.midx + ((T.midx_off[m] + idx * T.midx_pitch[m]) << T.midx_shift);
}

// Entry j of an index row of either element width.
__device__ __forceinline__ int32_t ueg_bracket(const Tables &T, const uint8_t *row, int j)
{
	return T.midx_shift == 1 ? (int32_t)reinterpret_cast<const uint16_t *>(row)[j]
	                         : reinterpret_cast<const int32_t *>(row)[j];
}

// Hash grid, one lane's bracket of nuclide nuc in bin idx when the lanes'
// bins differ (openmp-threading/Simulation.c:280-302).  One 96-byte record
// holds u_low, u_high and the energies E[b .. b + 9], b = u_low rounded
// down to even (six aligned 16-byte loads, one dependent step where the
// index entries, then the bisection's probes, would take several); eight
// more energies come from the nuclide's row where the bin spans further
// (with n_gp / bins points per bin on average, a bin of the 45000-point grid
// over 10000 bins exceeds ten in about 4% of cases).  With E[b .. u_high]
// sorted, the count c of those <= p decides the reference's three cases:
// p <= E[u_low] gives 0; c counting every one of them means p >= E[u_high],
// giving n_gp - 1; else the bisection of [u_low, u_high] ends at b + c - 1.
// A bin spanning more than eighteen energies, or unsorted rows, bisects.
__device__ __forceinline__ int32_t hash_lower_window(const Tables &T, int nuc, int64_t idx, double p)
{
	const int32_t n_gp = T.n_gp;
	const double2 *__restrict__ R =
	    reinterpret_cast<const double2 *>(T.hwin + (idx * T.idx_pitch + nuc) * kHwinDoubles);
	const double2 h = R[0];
	const int32_t u_low = __double2loint(h.x), u_high = __double2hiint(h.x);
	double v[10];
#pragma unroll
	for (int q = 0; q < 5; q++) {
		const double2 x = R[1 + q];
		v[2 * q] = x.x;
		v[2 * q + 1] = x.y;
	}
	const int32_t b = u_low & ~1;
	const int32_t top = u_high - b;  // position of u_high in the window
	const double e_low = (u_low & 1) ? v[1] : v[0];
	const double *__restrict__ E = T.esoa + (int64_t)nuc * T.e_pitch;
	const bool sorted = (T.variant & kVariantSortedRows) != 0;
	if (!sorted || top > 17) {
		const double e_high = E[u_high];
		return (p <= e_low)    ? 0
		       : (p >= e_high) ? n_gp - 1
		                       : grid_search_nuclide<1>(p, E, u_low, u_high);
	}
	int32_t c = 0;
#pragma unroll
	for (int m = 0; m < 10; m++)
		c += (m <= top && v[m] <= p) ? 1 : 0;
	if (top > 9) {
		// in bounds: b + 17 <= n_gp + 15 < e_pitch
		const double2 *__restrict__ w = reinterpret_cast<const double2 *>(E + b);
#pragma unroll
		for (int q = 5; q < 9; q++) {
			const double2 x = w[q];
			c += (2 * q <= top && x.x <= p) ? 1 : 0;
			c += (2 * q + 1 <= top && x.y <= p) ? 1 : 0;
		}
	}
	return !(e_low < p) ? 0 : (c == top + 1) ? n_gp - 1 : b + c - 1;
}

// mrow: the material's unionized index row (ueg_row), j: the nuclide's
// position in the material list; both used on the unionized grid only.
// COUNT (hash grid): lanes whose bins differ (the history rounds) take
// hash_lower_window; where a wave's lanes mostly share a bin the
// bisection's probes coalesce and are cheaper.
template <int GRID, int STRIDE, bool COUNT = false>
__device__ __forceinline__ int32_t nuclide_lower(const Tables &T, const double *__restrict__ E_row,
                                                 int nuc, const uint8_t *__restrict__ mrow, int j,
                                                 int64_t idx, double p)
{
	const int32_t n_gp = T.n_gp;
	int32_t lower;
	if (GRID == kHash && COUNT) {
		lower = hash_lower_window(T, nuc, idx, p);
	} else if (GRID == kUnionized) {
		lower = ueg_bracket(T, mrow, j);
	} else if (GRID == kNuclide) {
		lower = grid_search_nuclide<STRIDE>(p, E_row, 0, n_gp - 1);
	} else {
		const int32_t u_low = T.index_grid[idx * T.idx_pitch + nuc];
		const int32_t u_high = (idx == T.hash_bins - 1)
		                           ? n_gp - 1
		                           : T.index_grid[(idx + 1) * T.idx_pitch + nuc] + 1;
		const double e_low = E_row[(int64_t)u_low * STRIDE];
		const double e_high = E_row[(int64_t)u_high * STRIDE];
		if (p <= e_low) {
			lower = 0;
		} else if (p >= e_high) {
			lower = n_gp - 1;
		} else {
			lower = grid_search_nuclide<STRIDE>(p, E_row, u_low, u_high);
		}
	}
	if (lower == n_gp - 1)
		lower--;
	return lower;
}

// Grid-level index for a lookup: unionized-grid bracket or hash bin
// (openmp-threading/Simulation.c:346-352).  The hash bin keeps the
// reference's division p / (1.0 / bins) and truncation; it is clamped to
// [0, bins - 1] so that p == 1.0 (which the reference would read past its
// table with), or a caller's negative or NaN energy, stays in bounds.
template <int GRID>
__device__ __forceinline__ int64_t lookup_index(const Tables &T, double p)
{
	if (GRID == kUnionized) {
		if (p >= 0.0 && p < 1.0) {
			const int b = (int)(p * (double)(1 << kUegAccBits));  // exact scaling, truncation
			return grid_search_in((int32_t)T.n_union, p, T.ueg, T.ueg_acc[b], T.ueg_acc[b + 1]);
		}
		return grid_search((int32_t)T.n_union, p, T.ueg);
	}
	if (GRID == kHash) {
		const double q = p / T.hash_du;
		if (!(q >= 0.0))
			return 0;
		return q < (double)(T.hash_bins - 1) ? (int32_t)q : T.hash_bins - 1;
	}
	return -1;
}

// FP64 interpolation of the 5 channels between points lo and hi
// (openmp-threading/Simulation.c:306-322) times the concentration: one
// nuclide's terms of calculate_macro_xs (Simulation.c:372-373).
__device__ __forceinline__ void term_fp64(const double *__restrict__ pt, double p, double conc,
                                          double t[5])
{
	const double2 *q = reinterpret_cast<const double2 *>(pt);
	const double2 a = q[0], b = q[1], c = q[2];  // lo: E t | el ab | fi nf
	const double2 d = q[3], e = q[4], f = q[5];  // hi: E t | el ab | fi nf
	const double fr = (d.x - p) / (d.x - a.x);
	t[0] = (d.y - fr * (d.y - a.y)) * conc;
	t[1] = (e.x - fr * (e.x - b.x)) * conc;
	t[2] = (e.y - fr * (e.y - b.y)) * conc;
	t[3] = (f.x - fr * (f.x - c.x)) * conc;
	t[4] = (f.y - fr * (f.y - c.y)) * conc;
}

// The terms accumulated into macro in the reference's order.
__device__ __forceinline__ void accumulate_fp64(const double *__restrict__ pt, double p, double conc,
                                                double macro[5])
{
	double t[5];
	term_fp64(pt, p, conc, t);
#pragma unroll
	for (int k = 0; k < 5; k++)
		macro[k] += t[k];
}

// The tables as the constant address space: a wave-uniform load through
// such a pointer becomes a scalar-cache load even in kernels that also
// write global memory (the persistent batched kernel's atomics), where the
// compiler cannot otherwise show that a table load is not clobbered.  The
// tables are read-only while any kernel runs.
typedef const __attribute__((address_space(4))) double *const_table_ptr;
typedef const __attribute__((address_space(4))) int32_t *const_index_ptr;
__device__ __forceinline__ const_table_ptr as_table(const double *p)
{
	return (const_table_ptr)p;
}
__device__ __forceinline__ const_index_ptr as_table(const int32_t *p)
{
	return (const_index_ptr)p;
}

// accumulate_fp64 from an interval record (Tables::ivl): the differences
// hi - lo were rounded once when the record was built, as the reference
// rounds them per lookup, so the result is the same.  Every operand but p
// and the sums may come from the scalar cache: no copies into VGPRs.
__device__ __forceinline__ void accumulate_ivl(const double *__restrict__ R, double p, double conc,
                                               double macro[5])
{
	const const_table_ptr r = as_table(R);  // hi: E t el ab fi nf, then hi - lo
	const double fr = (r[0] - p) / r[6];
	macro[0] += (r[1] - fr * r[7]) * conc;
	macro[1] += (r[2] - fr * r[8]) * conc;
	macro[2] += (r[3] - fr * r[9]) * conc;
	macro[3] += (r[4] - fr * r[10]) * conc;
	macro[4] += (r[5] - fr * r[11]) * conc;
}

// calculate_macro_xs (openmp-threading/Simulation.c:326-375), FP64.
template <int GRID, bool COUNT = false>
__device__ __forceinline__ void macro_xs_fp64(const Tables &T, double p, int mat, int64_t idx,
                                              double macro[5])
{
#pragma unroll
	for (int c = 0; c < 5; c++)
		macro[c] = 0.0;
	const int nn = T.num_nucs[mat];
	const int32_t *__restrict__ ml = T.mats + mat * T.max_num_nucs;
	const double *__restrict__ cl = T.concs + mat * T.max_num_nucs;
	const uint8_t *__restrict__ mrow = (GRID == kUnionized) ? ueg_row(T, mat, idx) : nullptr;
	auto step = [&](int j) {
		const int nuc = ml[j];
		const double conc = cl[j];
		const double *row = T.grid + (int64_t)nuc * T.n_gp * kPointDoubles;
		const int32_t lower = nuclide_lower<GRID, kPointDoubles, COUNT>(T, row, nuc, mrow, j, idx, p);
		accumulate_fp64(row + (int64_t)lower * kPointDoubles, p, conc, macro);
	};
	if constexpr (COUNT) {
		// two nuclides per trip: their load chains overlap (the history
		// rounds' lanes rarely share data, so each lane waits on its own)
#pragma unroll 2
		for (int j = 0; j < nn; j++)
			step(j);
	} else {
		for (int j = 0; j < nn; j++)
			step(j);
	}
}

// Points k0, k0+1, k0+2 of one nuclide are shared by the wave (scalar loads);
// each lane interpolates between points sel and sel+1 (sel = 0 or 1).
__device__ __forceinline__ void accumulate_fp64_select(const double *__restrict__ P, int sel,
                                                       double p, double conc, double macro[5])
{
	double a[3 * kPointDoubles];
#pragma unroll
	for (int w = 0; w < 3 * kPointDoubles; w++)
		a[w] = P[w];
	double lo[kPointDoubles], hi[kPointDoubles];
#pragma unroll
	for (int w = 0; w < kPointDoubles; w++) {
		lo[w] = sel ? a[kPointDoubles + w] : a[w];
		hi[w] = sel ? a[2 * kPointDoubles + w] : a[kPointDoubles + w];
	}
	const double fr = (hi[0] - p) / (hi[0] - lo[0]);
#pragma unroll
	for (int c = 0; c < 5; c++)
		macro[c] += (hi[1 + c] - fr * (hi[1 + c] - lo[1 + c])) * conc;
}

// Interval records k0 and k0+1 of one nuclide are shared by the wave (scalar
// loads); each lane needs record sel (0 or 1).  Both interpolations run on
// every lane with the records' values straight from SGPRs and each lane keeps
// its own: gfx950 reads one SGPR per VALU instruction, so selecting the
// records' 24 dwords per lane would cost three instructions each, more than
// the second interpolation.  The kept sums are those of the lane's record
// alone, bit for bit.
__device__ __forceinline__ void accumulate_ivl_select(const double *__restrict__ R, uint32_t sel,
                                                      double p, double conc, double macro[5])
{
	double a[5], b[5];
#pragma unroll
	for (int c = 0; c < 5; c++) {
		a[c] = macro[c];
		b[c] = macro[c];
	}
	accumulate_ivl(R, p, conc, a);
	accumulate_ivl(R + kIvlDoubles, p, conc, b);
#pragma unroll
	for (int c = 0; c < 5; c++)
		macro[c] = sel ? b[c] : a[c];
}

// One nuclide of the wave-uniform paths, bracket k per lane.  The grid data
// come through the scalar cache when the lanes' brackets coincide (one
// interval for the wave), and with SHARE3 also when they span two
// neighbouring intervals (each lane picks its own); else per lane.  IVL
// reads the interval records (unionized grid: brackets stored with the
// lower == n_gp - 1 rule applied), else the points (hash grid: the rule is
// applied here).  The scheduling barrier after a two-interval step keeps the
// compiler from overlapping several of them (which held ~200 VGPRs live
// across the unrolled chunk: two waves per SIMD instead of eight).
template <bool SHARE3, bool IVL>
__device__ __forceinline__ void ueg_wave_nuclide(const Tables &T, int nuc, int k, double p,
                                                 double conc, double macro[5])
{
	const int32_t n_gp = T.n_gp;
	if (!IVL && k == n_gp - 1)
		k--;
	const int k0 = __builtin_amdgcn_readfirstlane(k);
	const uint32_t d = (uint32_t)(k - k0);
	const int64_t i0 = (int64_t)nuc * n_gp + k0;
	if (IVL) {
		const double *R = T.ivl + i0 * kIvlDoubles;
		if (__ballot(d != 0u) == 0ULL) {
			// the barrier keeps this call from being merged with the per-lane
			// one below into one call on a lane-varying pointer (vector loads)
			accumulate_ivl(R, p, conc, macro);
			__builtin_amdgcn_sched_barrier(0);
		} else if (SHARE3 && __ballot(d > 1u) == 0ULL) {
			accumulate_ivl_select(R, d, p, conc, macro);
			__builtin_amdgcn_sched_barrier(0);
		} else if (SHARE3 && __ballot(d + 1u > 1u) == 0ULL) {
			accumulate_ivl_select(R - kIvlDoubles, d + 1u, p, conc, macro);
			__builtin_amdgcn_sched_barrier(0);
		} else {
			accumulate_ivl(T.ivl + ((int64_t)nuc * n_gp + k) * kIvlDoubles, p, conc, macro);
		}
		return;
	}
	const double *P = T.grid + i0 * kPointDoubles;
	if (__ballot(d != 0u) == 0ULL) {
		accumulate_fp64(P, p, conc, macro);
	} else if (SHARE3 && __ballot(d > 1u) == 0ULL) {
		accumulate_fp64_select(P, (int)d, p, conc, macro);
		__builtin_amdgcn_sched_barrier(0);
	} else if (SHARE3 && __ballot(d + 1u > 1u) == 0ULL) {
		accumulate_fp64_select(P - kPointDoubles, (int)(d + 1u), p, conc, macro);
		__builtin_amdgcn_sched_barrier(0);
	} else {
		accumulate_fp64(T.grid + ((int64_t)nuc * n_gp + k) * kPointDoubles, p, conc, macro);
	}
}

// Walks material m's nuclides in list order for the wave-uniform unionized
// paths, calling f(j, nuclide, bracket k of this lane) for each.  The lane's
// brackets are contiguous in its row of the material's index table, so they
// are fetched as 16-byte vectors (8 uint16 or 4 int32 brackets each), CHUNK
// (the staging depth, a multiple of 8) at a time, all in flight together.
// The element width is a wave-uniform branch per chunk.
// Eight uint16 brackets of one 16-byte vector as two int4 of brackets.
__device__ __forceinline__ void unpack_u16x8(const int4 w, int4 &lo, int4 &hi)
{
	lo = make_int4((int)((uint32_t)w.x & 0xFFFFu), (int)((uint32_t)w.x >> 16),
	               (int)((uint32_t)w.y & 0xFFFFu), (int)((uint32_t)w.y >> 16));
	hi = make_int4((int)((uint32_t)w.z & 0xFFFFu), (int)((uint32_t)w.z >> 16),
	               (int)((uint32_t)w.w & 0xFFFFu), (int)((uint32_t)w.w >> 16));
}

template <int CHUNK, class F>
__device__ __forceinline__ void ueg_wave_rows(const Tables &T, int m, int64_t idx, F &&f)
{
	static_assert(CHUNK % 8 == 0, "staging depth is whole 16-byte vectors of either width");
	const int nn = as_table(T.num_nucs)[m];
	const const_index_ptr ml = as_table(T.mats) + m * T.max_num_nucs;
	const uint8_t *__restrict__ row = ueg_row(T, m, idx);
	const bool narrow = T.midx_shift == 1;
	int j0 = 0;
	for (; j0 + CHUNK <= nn; j0 += CHUNK) {
		int4 kk[CHUNK / 4];
		if (narrow) {
			const int4 *__restrict__ v = reinterpret_cast<const int4 *>(row) + j0 / 8;
			int4 w[CHUNK / 8];
#pragma unroll
			for (int u = 0; u < CHUNK / 8; u++)
				w[u] = v[u];
#pragma unroll
			for (int u = 0; u < CHUNK / 8; u++)
				unpack_u16x8(w[u], kk[2 * u], kk[2 * u + 1]);
		} else {
			const int4 *__restrict__ v = reinterpret_cast<const int4 *>(row) + j0 / 4;
#pragma unroll
			for (int u = 0; u < CHUNK / 4; u++)
				kk[u] = v[u];
		}
#pragma unroll
		for (int u = 0; u < CHUNK / 4; u++) {
			const int j = j0 + 4 * u;
			f(j, ml[j], kk[u].x);
			f(j + 1, ml[j + 1], kk[u].y);
			f(j + 2, ml[j + 2], kk[u].z);
			f(j + 3, ml[j + 3], kk[u].w);
		}
	}
	// rows are padded to 16 bytes, so the last 4-entry vector is in bounds
	for (; j0 < nn; j0 += 4) {
		int4 k4;
		if (narrow) {
			const uint2 w = reinterpret_cast<const uint2 *>(row)[j0 / 4];
			k4 = make_int4((int)(w.x & 0xFFFFu), (int)(w.x >> 16), (int)(w.y & 0xFFFFu),
			               (int)(w.y >> 16));
		} else {
			k4 = reinterpret_cast<const int4 *>(row)[j0 / 4];
		}
		f(j0, ml[j0], k4.x);
		if (j0 + 1 < nn)
			f(j0 + 1, ml[j0 + 1], k4.y);
		if (j0 + 2 < nn)
			f(j0 + 2, ml[j0 + 2], k4.z);
		if (j0 + 3 < nn)
			f(j0 + 3, ml[j0 + 3], k4.w);
	}
}

// calculate_macro_xs, unionized grid, FP64, for a wave whose lanes all hold
// material m (the material-sorted schedules).  The material list, the
// concentrations and the shared grid points are wave-uniform (scalar
// cache).  Arithmetic and summation order are those of macro_xs_fp64.
template <int CHUNK, bool SHARE3 = true>
__device__ __forceinline__ void macro_ueg_wave(const Tables &T, double p, int m, int64_t idx,
                                               double macro[5])
{
#pragma unroll
	for (int c = 0; c < 5; c++)
		macro[c] = 0.0;
	const const_table_ptr cl = as_table(T.concs) + m * T.max_num_nucs;
	ueg_wave_rows<CHUNK>(T, m, idx, [&](int j, int nuc, int k) {
		ueg_wave_nuclide<SHARE3, true>(T, nuc, k, p, cl[j], macro);
	});
}

constexpr int kIdxChunk = 8;

// Hash grid: the bracket of one nuclide for a lookup in bin idx
// (Simulation.c:273-302), with the bin's bounds u_low / u_high given.  When
// every lane of the wave shares the bin, the bounds and their energies are
// wave-uniform (scalar loads) and only the short bisection is per lane.
// With sorted rows and a short range, the bisection's result (the last index
// in [u_low, u_high) whose energy is <= p) is u_low plus the number of
// interior energies <= p: those are read with scalar loads (the range is the
// wave's) and counted, with no chain of per-lane probes.
template <int STRIDE>
__device__ __forceinline__ int32_t hash_bracket(const double *__restrict__ E_row, int32_t u_low,
                                                int32_t u_high, int32_t n_gp, double p,
                                                bool sorted_rows)
{
	constexpr int32_t kCountMax = 16;
	const double e_low = E_row[(int64_t)u_low * STRIDE];
	const double e_high = E_row[(int64_t)u_high * STRIDE];
	if (p <= e_low)
		return 0;
	if (p >= e_high)
		return n_gp - 1;
	if (!sorted_rows || u_high - u_low > kCountMax)
		return grid_search_nuclide<STRIDE>(p, E_row, u_low, u_high);
	int32_t lower = u_low;
	for (int32_t i = u_low + 1; i < u_high; i++)
		lower += (E_row[(int64_t)i * STRIDE] <= p) ? 1 : 0;
	return lower;
}

// Walks material m's nuclides for a wave whose lanes share the material and
// the hash bin idx, calling f(j, nuclide, bracket k of this lane).
template <int STRIDE, class F>
__device__ __forceinline__ void hash_wave_rows(const Tables &T, const double *__restrict__ base,
                                               double p, int m, int64_t idx, F &&f)
{
	const int nn = T.num_nucs[m];
	const int32_t n_gp = T.n_gp;
	const int32_t *__restrict__ ml = T.mats + m * T.max_num_nucs;
	const int32_t *__restrict__ lo_row = T.index_grid + idx * T.idx_pitch;
	const bool last = idx == T.hash_bins - 1;
	const bool sorted_rows = (T.variant & kVariantSortedRows) != 0;
#pragma unroll 4
	for (int j = 0; j < nn; j++) {
		const int nuc = ml[j];
		const int32_t u_low = lo_row[nuc];
		const int32_t u_high = last ? n_gp - 1 : lo_row[T.idx_pitch + nuc] + 1;
		const double *E_row = base + (int64_t)nuc * n_gp * STRIDE;
		f(j, nuc, hash_bracket<STRIDE>(E_row, u_low, u_high, n_gp, p, sorted_rows));
	}
}

// calculate_macro_xs, hash grid, FP64, for a wave whose lanes share the
// material and the hash bin.  Arithmetic and summation order are those of
// macro_xs_fp64.
template <bool SHARE3 = true>
__device__ __forceinline__ void macro_hash_wave(const Tables &T, double p, int m, int64_t idx,
                                                double macro[5])
{
#pragma unroll
	for (int c = 0; c < 5; c++)
		macro[c] = 0.0;
	const double *__restrict__ cl = T.concs + m * T.max_num_nucs;
	hash_wave_rows<kPointDoubles>(T, T.grid, p, m, idx, [&](int j, int nuc, int k) {
		ueg_wave_nuclide<SHARE3, false>(T, nuc, k, p, cl[j], macro);
	});
}

// argmax with max = -1.0 and strict '>' (Simulation.c:100-110)
__device__ __forceinline__ int argmax5(const double m[5])
{
	double mx = -1.0;
	int am = 0;
#pragma unroll
	for (int j = 0; j < 5; j++) {
		if (m[j] > mx) {
			mx = m[j];
			am = j;
		}
	}
	return am;
}

// Block order for kernels whose neighbouring blocks share data: the
// dispatcher sends block b to XCD b % 8, so consecutive logical blocks are
// regrouped into runs of `run` blocks per XCD, and each XCD's L2 sees a
// contiguous stretch of the (sorted) work.  Bijective on [0, gridDim.x).
__device__ __forceinline__ unsigned xcd_block(unsigned run)
{
	const unsigned b = blockIdx.x, G = gridDim.x, per = 8u * run;
	const unsigned R = (G / per) * per;
	if (b >= R)
		return b;
	const unsigned xcd = b & 7u, i = b >> 3;
	return (i / run) * per + xcd * run + (i % run);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off, 64);
	return v;
}

// Sum v over the block and add it to *out with one device-scope atomic.
template <int BLOCK>
__device__ __forceinline__ void block_add_u64(uint32_t v, unsigned long long *out)
{
	static_assert(BLOCK % 64 == 0, "block must be whole waves");
	__shared__ uint32_t part[BLOCK / 64];
	v = wave_sum_u32(v);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	__syncthreads();  // a previous call's reader (thread 0) is done with part[]
	if (lane == 0)
		part[wave] = v;
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long s = 0;
#pragma unroll
		for (int w = 0; w < BLOCK / 64; w++)
			s += part[w];
		if (s)
			atomicAdd(out, s);
	}
}

}  // namespace xsb
