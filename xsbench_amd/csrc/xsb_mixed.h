// xsb_mixed.h -- mixed-precision macroscopic XS (FP32 or BF16 XS tables,
// FP64 energies, FP32 terms summed in FP64) with an FP64 fallback that keeps
// every decision identical to the FP64 reference path.
//
// Table records (built on the device from the FP64 AoS at upload):
//   FP32: 32 B = { double E; float xs[5]; float pad; }
//   BF16: 24 B = { double E; bf16 xs[5]; u16 pad[3]; }  xs scaled by a power
//         of two per nuclide, s_n >= max |xs| of that nuclide
// Energies stay FP64 so that brackets and the interpolation fraction are the
// reference's exactly; only the 5 channel values are rounded.
//
// Each nuclide's term is conc * ((1 - f) xs_hi + f xs_lo), the reference's
// hi - f (hi - lo) regrouped, with f = (E_hi - p) / (E_hi - E_lo) and 1 - f both
// formed in FP64 and rounded once to FP32: the two weights carry the relative
// error of one rounding each, so a term's error is relative to the term.
//
// Error bounds.  With u the unit roundoff of the table format (2^-24 FP32,
// 2^-8 BF16), u32 = 2^-24, u64 = 2^-53, M_j an upper bound of |xs| for
// nuclide j and f_j the fraction, every FP32 term of ours differs from the
// reference's conc_j (hi - f_j (hi - lo)) by at most
// conc_j (|hi| |1 - f_j| + |lo| |f_j|) (u + 5 u32 + 8 u64) + 5 u64 conc_j M_j (the
// table values, the two weights, the concentration, the weights' products and
// the term's two roundings; the weights' FP64 fraction within a few u64 of
// the reference's), the FP64 sum of the terms adds (n-1) u64 sum|t|, and the
// reference's own ordered FP64 sum differs from the exact sum by
// (n-1) u64 sum|r|.  So, for every channel, |S_c - R_c| <= B with
//   B = 1.01 * A * (u + 12 u32),   A = sum_j conc_j M_j (1 + 2|f_j|)
// (the 1.01 absorbs the FP32 evaluation of A itself and the rounding of
// conc_j M_j, for n < 10^4).  When the tables hold no negative XS or
// concentration (checked at upload) and every weight of the lookup lies in
// [0, 1] (p inside each bracket: always on the unionized grid), each term is
// non-negative and the first part sums to (u + 5 u32) sum_j T_j with T_j >= 0
// the exact terms, i.e. channel by channel
//   b_c = 1.01 * (u + 8 u32) |S_c| + kRefTol * A
// (kRefTol covers the FP64 roundings, (2n + 13) u64 A, for n < 4 * 10^4;
// u + 8 u32 covers BF16's double rounding FP64 -> FP32 -> BF16).
// Otherwise b_c = B.  A lookup is recomputed in FP64 when its argmax is not
// separated (S_am - S_k <= b_am + b_k for some k), in history mode when some
// |S_c - 1| <= b_c, and, with FP32 tables, when some b_c > kRelTol |S_c|
// (sums whose terms nearly cancel, from extrapolation outside a nuclide's
// points): every FP32 result is then within relative error kRelTol of the
// FP64 reference, channel by channel, and the hash is the reference's.  B is
// a worst case over signs and rounding directions (about 8e-7 A for FP32);
// the errors actually seen are ~3e-8 relative on average.  A recomputed
// lookup returns the FP64 sums.
#pragma once

#include "xsb_device.h"

namespace xsb {

constexpr int kRec32Doubles = 4;  // 32 B record
constexpr double kRelTol = 1e-6;  // FP32 tables: certified relative error per channel
constexpr double kRefTol = 1e-11; // the FP64 roundings of both sums, relative to A
constexpr int kRec16Doubles = 3;  // 24 B record
constexpr int kIvl32Doubles = 8;  // 64 B FP32 interval record (Tables::ivl32)

// Per (material, position) constants for the mixed kernels, [12][max_num_nucs]
// float2 {c_eff, c_bound}: c_eff = conc * scale (scale = 1 for FP32),
// c_bound = conc * M (rounded up) for the error bound.  Stored in T.concs32
// as two consecutive [12][max] float planes per precision:
//   plane 0/1: FP32 c_eff / c_bound, plane 2/3: BF16 c_eff / c_bound.

__device__ __forceinline__ float bf16_to_f32(uint32_t h)
{
	return __uint_as_float(h << 16);
}

template <int PREC>
__device__ __forceinline__ const double *mixed_row(const Tables &T, int nuc)
{
	if (PREC == XSB_XS_FP32)
		return reinterpret_cast<const double *>(T.xs32) + (int64_t)nuc * T.n_gp * kRec32Doubles;
	return reinterpret_cast<const double *>(T.xs16) + (int64_t)nuc * T.n_gp * kRec16Doubles;
}

// Load the 5 channel values of a record.
template <int PREC>
__device__ __forceinline__ void load_xs(const double *rec, float x[5])
{
	if (PREC == XSB_XS_FP32) {
		const float4 a = *reinterpret_cast<const float4 *>(rec + 1);  // xs0..xs3
		const float b = *reinterpret_cast<const float *>(rec + 3);    // xs4
		x[0] = a.x;
		x[1] = a.y;
		x[2] = a.z;
		x[3] = a.w;
		x[4] = b;
	} else {
		const uint2 w = *reinterpret_cast<const uint2 *>(rec + 1);          // xs0..xs3
		const uint32_t w2 = *reinterpret_cast<const uint32_t *>(rec + 2);   // xs4 | pad
		x[0] = bf16_to_f32(w.x & 0xFFFFu);
		x[1] = bf16_to_f32(w.x >> 16);
		x[2] = bf16_to_f32(w.y & 0xFFFFu);
		x[3] = bf16_to_f32(w.y >> 16);
		x[4] = bf16_to_f32(w2 & 0xFFFFu);
	}
}

// FP32 records lo and lo + 1 (32-byte aligned, adjacent: {E, xs[5], pad}
// each) as four 16-byte loads.
__device__ __forceinline__ void load_pair32(const double *lo, double &e_lo, float xl[5],
                                            double &e_hi, float xh[5])
{
	const uint4 *q = reinterpret_cast<const uint4 *>(lo);
	const uint4 a = q[0], b = q[1], c = q[2], d = q[3];
	e_lo = __hiloint2double((int)a.y, (int)a.x);
	xl[0] = __uint_as_float(a.z);
	xl[1] = __uint_as_float(a.w);
	xl[2] = __uint_as_float(b.x);
	xl[3] = __uint_as_float(b.y);
	xl[4] = __uint_as_float(b.z);
	e_hi = __hiloint2double((int)c.y, (int)c.x);
	xh[0] = __uint_as_float(c.z);
	xh[1] = __uint_as_float(c.w);
	xh[2] = __uint_as_float(d.x);
	xh[3] = __uint_as_float(d.y);
	xh[4] = __uint_as_float(d.z);
}

// The two FP32 weights of a nuclide's term from the FP64 fraction f:
// fr = RN(f), gr = RN(1 - f); wmin keeps the smallest weight of the lookup
// (every term is a non-negative combination when it stays >= 0).
__device__ __forceinline__ void mixed_weights(double f64, float &fr, float &gr, float &wmin)
{
	fr = (float)f64;
	gr = (float)(1.0 - f64);
	wmin = fminf(wmin, fminf(fr, gr));
}

// Adds one nuclide's FP32 term c ((1 - f) xh + f xl) to the FP64 sums and its
// bound weight c_bound (1 + 2|f|) to A.
__device__ __forceinline__ void mixed_add(const float xl[5], const float xh[5], float fr, float gr,
                                          float c, float cb, double sum[5], float &A)
{
	const float fc = fr * c, gc = gr * c;
#pragma unroll
	for (int k = 0; k < 5; k++)
		sum[k] += (double)__builtin_fmaf(xl[k], fc, xh[k] * gc);
	A += cb * (1.0f + 2.0f * fabsf(fr));
}

// The FP64 fraction (e_hi - p) / (e_hi - e_lo) through the hardware
// reciprocal and two Newton steps: relative error far below 2^-40 (the
// bounds above leave 2 u32 of slack), a few instructions instead of the
// correctly rounded division; a zero width gives an infinite or NaN fraction,
// and so an infinite or NaN bound: the lookup falls back to FP64.
__device__ __forceinline__ double mixed_fraction(double e_lo, double e_hi, double p)
{
	const double w = e_hi - e_lo;
	double r = __builtin_amdgcn_rcp(w);
	r = __builtin_fma(r, __builtin_fma(-w, r, 1.0), r);
	r = __builtin_fma(r, __builtin_fma(-w, r, 1.0), r);
	return (e_hi - p) * r;
}

// One nuclide's term between the points (e_lo, xl) and (e_hi, xh).
__device__ __forceinline__ void mixed_term(double e_lo, const float xl[5], double e_hi,
                                           const float xh[5], double p, float c, float cb,
                                           double sum[5], float &A, float &wmin)
{
	float fr, gr;
	mixed_weights(mixed_fraction(e_lo, e_hi, p), fr, gr, wmin);
	mixed_add(xl, xh, fr, gr, c, cb, sum, A);
}

// One nuclide's term from an FP32 interval record R = {E_hi, RN(1 / (E_hi -
// E_lo)), xs_hi[5], xs_lo[5]}: the FP64 fraction is the numerator times the
// rounded reciprocal, within 3 u64 (relative) of the quotient; the XS are the
// points' own FP32 values.
// (The records are read through the constant address space, as_table: a
// record shared by the wave comes through the scalar cache.)
__device__ __forceinline__ double ivl32_fraction(const double *__restrict__ R, double p)
{
	const const_table_ptr r = as_table(R);
	return (r[0] - p) * r[1];
}

__device__ __forceinline__ void ivl32_terms(const double *__restrict__ R, float fc, float gc,
                                            float t[5])
{
	typedef const __attribute__((address_space(4))) float *const_float_ptr;
	const const_float_ptr f = (const_float_ptr)as_table(R + 2);  // xh[5], xl[5]
#pragma unroll
	for (int k = 0; k < 5; k++)
		t[k] = __builtin_fmaf(f[5 + k], fc, f[k] * gc);
}

// mixed_term on an FP32 interval record: the record shared by the wave (R),
// or one of the two consecutive records Q, Q + kIvl32Doubles (the second
// where sel).  With two, both terms are formed straight from the shared
// (scalar) records and each lane keeps its own, which is cheaper than first
// selecting the records' 30 dwords per lane at three instructions each.
__device__ __forceinline__ void mixed_term_ivl(const double *__restrict__ R, double p, float c,
                                               float cb, double sum[5], float &A, float &wmin)
{
	float fr, gr;
	mixed_weights(ivl32_fraction(R, p), fr, gr, wmin);
	float t[5];
	ivl32_terms(R, fr * c, gr * c, t);
#pragma unroll
	for (int k = 0; k < 5; k++)
		sum[k] += (double)t[k];
	A += cb * (1.0f + 2.0f * fabsf(fr));
}

__device__ __forceinline__ void mixed_term_ivl2(const double *__restrict__ Q, bool sel, double p,
                                                float c, float cb, double sum[5], float &A,
                                                float &wmin)
{
	const double fa64 = ivl32_fraction(Q, p), fb64 = ivl32_fraction(Q + kIvl32Doubles, p);
	const float fa = (float)fa64, ga = (float)(1.0 - fa64);
	const float fb = (float)fb64, gb = (float)(1.0 - fb64);
	float ta[5], tb[5];
	ivl32_terms(Q, fa * c, ga * c, ta);
	ivl32_terms(Q + kIvl32Doubles, fb * c, gb * c, tb);
#pragma unroll
	for (int k = 0; k < 5; k++)
		sum[k] += (double)(sel ? tb[k] : ta[k]);
	const float fr = sel ? fb : fa, gr = sel ? gb : ga;
	A += cb * (1.0f + 2.0f * fabsf(fr));
	wmin = fminf(wmin, fminf(fr, gr));
}

// The decision on the mixed sums (see the bounds above): returns the argmax
// the FP64 reference would return; macro[] receives the mixed-precision sums,
// or the FP64 sums when the lookup fell back (*fell_back = 1); *bound the
// largest channel bound.  nonneg: the tables hold no negative XS or
// concentration.  fp64(m) writes this lane's FP64 sums.  WAVE: fp64 is a
// wave-cooperative path that every lane of the wave must enter, so it runs
// for the whole wave as soon as one lane needs it (cheaper than one lane's
// FP64 loop with its wave idle).  DEFER: fp64 is not called; the caller
// re-evaluates the lookups with *fell_back = 1 itself.
template <int PREC, bool HISTORY, bool WAVE, bool DEFER = false, class F64>
__device__ __forceinline__ int mixed_finish(const double s[5], float A, float wmin, bool nonneg,
                                            double macro[5], uint32_t *fell_back, double *bound,
                                            F64 &&fp64)
{
	constexpr double kU = (PREC == XSB_XS_FP32) ? 0x1p-24 : 0x1p-8;
	// argmax of the mixed sums (max = -1.0, strict '>', as the reference)
	int am = argmax5(s);
	const double B = 1.01 * (double)A * (kU + 12.0 * 0x1p-24);
	const bool rel = nonneg && wmin >= 0.f;
	double b[5], bmax = 0.0;
#pragma unroll
	for (int k = 0; k < 5; k++) {
		b[k] = rel ? 1.01 * (kU + 8.0 * 0x1p-24) * fabs(s[k]) + kRefTol * (double)A : B;
		bmax = fmax(bmax, b[k]);
	}
	if (bound)
		*bound = rel ? bmax : B;
	bool ambiguous = false;
#pragma unroll
	for (int k = 0; k < 5; k++) {
		if (k != am && !(s[am] - s[k] > b[am] + b[k]))
			ambiguous = true;
		if (HISTORY && !(fabs(s[k] - 1.0) > b[k]))
			ambiguous = true;
		// b_c <= kRelTol (1 - kRelTol) |S_c| certifies |S_c - R_c| <= kRelTol |R_c|
		if (PREC == XSB_XS_FP32 && !(b[k] <= kRelTol * (1.0 - kRelTol) * fabs(s[k])))
			ambiguous = true;
	}
	// the reference's argmax needs max > -1.0 for a non-zero index
	if (!(s[am] > -1.0 + b[am]))
		ambiguous = true;
#pragma unroll
	for (int k = 0; k < 5; k++)
		macro[k] = s[k];
	*fell_back = ambiguous ? 1u : 0u;
	if (DEFER || (WAVE ? __ballot(ambiguous) == 0ULL : !ambiguous))
		return am;
	double m64[5];
	fp64(m64);
	if (ambiguous) {
#pragma unroll
		for (int k = 0; k < 5; k++)
			macro[k] = m64[k];
		am = argmax5(m64);
	}
	return am;
}

typedef const __attribute__((address_space(4))) float *const_float_table_ptr;

template <int PREC>
__device__ __forceinline__ const_float_table_ptr mixed_consts(const Tables &T, int mat)
{
	const int plane = T.max_num_nucs * kMaterials;
	return (const_float_table_ptr)T.concs32 + (PREC == XSB_XS_FP32 ? 0 : 2) * plane + mat * T.max_num_nucs;
}

// calculate_macro_xs on the mixed tables, one lane per lookup.
template <int GRID, int PREC, bool HISTORY = false, bool DEFER = false>
__device__ __forceinline__ int macro_xs_mixed(const Tables &T, double p, int mat, int64_t idx,
                                              double macro[5], uint32_t *fell_back,
                                              double *bound = nullptr)
{
	constexpr int REC = (PREC == XSB_XS_FP32) ? kRec32Doubles : kRec16Doubles;
	const int nn = T.num_nucs[mat];
	const const_float_table_ptr ceff = mixed_consts<PREC>(T, mat);
	const const_float_table_ptr cbnd = ceff + T.max_num_nucs * kMaterials;
	const int32_t *__restrict__ ml = T.mats + mat * T.max_num_nucs;
	const uint8_t *__restrict__ mrow = (GRID == kUnionized) ? ueg_row(T, mat, idx) : nullptr;
	double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
	float A = 0.f, wmin = 1.f;
	auto step = [&](int j) {
		const int nuc = ml[j];
		const double *row = mixed_row<PREC>(T, nuc);
		const int32_t lower = nuclide_lower<GRID, REC, HISTORY>(T, row, nuc, mrow, j, idx, p);
		if (PREC == XSB_XS_FP32 && GRID == kUnionized) {  // FP32 interval records
			mixed_term_ivl(T.ivl32 + ((int64_t)nuc * T.n_gp + lower) * kIvl32Doubles, p, ceff[j],
			               cbnd[j], sum, A, wmin);
			return;
		}
		const double *lo = row + (int64_t)lower * REC;
		float xl[5], xh[5];
		if (PREC == XSB_XS_FP32) {
			double e_lo, e_hi;
			load_pair32(lo, e_lo, xl, e_hi, xh);
			mixed_term(e_lo, xl, e_hi, xh, p, ceff[j], cbnd[j], sum, A, wmin);
			return;
		}
		const double *hi = lo + REC;
		load_xs<PREC>(lo, xl);
		load_xs<PREC>(hi, xh);
		mixed_term(lo[0], xl, hi[0], xh, p, ceff[j], cbnd[j], sum, A, wmin);
	};
	if constexpr (HISTORY) {
		// two nuclides per trip, as macro_xs_fp64 in the history kernels
#pragma unroll 2
		for (int j = 0; j < nn; j++)
			step(j);
	} else {
		for (int j = 0; j < nn; j++)
			step(j);
	}
	return mixed_finish<PREC, HISTORY, false, DEFER>(sum, A, wmin, (T.variant & kVariantNonNeg) != 0,
	                                                 macro, fell_back, bound, [&](double m64[5]) {
		                                                 macro_xs_fp64<GRID, HISTORY>(T, p, mat, idx, m64);
	                                                 });
}

// One nuclide of the wave-uniform path on the mixed tables, bracket k per
// lane: records through the scalar cache when the lanes' brackets coincide,
// and with SHARE3 also when they span two neighbouring intervals, else per
// lane (as ueg_wave_nuclide).
// With IVL (unionized grid, FP32 tables) the FP32 interval records are read
// instead, as ueg_wave_nuclide does in FP64.
template <int PREC, bool SHARE3 = true, bool IVL = false>
__device__ __forceinline__ void ueg_wave_nuclide_mixed(const Tables &T, int nuc, int k, double p,
                                                       float c, float cb, double sum[5],
                                                       float &A, float &wmin)
{
	constexpr int REC = (PREC == XSB_XS_FP32) ? kRec32Doubles : kRec16Doubles;
	const int32_t n_gp = T.n_gp;
	if (k == n_gp - 1)
		k--;
	const int k0 = __builtin_amdgcn_readfirstlane(k);
	const int d = k - k0;
	if (IVL && PREC == XSB_XS_FP32) {
		const uint32_t du = (uint32_t)d;
		const double *R = T.ivl32 + ((int64_t)nuc * n_gp + k0) * kIvl32Doubles;
		if (__ballot(du != 0u) == 0ULL) {
			mixed_term_ivl(R, p, c, cb, sum, A, wmin);
			__builtin_amdgcn_sched_barrier(0);  // not merged with the per-lane call
		} else if (SHARE3 && (__ballot(du > 1u) == 0ULL || __ballot(du + 1u > 1u) == 0ULL)) {
			// two records shared; each lane takes its own
			const bool back = __ballot(du > 1u) != 0ULL;
			mixed_term_ivl2(back ? R - kIvl32Doubles : R, (back ? du + 1u : du) != 0u, p, c, cb,
			                sum, A, wmin);
			__builtin_amdgcn_sched_barrier(0);
		} else {
			mixed_term_ivl(T.ivl32 + ((int64_t)nuc * n_gp + k) * kIvl32Doubles, p, c, cb, sum, A,
			               wmin);
		}
		return;
	}
	const double *row = mixed_row<PREC>(T, nuc);
	const double *P = row + (int64_t)k0 * REC;
	float x0[5], x1[5];
	if (__ballot(d != 0) == 0ULL) {
		load_xs<PREC>(P, x0);
		load_xs<PREC>(P + REC, x1);
		mixed_term(P[0], x0, P[REC], x1, p, c, cb, sum, A, wmin);
	} else if (SHARE3 &&
	           (__ballot(d != 0 && d != 1) == 0ULL || __ballot(d != 0 && d != -1) == 0ULL)) {
		// three records Q, Q+1, Q+2 shared; each lane takes its pair
		const double *Q = (__ballot(d == -1) != 0ULL) ? P - REC : P;
		const int sel = (__ballot(d == -1) != 0ULL) ? d + 1 : d;
		float x2[5];
		load_xs<PREC>(Q, x0);
		load_xs<PREC>(Q + REC, x1);
		load_xs<PREC>(Q + 2 * REC, x2);
		const double e0 = Q[0], e1 = Q[REC], e2 = Q[2 * REC];
		float xl[5], xh[5];
#pragma unroll
		for (int w = 0; w < 5; w++) {
			xl[w] = sel ? x1[w] : x0[w];
			xh[w] = sel ? x2[w] : x1[w];
		}
		mixed_term(sel ? e1 : e0, xl, sel ? e2 : e1, xh, p, c, cb, sum, A, wmin);
		__builtin_amdgcn_sched_barrier(0);
	} else {
		const double *L = row + (int64_t)k * REC;
		load_xs<PREC>(L, x0);
		load_xs<PREC>(L + REC, x1);
		mixed_term(L[0], x0, L[REC], x1, p, c, cb, sum, A, wmin);
	}
}

// calculate_macro_xs, unionized grid, mixed tables, for a wave whose lanes
// all hold material m; same terms and decision as macro_xs_mixed.
template <int PREC, int CHUNK, bool HISTORY, bool SHARE3 = true, bool DEFER = false>
__device__ __forceinline__ int macro_ueg_wave_mixed(const Tables &T, double p, int m, int64_t idx,
                                                    double macro[5], uint32_t *fell_back,
                                                    double *bound)
{
	const const_float_table_ptr ceff = mixed_consts<PREC>(T, m);
	const const_float_table_ptr cbnd = ceff + T.max_num_nucs * kMaterials;
	double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
	float A = 0.f, wmin = 1.f;
	ueg_wave_rows<CHUNK>(T, m, idx, [&](int j, int nuc, int k) {
		ueg_wave_nuclide_mixed<PREC, SHARE3, true>(T, nuc, k, p, ceff[j], cbnd[j], sum, A, wmin);
	});
	return mixed_finish<PREC, HISTORY, true, DEFER>(sum, A, wmin, (T.variant & kVariantNonNeg) != 0,
	                                                macro, fell_back, bound, [&](double m64[5]) {
		                                                macro_ueg_wave<CHUNK, SHARE3>(T, p, m, idx, m64);
	                                                });
}

// calculate_macro_xs, hash grid, mixed tables, for a wave whose lanes share
// the material and the hash bin; same terms and decision as macro_xs_mixed.
template <int PREC, bool HISTORY, bool SHARE3 = true, bool DEFER = false>
__device__ __forceinline__ int macro_hash_wave_mixed(const Tables &T, double p, int m, int64_t idx,
                                                     double macro[5], uint32_t *fell_back,
                                                     double *bound)
{
	constexpr int REC = (PREC == XSB_XS_FP32) ? kRec32Doubles : kRec16Doubles;
	const const_float_table_ptr ceff = mixed_consts<PREC>(T, m);
	const const_float_table_ptr cbnd = ceff + T.max_num_nucs * kMaterials;
	double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
	float A = 0.f, wmin = 1.f;
	hash_wave_rows<REC>(T, mixed_row<PREC>(T, 0), p, m, idx, [&](int j, int nuc, int k) {
		ueg_wave_nuclide_mixed<PREC, SHARE3>(T, nuc, k, p, ceff[j], cbnd[j], sum, A, wmin);
	});
	return mixed_finish<PREC, HISTORY, true, DEFER>(sum, A, wmin, (T.variant & kVariantNonNeg) != 0,
	                                                macro, fell_back, bound, [&](double m64[5]) {
		                                                macro_hash_wave<SHARE3>(T, p, m, idx, m64);
	                                                });
}

}  // namespace xsb
