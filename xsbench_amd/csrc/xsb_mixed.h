// xsb_mixed.h -- mixed-precision macroscopic XS (FP32 or BF16 XS tables,
// FP64 energies, FP32 Kahan accumulation) with an FP64 fallback that keeps
// every decision identical to the FP64 reference path.
//
// Table records (built on the device from the FP64 AoS at upload):
//   FP32: 32 B = { double E; float xs[5]; float pad; }
//   BF16: 24 B = { double E; bf16 xs[5]; u16 pad[3]; }  xs scaled by a power
//         of two per nuclide, s_n >= max |xs| of that nuclide
// Energies stay FP64 so that brackets and the interpolation fraction are the
// reference's exactly; only the 5 channel values are rounded.
//
// Error bound (per lookup, shared by the 5 channels).  With u the unit
// roundoff of the table format (2^-24 FP32, 2^-8 BF16), u32 = 2^-24, M_j an
// upper bound of |xs| for nuclide j and f_j the FP64 fraction, every term
// conc_j * xs_j of the reference differs from ours by at most
// conc_j M_j (1 + 2|f_j|) (u + 7 u32 + 8 u64) (5 u64 with the quotient as
// fraction, 3 more with the interval records' reciprocal), and the Kahan sum adds at most
// 2 u32 |S| + n u32^2 sum|t|; the reference's own ordered FP64 sum differs from
// the exact sum by (n-1) u64 sum|r|.  So |S_c - R_c| <= B with
//   B = 1.01 * A * (u + 12 u32),   A = sum_j conc_j M_j (1 + 2|f_j|)
// (the 1.01 absorbs the FP32 evaluation of A itself and the rounding of
// conc_j M_j, for n < 10^5).  A lookup whose argmax gap is <= 2B, or (history
// mode) with some |S_c - 1| <= B, is recomputed in FP64, and so is, with
// FP32 tables, a lookup with some channel where B > kRelTol |S_c| (sums whose
// terms nearly cancel, from extrapolation below a nuclide's first point):
// every FP32 result is then within relative error kRelTol of the FP64
// reference, channel by channel.  B is a worst case over the term signs and
// rounding directions (about 8e-7 A); the errors actually seen are ~3e-8
// relative on average.  A recomputed lookup returns the FP64 sums.
#pragma once

#include "xsb_device.h"

namespace xsb {

constexpr int kRec32Doubles = 4;  // 32 B record
constexpr double kRelTol = 1e-5;  // FP32 tables: certified relative error per channel
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

// The Kahan step for a corrected term y = t - comp already formed.
__device__ __forceinline__ void kahan_add_corrected(float &sum, float &comp, float y)
{
	const float s = sum + y;
	comp = (s - sum) - y;
	sum = s;
}

// One nuclide's term: interpolate the 5 channels between (e_lo, xl) and
// (e_hi, xh) with the FP64 fraction, add x * c to the Kahan sums, and c_bound
// (1 + 2|f|) to the bound accumulator A.  The Kahan step's corrected term
// y = x c - comp is formed as (xh c - comp) - (fr c) d, d = xh - xl, with two
// fused operations: before the Kahan step's own rounding of y, the roundings
// of fr c and of xh c - comp err by at most u32 c M (1 + 2|f|) + u32 |comp|,
// where rounding x and x c separately may err by 2 u32 c M (1 + 2|f|); the
// bound above counts the latter.
__device__ __forceinline__ void mixed_term(double e_lo, const float xl[5], double e_hi,
                                           const float xh[5], double p, float c, float cb,
                                           float sum[5], float comp[5], float &A)
{
	const double fr64 = (e_hi - p) / (e_hi - e_lo);
	const float fr = (float)fr64;
	const float frc = fr * c;
#pragma unroll
	for (int k = 0; k < 5; k++)
		kahan_add_corrected(sum[k], comp[k],
		                    __builtin_fmaf(-frc, xh[k] - xl[k], __builtin_fmaf(xh[k], c, -comp[k])));
	A += cb * (1.0f + 2.0f * fabsf(fr));
}

// One nuclide's term from an FP32 interval record R, added to the Kahan sums
// with mixed_term's operations.  The FP32 differences were rounded once when
// the record was built, exactly as mixed_term rounds them; the FP64 fraction
// is the numerator times the record's rounded reciprocal width, within 3 u64
// (relative) of the quotient mixed_term takes.
__device__ __forceinline__ float ivl32_fraction(const double *__restrict__ R, double p)
{
	return (float)((R[0] - p) * R[1]);
}

__device__ __forceinline__ void ivl32_terms(const double *__restrict__ R, float frc, float c,
                                            const float comp[5], float y[5])
{
	const float *f = reinterpret_cast<const float *>(R + 2);  // xh[5], d[5]
#pragma unroll
	for (int k = 0; k < 5; k++)
		y[k] = __builtin_fmaf(-frc, f[5 + k], __builtin_fmaf(f[k], c, -comp[k]));
}

// mixed_term on an FP32 interval record: the record shared by the wave (R),
// or one of the two consecutive records Q, Q + kIvl32Doubles (the second
// where sel).  With two, both terms are formed straight from the shared
// (scalar) records and each lane keeps its own, which is cheaper than first
// selecting the records' 30 dwords per lane at three instructions each.
__device__ __forceinline__ void mixed_term_ivl(const double *__restrict__ R, double p, float c,
                                               float cb, float sum[5], float comp[5], float &A)
{
	const float fr = ivl32_fraction(R, p);
	float y[5];
	ivl32_terms(R, fr * c, c, comp, y);
#pragma unroll
	for (int k = 0; k < 5; k++)
		kahan_add_corrected(sum[k], comp[k], y[k]);
	A += cb * (1.0f + 2.0f * fabsf(fr));
}

__device__ __forceinline__ void mixed_term_ivl2(const double *__restrict__ Q, bool sel, double p,
                                                float c, float cb, float sum[5], float comp[5],
                                                float &A)
{
	const float fa = ivl32_fraction(Q, p);
	const float fb = ivl32_fraction(Q + kIvl32Doubles, p);
	float ya[5], yb[5];
	ivl32_terms(Q, fa * c, c, comp, ya);
	ivl32_terms(Q + kIvl32Doubles, fb * c, c, comp, yb);
#pragma unroll
	for (int k = 0; k < 5; k++)
		kahan_add_corrected(sum[k], comp[k], sel ? yb[k] : ya[k]);
	A += cb * (1.0f + 2.0f * fabsf(sel ? fb : fa));
}

// The decision on the mixed sums (see the bound above): returns the argmax
// the FP64 reference would return; macro[] receives the mixed-precision sums,
// or the FP64 sums when the lookup fell back (*fell_back = 1).  fp64(m)
// writes this lane's FP64 sums.  WAVE: fp64 is a wave-cooperative path that
// every lane of the wave must enter, so it runs for the whole wave as soon as
// one lane needs it (cheaper than one lane's FP64 loop with its wave idle).
template <int PREC, bool HISTORY, bool WAVE, class F64>
__device__ __forceinline__ int mixed_finish(const float sum[5], float A, double macro[5],
                                            uint32_t *fell_back, double *bound, F64 &&fp64)
{
	constexpr float kU = (PREC == XSB_XS_FP32) ? 0x1p-24f : 0x1p-8f;
	// argmax of the mixed sums (max = -1.0, strict '>', as the reference)
	double s[5];
#pragma unroll
	for (int k = 0; k < 5; k++)
		s[k] = (double)sum[k];
	int am = argmax5(s);
	const double B = 1.01 * (double)A * ((double)kU + 12.0 * 0x1p-24);
	if (bound)
		*bound = B;
	bool ambiguous = false;
#pragma unroll
	for (int k = 0; k < 5; k++) {
		if (k != am && !(s[am] - s[k] > 2.0 * B))
			ambiguous = true;
		if (HISTORY && fabs(s[k] - 1.0) <= B)
			ambiguous = true;
		// B <= kRelTol |S_c| certifies |S_c - R_c| <= kRelTol |S_c| / (1 - kRelTol)
		if (PREC == XSB_XS_FP32 && !(B <= kRelTol * (1.0 - kRelTol) * fabs(s[k])))
			ambiguous = true;
	}
	// the reference's argmax needs max > -1.0 for a non-zero index; every
	// sum is >= -B here only if the XS are non-negative, so treat a sum
	// within B of -1.0 as ambiguous too
	if (s[am] <= -1.0 + B)
		ambiguous = true;
#pragma unroll
	for (int k = 0; k < 5; k++)
		macro[k] = s[k];
	*fell_back = ambiguous ? 1u : 0u;
	if (WAVE ? __ballot(ambiguous) == 0ULL : !ambiguous)
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

template <int PREC>
__device__ __forceinline__ const float *mixed_consts(const Tables &T, int mat)
{
	const int plane = T.max_num_nucs * kMaterials;
	return T.concs32 + (PREC == XSB_XS_FP32 ? 0 : 2) * plane + mat * T.max_num_nucs;
}

// calculate_macro_xs on the mixed tables, one lane per lookup.
template <int GRID, int PREC, bool HISTORY = false>
__device__ __forceinline__ int macro_xs_mixed(const Tables &T, double p, int mat, int64_t idx,
                                              double macro[5], uint32_t *fell_back,
                                              double *bound = nullptr)
{
	constexpr int REC = (PREC == XSB_XS_FP32) ? kRec32Doubles : kRec16Doubles;
	const int nn = T.num_nucs[mat];
	const float *__restrict__ ceff = mixed_consts<PREC>(T, mat);
	const float *__restrict__ cbnd = ceff + T.max_num_nucs * kMaterials;
	const int32_t *__restrict__ ml = T.mats + mat * T.max_num_nucs;
	const uint8_t *__restrict__ mrow = (GRID == kUnionized) ? ueg_row(T, mat, idx) : nullptr;
	float sum[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
	float comp[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
	float A = 0.f;
	auto step = [&](int j) {
		const int nuc = ml[j];
		const double *row = mixed_row<PREC>(T, nuc);
		const int32_t lower = nuclide_lower<GRID, REC, HISTORY>(T, row, nuc, mrow, j, idx, p);
		if (PREC == XSB_XS_FP32 && GRID == kUnionized) {  // FP32 interval records
			mixed_term_ivl(T.ivl32 + ((int64_t)nuc * T.n_gp + lower) * kIvl32Doubles, p, ceff[j],
			               cbnd[j], sum, comp, A);
			return;
		}
		const double *lo = row + (int64_t)lower * REC;
		const double *hi = lo + REC;
		float xl[5], xh[5];
		load_xs<PREC>(lo, xl);
		load_xs<PREC>(hi, xh);
		mixed_term(lo[0], xl, hi[0], xh, p, ceff[j], cbnd[j], sum, comp, A);
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
	return mixed_finish<PREC, HISTORY, false>(sum, A, macro, fell_back, bound, [&](double m64[5]) {
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
                                                       float c, float cb, float sum[5],
                                                       float comp[5], float &A)
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
			mixed_term_ivl(R, p, c, cb, sum, comp, A);
		} else if (SHARE3 && (__ballot(du > 1u) == 0ULL || __ballot(du + 1u > 1u) == 0ULL)) {
			// two records shared; each lane takes its own
			const bool back = __ballot(du > 1u) != 0ULL;
			mixed_term_ivl2(back ? R - kIvl32Doubles : R, (back ? du + 1u : du) != 0u, p, c, cb,
			                sum, comp, A);
			__builtin_amdgcn_sched_barrier(0);
		} else {
			mixed_term_ivl(T.ivl32 + ((int64_t)nuc * n_gp + k) * kIvl32Doubles, p, c, cb, sum,
			               comp, A);
		}
		return;
	}
	const double *row = mixed_row<PREC>(T, nuc);
	const double *P = row + (int64_t)k0 * REC;
	float x0[5], x1[5];
	if (__ballot(d != 0) == 0ULL) {
		load_xs<PREC>(P, x0);
		load_xs<PREC>(P + REC, x1);
		mixed_term(P[0], x0, P[REC], x1, p, c, cb, sum, comp, A);
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
		mixed_term(sel ? e1 : e0, xl, sel ? e2 : e1, xh, p, c, cb, sum, comp, A);
		__builtin_amdgcn_sched_barrier(0);
	} else {
		const double *L = row + (int64_t)k * REC;
		load_xs<PREC>(L, x0);
		load_xs<PREC>(L + REC, x1);
		mixed_term(L[0], x0, L[REC], x1, p, c, cb, sum, comp, A);
	}
}

// calculate_macro_xs, unionized grid, mixed tables, for a wave whose lanes
// all hold material m; same terms and decision as macro_xs_mixed.
template <int PREC, int CHUNK, bool HISTORY, bool SHARE3 = true>
__device__ __forceinline__ int macro_ueg_wave_mixed(const Tables &T, double p, int m, int64_t idx,
                                                    double macro[5], uint32_t *fell_back,
                                                    double *bound)
{
	const float *__restrict__ ceff = mixed_consts<PREC>(T, m);
	const float *__restrict__ cbnd = ceff + T.max_num_nucs * kMaterials;
	float sum[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
	float comp[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
	float A = 0.f;
	ueg_wave_rows<CHUNK>(T, m, idx, [&](int j, int nuc, int k) {
		ueg_wave_nuclide_mixed<PREC, SHARE3, true>(T, nuc, k, p, ceff[j], cbnd[j], sum, comp, A);
	});
	return mixed_finish<PREC, HISTORY, true>(sum, A, macro, fell_back, bound, [&](double m64[5]) {
		macro_ueg_wave<CHUNK, SHARE3>(T, p, m, idx, m64);
	});
}

// calculate_macro_xs, hash grid, mixed tables, for a wave whose lanes share
// the material and the hash bin; same terms and decision as macro_xs_mixed.
template <int PREC, bool HISTORY, bool SHARE3 = true>
__device__ __forceinline__ int macro_hash_wave_mixed(const Tables &T, double p, int m, int64_t idx,
                                                     double macro[5], uint32_t *fell_back,
                                                     double *bound)
{
	constexpr int REC = (PREC == XSB_XS_FP32) ? kRec32Doubles : kRec16Doubles;
	const float *__restrict__ ceff = mixed_consts<PREC>(T, m);
	const float *__restrict__ cbnd = ceff + T.max_num_nucs * kMaterials;
	float sum[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
	float comp[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
	float A = 0.f;
	hash_wave_rows<REC>(T, mixed_row<PREC>(T, 0), p, m, idx, [&](int j, int nuc, int k) {
		ueg_wave_nuclide_mixed<PREC, SHARE3>(T, nuc, k, p, ceff[j], cbnd[j], sum, comp, A);
	});
	return mixed_finish<PREC, HISTORY, true>(sum, A, macro, fell_back, bound, [&](double m64[5]) {
		macro_hash_wave<SHARE3>(T, p, m, idx, m64);
	});
}

}  // namespace xsb
