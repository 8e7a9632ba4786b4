// xsb_lookup.hip -- lookup kernels of libxsb_hip.so (gfx950).
//
// One lookup per lane (wave64).  The FP64 kernels reproduce the reference's
// arithmetic exactly (see xsb_device.h); the verification sum of argmax+1 is
// reduced per block and added with one device-scope atomic, instead of the
// reference GPU port's lookups x 8 B array plus a separate reduction
// (cuda/Simulation.cu:34).
#include "xsb_internal.h"
#include "xsb_mixed.h"

namespace xsb {

constexpr int kBlock = 256;

static inline unsigned blocks_for(int64_t n, int per_block = kBlock)
{
	return (unsigned)((n + per_block - 1) / per_block);
}

// ---------------------------------------------------------------------------
// Schedule 0: fused sampling + lookup in lookup order
// (openmp-threading/Simulation.c:45-111, cuda/Simulation.cu:44-99)
// ---------------------------------------------------------------------------
template <int GRID, bool DETAIL>
__global__ __launch_bounds__(kBlock) void k_event_fused(Tables T, int64_t first, int64_t n,
                                                        uint64_t seed0, DetailOut det,
                                                        unsigned long long *vsum)
{
	const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
	uint32_t v = 0;
	if (i < n) {
		uint64_t s = ff_jump(seed0, 2ULL * (uint64_t)(first + i), T.ff_tab);
		const double p = lcg_double(s);
		const int mat = pick_mat(s, T.thr);
		const int64_t idx = lookup_index<GRID>(T, p);
		double macro[5];
		macro_xs_fp64<GRID>(T, p, mat, idx, macro);
		const int am = argmax5(macro);
		v = (uint32_t)am + 1u;
		if (DETAIL) {
			if (det.energy) det.energy[i] = p;
			if (det.mat) det.mat[i] = mat;
			if (det.index) det.index[i] = idx;
			if (det.macro)
				for (int c = 0; c < 5; c++) det.macro[5 * i + c] = macro[c];
			if (det.argmax) det.argmax[i] = am;
		}
	}
	block_add_u64<kBlock>(v, vsum);
}

int launch_event_fused(Device *d, int64_t first, int64_t n, uint64_t seed, const DetailOut *det,
                       unsigned long long *vsum, hipStream_t s)
{
	if (n <= 0)
		return XSB_OK;
	const unsigned nb = blocks_for(n);
	DetailOut dz{};
	const bool detail = det != nullptr;
	const DetailOut &dd = detail ? *det : dz;
#define XSB_LAUNCH(G)                                                                           \
	do {                                                                                        \
		if (detail)                                                                             \
			hipLaunchKernelGGL((k_event_fused<G, true>), dim3(nb), dim3(kBlock), 0, s, d->T,    \
			                   first, n, seed, dd, vsum);                                       \
		else                                                                                    \
			hipLaunchKernelGGL((k_event_fused<G, false>), dim3(nb), dim3(kBlock), 0, s, d->T,   \
			                   first, n, seed, dd, vsum);                                       \
	} while (0)
	switch (d->T.grid_type) {
	case kUnionized: XSB_LAUNCH(kUnionized); break;
	case kNuclide: XSB_LAUNCH(kNuclide); break;
	default: XSB_LAUNCH(kHash); break;
	}
#undef XSB_LAUNCH
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

// ---------------------------------------------------------------------------
// Sampling kernel (cuda/Simulation.cu:441-462): energies and materials of
// lookups [first, first+n) into arrays.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_sample(Tables T, int64_t first, int64_t n,
                                                   uint64_t seed0, double *E, int32_t *mat)
{
	const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
	if (i >= n)
		return;
	uint64_t s = ff_jump(seed0, 2ULL * (uint64_t)(first + i), T.ff_tab);
	E[i] = lcg_double(s);
	mat[i] = pick_mat(s, T.thr);
}

int launch_sample(Device *d, int64_t first, int64_t n, uint64_t seed, double *E, int32_t *mat,
                  hipStream_t s)
{
	if (n <= 0)
		return XSB_OK;
	hipLaunchKernelGGL(k_sample, dim3(blocks_for(n)), dim3(kBlock), 0, s, d->T, first, n, seed, E,
	                   mat);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

// ---------------------------------------------------------------------------
// Lookups over an explicit sample list (cuda/Simulation.cu:464-509).  With
// perm != nullptr, list position i holds original lookup perm[i] and the
// optional per-lookup outputs are written at that original position.
// ---------------------------------------------------------------------------
template <int GRID, int PREC, bool DETAIL>
__global__ __launch_bounds__(kBlock) void k_lookup_list(Tables T, int64_t n, const double *E,
                                                        const int32_t *M, const uint32_t *perm,
                                                        DetailOut det, unsigned long long *vsum,
                                                        unsigned long long *nfb)
{
	const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
	uint32_t v = 0, fb = 0;
	// a material outside [0, 12) is not looked up: it adds nothing and its
	// argmax output is -1
	if (i < n && (unsigned)M[i] >= (unsigned)kMaterials) {
		if (DETAIL && det.argmax)
			det.argmax[perm ? (int64_t)perm[i] : i] = -1;
	} else if (i < n) {
		const double p = E[i];
		const int mat = M[i];
		const int64_t idx = lookup_index<GRID>(T, p);
		double macro[5];
		double bnd = 0.0;
		int am;
		if (PREC == XSB_XS_FP64) {
			macro_xs_fp64<GRID>(T, p, mat, idx, macro);
			am = argmax5(macro);
		} else {
			am = macro_xs_mixed<GRID, PREC>(T, p, mat, idx, macro, &fb, DETAIL ? &bnd : nullptr);
		}
		v = (uint32_t)am + 1u;
		if (DETAIL) {
			const int64_t o = perm ? (int64_t)perm[i] : i;
			if (det.energy) det.energy[o] = p;
			if (det.mat) det.mat[o] = mat;
			if (det.index) det.index[o] = idx;
			if (det.macro)
				for (int c = 0; c < 5; c++) det.macro[5 * o + c] = macro[c];
			if (det.argmax) det.argmax[o] = am;
			if (det.flag) det.flag[o] = (int32_t)fb;
			if (det.bound) det.bound[o] = bnd;
		}
	}
	block_add_u64<kBlock>(v, vsum);
	if (PREC != XSB_XS_FP64 && nfb)
		block_add_u64<kBlock>(fb, nfb);
}

template <int PREC>
static void launch_list_prec(Device *d, unsigned nb, hipStream_t s, int64_t n, const double *E,
                             const int32_t *M, const uint32_t *perm, const DetailOut &dd,
                             bool detail, unsigned long long *vsum, unsigned long long *nfb)
{
#define XSB_LAUNCH(G)                                                                           \
	do {                                                                                        \
		if (detail)                                                                             \
			hipLaunchKernelGGL((k_lookup_list<G, PREC, true>), dim3(nb), dim3(kBlock), 0, s,    \
			                   d->T, n, E, M, perm, dd, vsum, nfb);                             \
		else                                                                                    \
			hipLaunchKernelGGL((k_lookup_list<G, PREC, false>), dim3(nb), dim3(kBlock), 0, s,   \
			                   d->T, n, E, M, perm, dd, vsum, nfb);                             \
	} while (0)
	switch (d->T.grid_type) {
	case kUnionized: XSB_LAUNCH(kUnionized); break;
	case kNuclide: XSB_LAUNCH(kNuclide); break;
	default: XSB_LAUNCH(kHash); break;
	}
#undef XSB_LAUNCH
}

int launch_lookup_list(Device *d, int64_t n, const double *E, const int32_t *mat,
                       const uint32_t *perm, int precision, const DetailOut *det,
                       unsigned long long *vsum, unsigned long long *nfallback, hipStream_t s)
{
	if (n <= 0)
		return XSB_OK;
	if (precision == XSB_XS_FP32 && !d->d_xs32)
		return fail(XSB_ERR_UNSUPPORTED, "FP32 tables were not built at upload (XSB_BUILD_FP32)");
	if (precision == XSB_XS_BF16 && !d->d_xs16)
		return fail(XSB_ERR_UNSUPPORTED, "BF16 tables were not built at upload (XSB_BUILD_BF16)");
	const unsigned nb = blocks_for(n);
	DetailOut dz{};
	const bool detail = det != nullptr;
	const DetailOut &dd = detail ? *det : dz;
	if (precision == XSB_XS_FP64)
		launch_list_prec<XSB_XS_FP64>(d, nb, s, n, E, mat, perm, dd, detail, vsum, nfallback);
	else if (precision == XSB_XS_FP32)
		launch_list_prec<XSB_XS_FP32>(d, nb, s, n, E, mat, perm, dd, detail, vsum, nfallback);
	else
		launch_list_prec<XSB_XS_BF16>(d, nb, s, n, E, mat, perm, dd, detail, vsum, nfallback);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

// ---------------------------------------------------------------------------
// calculate_micro_xs (openmp-threading/Simulation.c:241-323) over explicit
// samples: energy, nuclide, and the grid-level data the reference reads from
// index_data at the sample's grid index -- the bracket on the unionized grid,
// (u_low, u_high) on the hash grid; the nuclide grid bisects the whole row.
// Bounds are checked on the host before the launch.
// ---------------------------------------------------------------------------
template <int GRID>
__global__ __launch_bounds__(kBlock) void k_micro(Tables T, int64_t n, const double *__restrict__ E,
                                                  const int32_t *__restrict__ nuc,
                                                  const int32_t *__restrict__ lo,
                                                  const int32_t *__restrict__ hi,
                                                  double *__restrict__ xs)
{
	const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
	if (i >= n)
		return;
	const double p = E[i];
	const int32_t n_gp = T.n_gp;
	const double *row = T.grid + (int64_t)nuc[i] * n_gp * kPointDoubles;
	int32_t lower;
	if (GRID == kNuclide) {
		lower = grid_search_nuclide<kPointDoubles>(p, row, 0, n_gp - 1);
	} else if (GRID == kUnionized) {
		lower = lo[i];
	} else {
		const int32_t u_low = lo[i], u_high = hi[i];
		if (p <= row[(int64_t)u_low * kPointDoubles])
			lower = 0;
		else if (p >= row[(int64_t)u_high * kPointDoubles])
			lower = n_gp - 1;
		else
			lower = grid_search_nuclide<kPointDoubles>(p, row, u_low, u_high);
	}
	if (lower == n_gp - 1)
		lower--;
	const double *L = row + (int64_t)lower * kPointDoubles, *H = L + kPointDoubles;
	const double f = (H[0] - p) / (H[0] - L[0]);
#pragma unroll
	for (int c = 0; c < 5; c++)
		xs[5 * i + c] = H[1 + c] - f * (H[1 + c] - L[1 + c]);
}

int launch_micro(Device *d, int64_t n, const double *E, const int32_t *nuc, const int32_t *lo,
                 const int32_t *hi, double *xs, hipStream_t s)
{
	if (n <= 0)
		return XSB_OK;
	const unsigned nb = blocks_for(n);
	switch (d->T.grid_type) {
	case kUnionized:
		hipLaunchKernelGGL(k_micro<kUnionized>, dim3(nb), dim3(kBlock), 0, s, d->T, n, E, nuc, lo,
		                   hi, xs);
		break;
	case kNuclide:
		hipLaunchKernelGGL(k_micro<kNuclide>, dim3(nb), dim3(kBlock), 0, s, d->T, n, E, nuc, lo,
		                   hi, xs);
		break;
	default:
		hipLaunchKernelGGL(k_micro<kHash>, dim3(nb), dim3(kBlock), 0, s, d->T, n, E, nuc, lo, hi,
		                   xs);
		break;
	}
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

// ---------------------------------------------------------------------------
// History-based transport, one particle per lane
// (openmp-threading/Simulation.c:141-237).
// ---------------------------------------------------------------------------

template <int GRID, int PREC>
__global__ __launch_bounds__(kBlock) void k_history(Tables T, int64_t particles, int lookups,
                                                    uint64_t seed0, uint64_t *out_sum,
                                                    uint64_t *out_seed, unsigned long long *vsum,
                                                    unsigned long long *nfb)
{
	const int64_t pid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
	uint32_t v = 0, fb = 0;
	if (pid < particles) {
		uint64_t s = ff_jump(seed0, history_offset(pid, lookups), T.ff_tab);
		double p = lcg_double(s);
		int mat = pick_mat(s, T.thr);
		for (int it = 0; it < lookups; it++) {
			const int64_t idx = lookup_index<GRID>(T, p);
			double macro[5];
			int am;
			if (PREC == XSB_XS_FP64) {
				macro_xs_fp64<GRID>(T, p, mat, idx, macro);
				am = argmax5(macro);
			} else {
				uint32_t f1 = 0;
				am = macro_xs_mixed<GRID, PREC, true>(T, p, mat, idx, macro, &f1);
				fb += f1;
			}
			v += (uint32_t)am + 1u;
			uint32_t n_forward = 0;
#pragma unroll
			for (int c = 0; c < 5; c++)
				n_forward += (macro[c] > 1.0) ? 1u : 0u;
			if (n_forward > 0)
				s = ff_lcg(s, n_forward);
			p = lcg_double(s);
			mat = pick_mat(s, T.thr);
		}
		if (out_sum) out_sum[pid] = v;
		if (out_seed) out_seed[pid] = s;
	}
	block_add_u64<kBlock>(v, vsum);
	if (PREC != XSB_XS_FP64 && nfb)
		block_add_u64<kBlock>(fb, nfb);
}

int launch_history(Device *d, int64_t particles, int lookups, uint64_t seed, int precision,
                   uint64_t *out_sum, uint64_t *out_seed, unsigned long long *vsum,
                   unsigned long long *nfallback, hipStream_t s)
{
	if (particles <= 0)
		return XSB_OK;
	if (precision == XSB_XS_FP32 && !d->d_xs32)
		return fail(XSB_ERR_UNSUPPORTED, "FP32 tables were not built at upload (XSB_BUILD_FP32)");
	if (precision == XSB_XS_BF16 && !d->d_xs16)
		return fail(XSB_ERR_UNSUPPORTED, "BF16 tables were not built at upload (XSB_BUILD_BF16)");
	const unsigned nb = blocks_for(particles);
#define XSB_LAUNCH(G, P)                                                                       \
	hipLaunchKernelGGL((k_history<G, P>), dim3(nb), dim3(kBlock), 0, s, d->T, particles,       \
	                   lookups, seed, out_sum, out_seed, vsum, nallback_)
	unsigned long long *nallback_ = nfallback;
	const int g = d->T.grid_type;
	if (precision == XSB_XS_FP64) {
		if (g == kUnionized) XSB_LAUNCH(kUnionized, XSB_XS_FP64);
		else if (g == kNuclide) XSB_LAUNCH(kNuclide, XSB_XS_FP64);
		else XSB_LAUNCH(kHash, XSB_XS_FP64);
	} else if (precision == XSB_XS_FP32) {
		if (g == kUnionized) XSB_LAUNCH(kUnionized, XSB_XS_FP32);
		else if (g == kNuclide) XSB_LAUNCH(kNuclide, XSB_XS_FP32);
		else XSB_LAUNCH(kHash, XSB_XS_FP32);
	} else {
		if (g == kUnionized) XSB_LAUNCH(kUnionized, XSB_XS_BF16);
		else if (g == kNuclide) XSB_LAUNCH(kNuclide, XSB_XS_BF16);
		else XSB_LAUNCH(kHash, XSB_XS_BF16);
	}
#undef XSB_LAUNCH
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

// ---------------------------------------------------------------------------
// Algorithmic-byte counter: per lookup, the stored size of the distinct words
// the reference algorithm reads (search probes, num_nucs, material list and
// concentration words, index words, bracket points and any per-nuclide
// bisection probes).  Word sizes follow the table format of `precision`.
// ---------------------------------------------------------------------------
struct Fmt {
	int e, xs, idx, conc, mat;
};

__device__ __forceinline__ int probes_between(const double *__restrict__ E, int stride, int32_t low,
                                              int32_t high, double q)
{
	int n = 0;
	int32_t lo = low, hi = high;
	while (hi - lo > 1) {
		int32_t mid = lo + ((hi - lo) >> 1);
		n++;
		if (E[(int64_t)mid * stride] > q)
			hi = mid;
		else
			lo = mid;
	}
	return n;
}

template <int GRID>
__global__ __launch_bounds__(kBlock) void k_balg(Tables T, int64_t first, int64_t n, uint64_t seed0,
                                                 Fmt f, unsigned long long *out)
{
	const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
	uint32_t b = 0;
	if (i < n) {
		uint64_t s = ff_jump(seed0, 2ULL * (uint64_t)(first + i), T.ff_tab);
		const double p = lcg_double(s);
		const int mat = pick_mat(s, T.thr);
		const int nn = T.num_nucs[mat];
		b = 4u + (uint32_t)nn * (uint32_t)(f.conc + f.mat);
		int64_t idx = -1;
		if (GRID == kUnionized) {
			b += 8u * (uint32_t)probes_between(T.ueg, 1, 0, (int32_t)T.n_union - 1, p);
			idx = grid_search((int32_t)T.n_union, p, T.ueg);
		} else if (GRID == kHash) {
			idx = lookup_index<kHash>(T, p);
		}
		const int32_t n_gp = T.n_gp;
		for (int j = 0; j < nn; j++) {
			const int nuc = T.mats[mat * T.max_num_nucs + j];
			const double *Er = T.grid + (int64_t)nuc * n_gp * kPointDoubles;
			if (GRID == kUnionized) {
				b += (uint32_t)(f.idx + 2 * (f.e + f.xs));
			} else if (GRID == kNuclide) {
				int32_t lower = grid_search_nuclide<kPointDoubles>(p, Er, 0, n_gp - 1);
				int pr = probes_between(Er, kPointDoubles, 0, n_gp - 1, p);
				int seen = (lower > 0) + (lower + 1 < n_gp - 1);
				b += (uint32_t)(pr * f.e + 2 * (f.e + f.xs) - seen * f.e);
			} else {
				const int32_t u_low = T.index_grid[idx * T.idx_pitch + nuc];
				int32_t u_high;
				b += (uint32_t)f.idx;
				if (idx == T.hash_bins - 1) {
					u_high = n_gp - 1;
				} else {
					u_high = T.index_grid[(idx + 1) * T.idx_pitch + nuc] + 1;
					b += (uint32_t)f.idx;
				}
				const double e_low = Er[(int64_t)u_low * kPointDoubles];
				const double e_high = Er[(int64_t)u_high * kPointDoubles];
				int32_t lower;
				int pr = 0;
				if (p <= e_low) {
					lower = 0;
				} else if (p >= e_high) {
					lower = n_gp - 1;
				} else {
					lower = grid_search_nuclide<kPointDoubles>(p, Er, u_low, u_high);
					pr = probes_between(Er, kPointDoubles, u_low, u_high, p);
				}
				if (lower == n_gp - 1)
					lower--;
				int eb = 2 * f.e + pr * f.e;
				const bool lo_seen = lower == u_low || lower == u_high ||
				                     (lower > u_low && lower < u_high && pr > 0);
				const bool hi_seen = lower + 1 == u_low || lower + 1 == u_high ||
				                     (lower + 1 > u_low && lower + 1 < u_high && pr > 0);
				eb += (lo_seen ? 0 : f.e) + (hi_seen ? 0 : f.e);
				b += (uint32_t)(eb + 2 * f.xs);
			}
		}
	}
	// per-lookup bytes can exceed 2^16; sum in 64 bits per wave
	unsigned long long w = b;
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		w += __shfl_xor(w, off, 64);
	if ((threadIdx.x & 63) == 0 && w)
		atomicAdd(out, w);
}

int launch_balg(Device *d, int64_t first, int64_t n, uint64_t seed, int precision, double *out_mean,
                hipStream_t s)
{
	// word sizes: energy, xs block per point, index word, conc, mats
	Fmt f{8, 40, 4, 8, 4};
	if (precision == XSB_XS_FP32)
		f = Fmt{8, 20, 4, 4, 4};
	else if (precision == XSB_XS_BF16)
		f = Fmt{8, 10, 4, 4, 4};
	unsigned long long *acc = d->d_acc + 32;
	XSB_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), s));
	const unsigned nb = blocks_for(n);
	switch (d->T.grid_type) {
	case kUnionized:
		hipLaunchKernelGGL(k_balg<kUnionized>, dim3(nb), dim3(kBlock), 0, s, d->T, first, n, seed,
		                   f, acc);
		break;
	case kNuclide:
		hipLaunchKernelGGL(k_balg<kNuclide>, dim3(nb), dim3(kBlock), 0, s, d->T, first, n, seed, f,
		                   acc);
		break;
	default:
		hipLaunchKernelGGL(k_balg<kHash>, dim3(nb), dim3(kBlock), 0, s, d->T, first, n, seed, f,
		                   acc);
		break;
	}
	XSB_HIP(hipGetLastError());
	unsigned long long tot = 0;
	XSB_HIP(hipMemcpyAsync(&tot, acc, sizeof(tot), hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	*out_mean = n > 0 ? (double)tot / (double)n : 0.0;
	return XSB_OK;
}

// ---------------------------------------------------------------------------
// Plain copy, used to calibrate the HBM byte counters.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_copy(const float4 *__restrict__ src,
                                                 float4 *__restrict__ dst, int64_t n4)
{
	for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n4;
	     i += (int64_t)gridDim.x * kBlock)
		dst[i] = src[i];
}

int launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t s)
{
	const int64_t n4 = bytes / 16;
	unsigned nb = blocks_for(n4);
	if (nb > 256u * 64u)
		nb = 256u * 64u;
	hipLaunchKernelGGL(k_copy, dim3(nb), dim3(kBlock), 0, s, (const float4 *)src, (float4 *)dst,
	                   n4);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

}  // namespace xsb
