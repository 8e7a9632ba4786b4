// xsb_sched.hip -- sorted lookup schedules (the reference's -k 2..6:
// openmp-threading/Simulation.c:698-871, cuda/Simulation.cu:514-1099).
//
// The samples are never materialised: a key kernel derives each lookup's
// (material, energy) from its LCG state and emits a sort key plus the lookup
// id; an LSD radix sort (10-bit digits, stable per-wave ranking with ballots)
// orders the ids; the lookup kernel re-derives (E, mat) from the id.  One
// launch covers all material segments, so no per-material tail waves.
//
//   k = 2, 4 : key = material                       (material segments)
//   k = 3, 5 : key = (material != fuel)             (fuel first)
//   k = 6    : key = material << 16 | energy bits   (material, then energy)
// Digits are 10 bits wide, so the material-energy key sorts in two passes.
#include "xsb_sorted.h"

namespace xsb {

constexpr int kEnergyBits = 16;
// Lookup kernel workgroup: eight waves of neighbouring sorted lookups share
// one CU's L1 (17M lookups: 256 threads 2.03 ms, 512 1.87, 768 1.99, 1024
// 1.91); runs of 4096 lookups per XCD.
constexpr int kIdsBlock = 512;
// Index-table entries staged per lane (16: 17M lookups FP64 1.89 -> 1.86 ms,
// FP32 1.58 -> 1.51 ms against 8; BF16 keeps 8: 3.74 -> 4.05 ms with 16).
constexpr int kIdsChunk = 16;
constexpr unsigned kXcdRun = 4096 / kIdsBlock;  // lookup blocks per XCD run

enum SortMode : int { kModeMat = 0, kModeFuel = 1, kModeMatE = 2 };

// Sort key of the lookup whose LCG state (before its two draws) is s.
__device__ __forceinline__ uint32_t sample_key(const Tables &T, uint64_t s, int mode)
{
	const double p = lcg_double(s);
	const int mat = pick_mat(s, T.thr);
	if (mode == kModeFuel)
		return mat != 0 ? 1u : 0u;
	if (mode == kModeMat)
		return (uint32_t)mat;
	uint32_t e = (uint32_t)(p * (double)(1u << kEnergyBits));  // exact scaling, truncation
	e = e < (1u << kEnergyBits) ? e : (1u << kEnergyBits) - 1u;
	return ((uint32_t)mat << kEnergyBits) | e;
}

// Keys + ids for lookups [first, first+n), and the first pass's per-tile
// digit histogram hist[digit * ntiles + tile].
// Output: {key, lookup id} pairs, interleaved (one 8-byte word per lookup:
// the sort moves each element with one load and one store).
__global__ __launch_bounds__(kSortBlock) void k_keys(Tables T, int64_t first, int64_t n,
                                                     uint64_t seed0, int mode, uint2 *pairs,
                                                     uint32_t *hist, int ntiles)
{
	__shared__ uint32_t h[kBins];
	for (int d = threadIdx.x; d < kBins; d += kSortBlock)
		h[d] = 0;
	__syncthreads();
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tile = sort_tile();
	const int64_t base = (int64_t)tile * kTile;
	// A wave's lanes hold consecutive lookups (tile_pos), two draws each:
	// the wave's first state s0 is one full skip-ahead, a lane's state is s0
	// advanced 2 * lane draws and the next item's s0 is 128 draws on, both
	// single radix-256 digits of the skip table (entry 0 is the identity).
	const int64_t e0 = base + tile_pos(wave, 0, 0);
	uint64_t s0 = ff_jump(seed0, 2ULL * (uint64_t)(first + e0), T.ff_tab);
	const ulonglong2 lane_skip = reinterpret_cast<const ulonglong2 *>(T.ff_tab)[2 * lane];
	const ulonglong2 item_skip = reinterpret_cast<const ulonglong2 *>(T.ff_tab)[128];
	static_assert(kSortItems > 1 && 2 * 63 < 256, "skips within one digit");
	for (int it = 0; it < kSortItems; it++) {
		const int64_t e = base + tile_pos(wave, it, lane);
		if (e < n) {
			const uint64_t s = (lane_skip.x * s0 + lane_skip.y) & kLcgMask;
			const uint32_t k = sample_key(T, s, mode);
			pairs[e] = make_uint2(k, (uint32_t)e);
			atomicAdd(&h[k & kDigitMask], 1u);
		}
		s0 = (item_skip.x * s0 + item_skip.y) & kLcgMask;
	}
	__syncthreads();
	write_tile_hist(h, hist, ntiles, tile);
}

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_hist(const K *__restrict__ keys, int64_t n,
                                                     int shift, uint32_t *hist, int ntiles)
{
	__shared__ uint32_t h[kBins];
	for (int d = threadIdx.x; d < kBins; d += kSortBlock)
		h[d] = 0;
	__syncthreads();
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tile = sort_tile();
	const int64_t base = (int64_t)tile * kTile;
	uint32_t dg[kSortItems];  // all loads in flight before the counting
#pragma unroll
	for (int it = 0; it < kSortItems; it++) {
		const int64_t e = base + tile_pos(wave, it, lane);
		dg[it] = e < n ? (uint32_t)(keys[e] >> shift) & kDigitMask : (uint32_t)kBins;
	}
#pragma unroll
	for (int it = 0; it < kSortItems; it++)
		if (dg[it] < (uint32_t)kBins)
			atomicAdd(&h[dg[it]], 1u);
	__syncthreads();
	write_tile_hist(h, hist, ntiles, tile);
}

// k_hist over interleaved {key, id} pairs.
__global__ __launch_bounds__(kSortBlock) void k_hist_pairs(const uint2 *__restrict__ pairs,
                                                           int64_t n, int shift, uint32_t *hist,
                                                           int ntiles)
{
	__shared__ uint32_t h[kBins];
	for (int d = threadIdx.x; d < kBins; d += kSortBlock)
		h[d] = 0;
	__syncthreads();
	// a histogram does not depend on the order: two pairs per 16-byte load
	const int tile = sort_tile();
	const int64_t base = (int64_t)tile * kTile;
	const uint4 *__restrict__ q = reinterpret_cast<const uint4 *>(pairs + base);
	uint32_t dg[kSortItems];
#pragma unroll
	for (int it = 0; it < kSortItems / 2; it++) {
		const int64_t e = base + 2 * (it * kSortBlock + threadIdx.x);  // pairs e, e + 1
		uint4 w = make_uint4(0u, 0u, 0u, 0u);
		if (e + 1 < n)
			w = q[it * kSortBlock + threadIdx.x];
		else if (e < n)
			w.x = pairs[e].x;
		dg[2 * it] = e < n ? (w.x >> shift) & kDigitMask : (uint32_t)kBins;
		dg[2 * it + 1] = e + 1 < n ? (w.z >> shift) & kDigitMask : (uint32_t)kBins;
	}
#pragma unroll
	for (int it = 0; it < kSortItems; it++)
		if (dg[it] < (uint32_t)kBins)
			atomicAdd(&h[dg[it]], 1u);
	__syncthreads();
	write_tile_hist(h, hist, ntiles, tile);
}

// ---- exclusive scan of the digit-major histogram (chunk totals, then apply) ----
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane)
{
#pragma unroll
	for (int off = 1; off < 64; off <<= 1) {
		const uint32_t y = __shfl_up(x, off, 64);
		if (lane >= off)
			x += y;
	}
	return x;
}

// exclusive scan of one value per thread over a BLOCK-thread block
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan_n(uint32_t x)
{
	constexpr int kW = BLOCK / 64;
	__shared__ uint32_t wsum[kW];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t incl = wave_incl_scan(x, lane);
	if (lane == 63)
		wsum[wave] = incl;
	__syncthreads();
	uint32_t off = 0;
#pragma unroll
	for (int w = 0; w < kW; w++)
		off += (w < wave) ? wsum[w] : 0u;
	__syncthreads();
	return off + incl - x;
}

// exclusive scan of one value per thread over a 256-thread block
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *total)
{
	__shared__ uint32_t wsum[kSortWaves];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t incl = wave_incl_scan(x, lane);
	if (lane == 63)
		wsum[wave] = incl;
	__syncthreads();
	uint32_t off = 0, tot = 0;
#pragma unroll
	for (int w = 0; w < kSortWaves; w++) {
		off += (w < wave) ? wsum[w] : 0u;
		tot += wsum[w];
	}
	__syncthreads();
	if (total)
		*total = tot;
	return off + incl - x;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t *__restrict__ a, int64_t n,
                                                     uint32_t *part)
{
	const int64_t base = (int64_t)blockIdx.x * kScanChunk;
	uint32_t s = 0;
	for (int it = 0; it < kScanItems; it++) {
		const int64_t i = base + (int64_t)it * 256 + threadIdx.x;
		if (i < n)
			s += a[i];
	}
	uint32_t tot;
	block_excl_scan(s, &tot);
	if (threadIdx.x == 0)
		part[blockIdx.x] = tot;
}

// Each block first adds up the chunk totals of the blocks before it (a few
// loads per thread: the totals are few and cached), so no separate pass
// scans them.  The histogram length is a multiple of kBins, so a thread's
// 16 words are either all in range or all out: read and written as four
// 16-byte vectors.
static_assert(kScanItems == 16 && kBins % kScanItems == 0, "whole vectors per thread");
__global__ __launch_bounds__(256) void k_scan_apply(uint32_t *a, int64_t n,
                                                    const uint32_t *__restrict__ part)
{
	uint32_t c = 0;
	for (int i = threadIdx.x; i < (int)blockIdx.x; i += 256)
		c += part[i];
	uint32_t carry;
	block_excl_scan(c, &carry);
	const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
	const bool in = base < n;
	uint4 w[kScanItems / 4];
#pragma unroll
	for (int q = 0; q < kScanItems / 4; q++)
		w[q] = in ? reinterpret_cast<const uint4 *>(a + base)[q] : make_uint4(0u, 0u, 0u, 0u);
	uint32_t s = 0;
#pragma unroll
	for (int q = 0; q < kScanItems / 4; q++)
		s += w[q].x + w[q].y + w[q].z + w[q].w;
	uint32_t off = block_excl_scan(s, nullptr) + carry;
	if (!in)
		return;
#pragma unroll
	for (int q = 0; q < kScanItems / 4; q++) {
		uint4 o;
		o.x = off;
		off += w[q].x;
		o.y = off;
		off += w[q].y;
		o.z = off;
		off += w[q].z;
		o.w = off;
		off += w[q].w;
		reinterpret_cast<uint4 *>(a + base)[q] = o;
	}
}

// Stable scatter of one digit.  Tile order is (wave, item, lane); each
// wave ranks its items in that order with ballot-built peer masks and
// per-wave digit counters in LDS, then waves are offset per digit.
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_scatter(const K *__restrict__ kin,
                                                        const uint32_t *__restrict__ vin,
                                                        K *__restrict__ kout,
                                                        uint32_t *__restrict__ vout, int64_t n,
                                                        int shift,
                                                        const uint32_t *__restrict__ hist,
                                                        int ntiles)
{
	__shared__ uint32_t cnt[kSortWaves][kBins];
	for (int w = 0; w < kSortWaves; w++)
		for (int d = threadIdx.x; d < kBins; d += kSortBlock)
			cnt[w][d] = 0;
	__syncthreads();
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint64_t lt_mask = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
	const int tile = sort_tile();
	const int64_t base = (int64_t)tile * kTile;
	K kk[kSortItems];
	uint32_t vv[kSortItems];
	uint16_t loc[kSortItems];
	for (int it = 0; it < kSortItems; it++) {
		const int64_t e = base + tile_pos(wave, it, lane);
		const bool valid = e < n;
		kk[it] = valid ? kin[e] : (K)0;
		vv[it] = valid ? vin[e] : 0u;
		const uint32_t d = (uint32_t)(kk[it] >> shift) & kDigitMask;
		uint64_t peers = __ballot(valid);
#pragma unroll
		for (int b = 0; b < kDigitBits; b++) {
			const uint64_t bal = __ballot((d >> b) & 1u);
			peers &= ((d >> b) & 1u) ? bal : ~bal;
		}
		if (valid) {
			const uint32_t c = cnt[wave][d];
			loc[it] = (uint16_t)(c + (uint32_t)__popcll(peers & lt_mask));
			if ((peers & lt_mask) == 0)  // lowest lane of the peer group
				cnt[wave][d] = c + (uint32_t)__popcll(peers);
		}
	}
	__syncthreads();
	for (int d = threadIdx.x; d < kBins; d += kSortBlock) {
		uint32_t run = hist[(int64_t)d * ntiles + tile];
		for (int w = 0; w < kSortWaves; w++) {
			const uint32_t t = cnt[w][d];
			cnt[w][d] = run;
			run += t;
		}
	}
	__syncthreads();
	for (int it = 0; it < kSortItems; it++) {
		const int64_t e = base + tile_pos(wave, it, lane);
		if (e < n) {
			const uint32_t d = (uint32_t)(kk[it] >> shift) & kDigitMask;
			const uint32_t pos = cnt[wave][d] + loc[it];
			kout[pos] = kk[it];
			vout[pos] = vv[it];
		}
	}
}

// Scatter of one digit of {key, value} pairs for the lookup schedules, which need the order only
// for locality: a tile's elements are grouped by digit in LDS (ranks from
// LDS counters, so equal digits of one tile may swap places), then written
// out in runs, consecutive lanes storing consecutive addresses.  After the
// last pass the keys are sorted; elements with equal keys may be permuted
// within a tile.
constexpr int kRunsBlock = 1024;                 // 16 waves: two blocks fill a CU
constexpr int kRunsItems = kTile / kRunsBlock;   // 8 elements per thread
static_assert(kBins == kRunsBlock, "one digit per thread in the local scan");

// IDS: the pass writes only the ids, packed (uint32_t[n] at pout), for a
// caller that needs the order and not the keys.  starts (if given) gets each
// digit's first position in the output.
template <bool IDS>
__global__ __launch_bounds__(kRunsBlock) void k_scatter_runs(const uint2 *__restrict__ pin,
                                                             uint2 *__restrict__ pout, int64_t n,
                                                             int shift,
                                                             const uint32_t *__restrict__ hist,
                                                             int ntiles, uint32_t *starts)
{
	__shared__ uint32_t cnt[kBins];   // digit counts, then the tile's local starts
	__shared__ uint32_t gbase[kBins]; // global start minus local start, per digit
	__shared__ uint32_t sk[kTile], sv[kTile];
	cnt[threadIdx.x] = 0;
	__syncthreads();
	const int tile = sort_tile();
	const int64_t base = (int64_t)tile * kTile;
	const int m = (int)min((int64_t)kTile, n - base);
	uint32_t kk[kRunsItems], vv[kRunsItems], rank[kRunsItems];
#pragma unroll
	for (int it = 0; it < kRunsItems; it++) {
		const int i = it * kRunsBlock + threadIdx.x;
		if (i < m) {
			const uint2 q = pin[base + i];
			kk[it] = q.x;
			vv[it] = q.y;
		}
	}
#pragma unroll
	for (int it = 0; it < kRunsItems; it++) {
		const int i = it * kRunsBlock + threadIdx.x;
		if (i < m)
			rank[it] = atomicAdd(&cnt[(kk[it] >> shift) & kDigitMask], 1u);
	}
	__syncthreads();
	const int d = threadIdx.x;
	const uint32_t c = cnt[d];
	const uint32_t run = block_excl_scan_n<kRunsBlock>(c);
	cnt[d] = run;
	const uint32_t start = hist[(int64_t)d * ntiles + tile];
	if (starts && tile == 0)
		starts[d] = start;
	gbase[d] = start - run;
	__syncthreads();
#pragma unroll
	for (int it = 0; it < kRunsItems; it++) {
		const int i = it * kRunsBlock + threadIdx.x;
		if (i < m) {
			const uint32_t pos = cnt[(kk[it] >> shift) & kDigitMask] + rank[it];
			sk[pos] = kk[it];
			sv[pos] = vv[it];
		}
	}
	__syncthreads();
#pragma unroll
	for (int it = 0; it < kRunsItems; it++) {
		const int i = it * kRunsBlock + threadIdx.x;
		if (i < m) {
			const uint32_t key = sk[i];
			const uint32_t g = gbase[(key >> shift) & kDigitMask] + (uint32_t)i;
			if (IDS)
				reinterpret_cast<uint32_t *>(pout)[g] = sv[i];
			else
				pout[g] = make_uint2(key, sv[i]);
		}
	}
}

// Lookups in sorted order: position i holds lookup id ids[i] (the sort's
// last pass writes the ids alone); (E, mat) are
// re-derived from the id's LCG state.  Lanes past n repeat lookup n-1 (in the
// same wave) and contribute nothing, so every lane stays active for the
// wave-wide search and the wave-uniform material path.
// Eight waves per SIMD (at most 64 VGPRs) hide the per-nuclide latency chain
// better than the six the register demand would otherwise give.
// On the unionized grid the wave path runs without its two-interval step:
// with the tables read through the scalar cache, the straddling waves'
// per-lane record loads cost less than interpolating two shared records on
// every lane.  The hash grid's wave path keeps it (its per-lane step also
// searches the bin per lane).
template <int GRID, int PREC, bool DETAIL>
__global__ __launch_bounds__(kIdsBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_lookup_ids(Tables T, int64_t n,
                                                    const uint32_t *__restrict__ ids,
                                                    int64_t first, uint64_t seed0, DetailOut det,
                                                    unsigned long long *vsum,
                                                    unsigned long long *nfb)
{
	const int64_t i = (int64_t)xcd_block(kXcdRun) * kIdsBlock + threadIdx.x;
	const bool valid = i < n;
	const uint32_t id = ids[valid ? i : n - 1];
	uint64_t s = ff_jump(seed0, 2ULL * (uint64_t)(first + id), T.ff_tab);
	const double p = lcg_double(s);
	const int mat = pick_mat(s, T.thr);
	double macro[5];
	double bnd = 0.0;
	uint32_t fb = 0;
	int64_t idx;
	const int am = sorted_lookup<GRID, PREC, PREC == XSB_XS_BF16 ? kIdxChunk : kIdsChunk, false,
	                             GRID != kUnionized>(
	    T, p, mat, T.variant & kVariantWave, idx, macro, fb, DETAIL ? &bnd : nullptr);
	if (!valid)
		fb = 0;
	const uint32_t v = valid ? (uint32_t)am + 1u : 0u;
	if (DETAIL && valid) {
		const int64_t o = id;
		if (det.energy) det.energy[o] = p;
		if (det.mat) det.mat[o] = mat;
		if (det.index) det.index[o] = idx;
		if (det.macro)
			for (int c = 0; c < 5; c++) det.macro[5 * o + c] = macro[c];
		if (det.argmax) det.argmax[o] = am;
		if (det.flag) det.flag[o] = (int32_t)fb;
		if (det.bound) det.bound[o] = bnd;
	}
	block_add_u64<kIdsBlock>(v, vsum);
	if (PREC != XSB_XS_FP64 && nfb)
		block_add_u64<kIdsBlock>(fb, nfb);
}

template <int PREC>
static void launch_ids_prec(Device *d, hipStream_t s, int64_t n, const uint32_t *ids, int64_t first,
                            uint64_t seed, const DetailOut *det, unsigned long long *vsum,
                            unsigned long long *nfb)
{
	const unsigned nb = (unsigned)((n + kIdsBlock - 1) / kIdsBlock);
	DetailOut dz{};
	const DetailOut &dd = det ? *det : dz;
#define XSB_LAUNCH(G)                                                                         \
	do {                                                                                      \
		if (det)                                                                              \
			hipLaunchKernelGGL((k_lookup_ids<G, PREC, true>), dim3(nb), dim3(kIdsBlock), 0, s, \
			                   d->T, n, ids, first, seed, dd, vsum, nfb);                     \
		else                                                                                  \
			hipLaunchKernelGGL((k_lookup_ids<G, PREC, false>), dim3(nb), dim3(kIdsBlock), 0, \
			                   s, d->T, n, ids, first, seed, dd, vsum, nfb);                  \
	} while (0)
	switch (d->T.grid_type) {
	case kUnionized: XSB_LAUNCH(kUnionized); break;
	case kNuclide: XSB_LAUNCH(kNuclide); break;
	default: XSB_LAUNCH(kHash); break;
	}
#undef XSB_LAUNCH
}

static int scan_hist(uint32_t *hist, int64_t len, uint32_t *part, hipStream_t s)
{
	const int nparts = (int)((len + kScanChunk - 1) / kScanChunk);
	hipLaunchKernelGGL(k_scan_reduce, dim3(nparts), dim3(256), 0, s, hist, len, part);
	hipLaunchKernelGGL(k_scan_apply, dim3(nparts), dim3(256), 0, s, hist, len, part);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

template <typename K>
static int sort_pairs_t(hipStream_t s, int64_t n, int passes, bool hist0_ready, K **k, uint32_t **v,
                        K **k2, uint32_t **v2, uint32_t *hist, uint32_t *part)
{
	const int ntiles = sort_tiles(n);
	const int64_t hlen = sort_hist_len(n);
	if (reinterpret_cast<uintptr_t>(hist) % 16 != 0)  // k_scan_apply's vectors
		return fail(XSB_ERR_INVALID, "sort histogram not 16-byte aligned");
	int rc;
	for (int pass = 0; pass < passes; pass++) {
		const int shift = kDigitBits * pass;
		if (pass > 0 || !hist0_ready) {
			hipLaunchKernelGGL(k_hist<K>, dim3(ntiles), dim3(kSortBlock), 0, s, *k, n, shift, hist,
			                   ntiles);
			XSB_HIP(hipGetLastError());
		}
		if ((rc = scan_hist(hist, hlen, part, s)))
			return rc;
		hipLaunchKernelGGL(k_scatter<K>, dim3(ntiles), dim3(kSortBlock), 0, s, *k, *v, *k2, *v2, n,
		                   shift, hist, ntiles);
		XSB_HIP(hipGetLastError());
		K *t = *k;
		*k = *k2;
		*k2 = t;
		uint32_t *tv = *v;
		*v = *v2;
		*v2 = tv;
	}
	return XSB_OK;
}

int sort_pairs(hipStream_t s, int64_t n, int passes, bool hist0_ready, uint32_t **k, uint32_t **v,
               uint32_t **k2, uint32_t **v2, uint32_t *hist, uint32_t *part)
{
	return sort_pairs_t<uint32_t>(s, n, passes, hist0_ready, k, v, k2, v2, hist, part);
}

int order_pairs(hipStream_t s, int64_t n, int passes, bool hist0_ready, uint2 **p,
                            uint2 **p2, uint32_t *hist, uint32_t *part, bool ids_last,
                            uint32_t *starts)
{
	const int ntiles = sort_tiles(n);
	const int64_t hlen = sort_hist_len(n);
	// k_scan_apply's and k_hist_pairs' 16-byte vectors
	if ((reinterpret_cast<uintptr_t>(hist) | reinterpret_cast<uintptr_t>(*p) |
	     reinterpret_cast<uintptr_t>(*p2)) % 16 != 0)
		return fail(XSB_ERR_INVALID, "sort buffers not 16-byte aligned");
	int rc;
	for (int pass = 0; pass < passes; pass++) {
		const int shift = kDigitBits * pass;
		if (pass > 0 || !hist0_ready) {
			hipLaunchKernelGGL(k_hist_pairs, dim3(ntiles), dim3(kSortBlock), 0, s, *p, n, shift,
			                   hist, ntiles);
			XSB_HIP(hipGetLastError());
		}
		if ((rc = scan_hist(hist, hlen, part, s)))
			return rc;
		const bool last = pass == passes - 1;
		if (ids_last && last)
			hipLaunchKernelGGL(k_scatter_runs<true>, dim3(ntiles), dim3(kRunsBlock), 0, s, *p, *p2,
			                   n, shift, hist, ntiles, starts);
		else
			hipLaunchKernelGGL(k_scatter_runs<false>, dim3(ntiles), dim3(kRunsBlock), 0, s, *p, *p2,
			                   n, shift, hist, ntiles, last ? starts : nullptr);
		XSB_HIP(hipGetLastError());
		uint2 *t = *p;
		*p = *p2;
		*p2 = t;
	}
	return XSB_OK;
}

int sort_pairs64(hipStream_t s, int64_t n, int passes, uint64_t **k, uint32_t **v, uint64_t **k2,
                 uint32_t **v2, uint32_t *hist, uint32_t *part)
{
	return sort_pairs_t<uint64_t>(s, n, passes, false, k, v, k2, v2, hist, part);
}

int run_sorted_event(Device *d, const xsb_event_args *a, const DetailOut *det, xsb_result *r,
                     unsigned long long *vsum, unsigned long long *nfb, void *work,
                     size_t work_bytes)
{
	const int64_t n = a->lookups;
	hipStream_t s = d->stream;
	const int k = r->kernel_id;
	const int mode = (k == XSB_K_MAT_E_SORT) ? kModeMatE
	                 : (k == XSB_K_FUEL_LAUNCH || k == XSB_K_FUEL_SORT) ? kModeFuel
	                                                                    : kModeMat;
	const int passes = (mode == kModeMatE) ? 2 : 1;
	if (n == 0) {
		XSB_HIP(hipEventRecord(d->ev[1], s));
		XSB_HIP(hipEventRecord(d->ev[2], s));
		return XSB_OK;
	}
	const int ntiles = sort_tiles(n);
	const int64_t hlen = sort_hist_len(n);
	const int nparts = sort_parts(n);
	const int64_t n2 = (n + 1) & ~(int64_t)1;  // pair buffers start 16-byte aligned
	const size_t bytes = sorted_event_bytes(n);
	void *buf = work;
	int rc;
	if (work) {
		if (work_bytes < bytes || reinterpret_cast<uintptr_t>(work) % 16 != 0)
			return fail(XSB_ERR_INVALID, "work buffer of %zu bytes for %lld lookups (%zu needed)",
			            work_bytes, (long long)n, bytes);
	} else if ((rc = scratch_reserve(d, 0, bytes, &buf))) {
		return rc;
	}
	uint2 *pA = (uint2 *)buf, *pB = pA + n2;
	uint32_t *hist = (uint32_t *)(pB + n2), *part = hist + hlen, *starts = part + nparts + 16;

	hipLaunchKernelGGL(k_keys, dim3(ntiles), dim3(kSortBlock), 0, s, d->T, a->first_lookup, n,
	                   a->seed, mode, pA, hist, ntiles);
	XSB_HIP(hipGetLastError());
	if ((rc = order_pairs(s, n, passes, true, &pA, &pB, hist, part, true, starts)))
		return rc;
	const uint32_t *vA = reinterpret_cast<const uint32_t *>(pA);  // the ids, packed
	XSB_HIP(hipEventRecord(d->ev[1], s));
	if (a->xs_precision == XSB_XS_FP64)
		launch_ids_prec<XSB_XS_FP64>(d, s, n, vA, a->first_lookup, a->seed, det, vsum, nfb);
	else if (a->xs_precision == XSB_XS_FP32)
		launch_ids_prec<XSB_XS_FP32>(d, s, n, vA, a->first_lookup, a->seed, det, vsum, nfb);
	else
		launch_ids_prec<XSB_XS_BF16>(d, s, n, vA, a->first_lookup, a->seed, det, vsum, nfb);
	XSB_HIP(hipGetLastError());
	XSB_HIP(hipEventRecord(d->ev[2], s));

	// the last pass's digit starts, for the material counts
	static_assert(kBins * sizeof(uint32_t) <= 4 * 1024, "fits the pinned result buffer");
	XSB_HIP(hipMemcpyAsync(d->h_out + 2, starts, kBins * sizeof(uint32_t), hipMemcpyDeviceToHost,
	                       s));
	return XSB_OK;
}

void finish_sorted_event(const Device *d, const xsb_event_args *a, xsb_result *r)
{
	static_assert(kEnergyBits + 4 <= 2 * kDigitBits, "material-energy key fits two digits");
	const int64_t n = a->lookups;
	if (n == 0)
		return;
	const int k = r->kernel_id;
	const int mode = (k == XSB_K_MAT_E_SORT) ? kModeMatE
	                 : (k == XSB_K_FUEL_LAUNCH || k == XSB_K_FUEL_SORT) ? kModeFuel
	                                                                    : kModeMat;
	const uint32_t *st = reinterpret_cast<const uint32_t *>(d->h_out + 2);
	if (mode != kModeFuel) {
		for (int dgt = 0; dgt < kBins; dgt++) {
			const uint32_t next = (dgt == kBins - 1) ? (uint32_t)n : st[dgt + 1];
			const int m = (mode == kModeMat) ? dgt : (dgt >> (kEnergyBits - kDigitBits));
			if (m < kMaterials)
				r->material_counts[m] += (int64_t)(next - st[dgt]);
		}
	} else {
		r->material_counts[0] = (int64_t)st[1];
	}
}

}  // namespace xsb
