// xsb_batch.hip -- many independent lookup batches (no reference
// counterpart): each batch has its own sample stream and material CDF.
//
// Batch b, lookup i:  s = FF(seed_b, 2i), E = LCG(s), roll = LCG(s),
//                     mat = first m with roll < thr_b[m]  (thr_b[11] >= 1).
// Global lookup id g = off_b + i numbers all lookups batch after batch.
//
// mode 0 (sorted pipeline, replayed as one hipGraph):
//   k_breset    clears the per-batch sums, counters and queue heads (a fill
//               kernel: no memset node in the graph)
//   k_bkeys     key = window << (4 + eb) | mat << eb | top eb bits of E, and
//               the lookup id, for every lookup of every batch
//   radix sort  (window, material, energy) order, ids as payload
//   k_bsorted   persistent lookup kernel: sort window after sort window,
//               workgroups take tiles of the window's part of the sorted
//               order from 8 per-XCD queues (the XCD's own slice first, so
//               its L2 sees neighbouring energies, then the other queues);
//               per-batch sums and counts are gathered in LDS and flushed
//               with one atomic per batch (after every tile when there are
//               several windows); a batch's completion time is taken when
//               its count reaches its size.
// Sort windows: by default the smallest batches (at most 55% of them and 20%
// of the lookups) form a first window and the rest a second, so that most
// batches complete after a small part of the work; or windows of about
// `window` lookups in batch order; or one window.
// mode 1: one fused launch per batch, in batch order.
// mode 2: one persistent kernel over a queue of tiles in lookup order; each
//         tile is ordered by material inside the workgroup.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "xsb_sorted.h"

namespace xsb {

constexpr int kBatchBlock = 256;
// The sorted pipeline's persistent kernel (512 threads: FP64 9.65 -> 10.7 ms
// for 122M lookups, 1024: 12.2 ms; the event kernel gains from 512).
constexpr int kBSortedBlock = 256;

__device__ __forceinline__ void batch_sample(const Tables &T, uint64_t seed, const double *thr,
                                             int64_t i, double &p, int &mat)
{
	uint64_t s = ff_jump(seed, 2ULL * (uint64_t)i, T.ff_tab);
	p = lcg_double(s);
	const double roll = lcg_double(s);
	int m = 0;
#pragma unroll
	for (int k = 0; k < kMaterials - 1; k++)
		m += (thr[k] <= roll) ? 1 : 0;  // thresholds increase; thr[11] >= 1 > roll
	mat = m;
}

template <int GRID, int PREC>
__device__ __forceinline__ int lookup_one(const Tables &T, double p, int mat, uint32_t *fb)
{
	const int64_t idx = lookup_index<GRID>(T, p);
	double macro[5];
	if (PREC == XSB_XS_FP64) {
		macro_xs_fp64<GRID>(T, p, mat, idx, macro);
		return argmax5(macro);
	}
	return macro_xs_mixed<GRID, PREC>(T, p, mat, idx, macro, fb);
}

// ---- mode 1: one launch per batch ----
template <int GRID, int PREC>
__global__ __launch_bounds__(kBatchBlock) void k_batch_one(Tables T, int64_t size, uint64_t seed,
                                                           const double *__restrict__ thr_g,
                                                           unsigned long long *vsum,
                                                           unsigned long long *nfb)
{
	double thr[kMaterials];
#pragma unroll
	for (int k = 0; k < kMaterials; k++)
		thr[k] = thr_g[k];
	const int64_t i = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
	uint32_t v = 0, fb = 0;
	if (i < size) {
		double p;
		int mat;
		batch_sample(T, seed, thr, i, p, mat);
		v = (uint32_t)lookup_one<GRID, PREC>(T, p, mat, &fb) + 1u;
	}
	block_add_u64<kBatchBlock>(v, vsum);
	if (PREC != XSB_XS_FP64)
		block_add_u64<kBatchBlock>(fb, nfb);
}

// ---- mode 2: persistent kernel over a queue of tiles in lookup order ----
struct BatchQueue {
	const int64_t *sizes;       // [nb]
	const uint64_t *seeds;      // [nb]
	const double *thr;          // [nb][12]
	const int64_t *item_off;    // [nb + 1] first work item of each batch
	unsigned long long *vsum;   // [nb]
	unsigned long long *nfb;    // [1]
	uint32_t *remaining;        // [nb] tiles left per batch
	uint64_t *done_clock;       // [nb] device clock when the batch completed
	uint32_t *next_item;        // queue head
	int64_t n_items;
	int32_t nb;
	int32_t tile;               // lookups per work item (multiple of kBatchBlock)
};

__global__ void k_queue_reset(BatchQueue q, const uint32_t *__restrict__ tiles_per_batch)
{
	for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < q.nb; b += gridDim.x * blockDim.x) {
		q.remaining[b] = tiles_per_batch[b];
		q.vsum[b] = 0;
		q.done_clock[b] = 0;
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		*q.next_item = 0;
		*q.nfb = 0;
	}
}

template <int GRID, int PREC>
__global__ __launch_bounds__(kBatchBlock) void k_batch_persistent(Tables T, BatchQueue q)
{
	__shared__ int64_t s_item;
	__shared__ uint32_t s_cnt[kMaterials + 1];
	__shared__ uint32_t s_slot[kBatchBlock * 16];  // ordered tile (lookup offsets)
	__shared__ uint8_t s_mat[kBatchBlock * 16];
	for (;;) {
		if (threadIdx.x == 0)
			s_item = (int64_t)atomicAdd(q.next_item, 1u);
		__syncthreads();
		const int64_t w = s_item;
		if (w >= q.n_items)
			break;  // every workgroup leaves through here once the queue is drained
		// batch of work item w: last b with item_off[b] <= w
		int lo = 0, hi = q.nb;
		while (hi - lo > 1) {
			const int mid = (lo + hi) >> 1;
			if (q.item_off[mid] <= w)
				lo = mid;
			else
				hi = mid;
		}
		const int b = lo;
		const int64_t t0 = (w - q.item_off[b]) * (int64_t)q.tile;
		const int64_t size = q.sizes[b];
		const int n = (int)min((int64_t)q.tile, size - t0);
		const uint64_t seed = q.seeds[b];
		double thr[kMaterials];
#pragma unroll
		for (int k = 0; k < kMaterials; k++)
			thr[k] = q.thr[b * kMaterials + k];
		// order the tile by material: counting sort in LDS
		if (threadIdx.x <= kMaterials)
			s_cnt[threadIdx.x] = 0;
		__syncthreads();
		for (int j = threadIdx.x; j < n; j += kBatchBlock) {
			double p;
			int mat;
			batch_sample(T, seed, thr, t0 + j, p, mat);
			s_mat[j] = (uint8_t)mat;
			atomicAdd(&s_cnt[mat + 1], 1u);
		}
		__syncthreads();
		if (threadIdx.x == 0)
			for (int m = 1; m <= kMaterials; m++)
				s_cnt[m] += s_cnt[m - 1];
		__syncthreads();
		for (int j = threadIdx.x; j < n; j += kBatchBlock) {
			const uint32_t pos = atomicAdd(&s_cnt[s_mat[j]], 1u);
			s_slot[pos] = (uint32_t)j;
		}
		__syncthreads();
		uint32_t v = 0, fb = 0;
		for (int j = threadIdx.x; j < n; j += kBatchBlock) {
			double p;
			int mat;
			batch_sample(T, seed, thr, t0 + s_slot[j], p, mat);
			uint32_t f1 = 0;
			v += (uint32_t)lookup_one<GRID, PREC>(T, p, mat, &f1) + 1u;
			fb += f1;
		}
		block_add_u64<kBatchBlock>(v, q.vsum + b);
		if (PREC != XSB_XS_FP64)
			block_add_u64<kBatchBlock>(fb, q.nfb);
		if (threadIdx.x == 0) {
			__threadfence();
			if (atomicSub(q.remaining + b, 1u) == 1u)
				q.done_clock[b] = wall_clock64();
		}
		__syncthreads();
	}
}

__global__ void k_clock(uint64_t *out)
{
	*out = wall_clock64();
}

template <int PREC>
static void launch_persistent(Device *d, const BatchQueue &q, unsigned grid, hipStream_t s)
{
	switch (d->T.grid_type) {
	case kUnionized:
		hipLaunchKernelGGL((k_batch_persistent<kUnionized, PREC>), dim3(grid), dim3(kBatchBlock), 0,
		                   s, d->T, q);
		break;
	case kNuclide:
		hipLaunchKernelGGL((k_batch_persistent<kNuclide, PREC>), dim3(grid), dim3(kBatchBlock), 0, s,
		                   d->T, q);
		break;
	default:
		hipLaunchKernelGGL((k_batch_persistent<kHash, PREC>), dim3(grid), dim3(kBatchBlock), 0, s,
		                   d->T, q);
		break;
	}
}

template <int PREC>
static void launch_one(Device *d, int64_t size, uint64_t seed, const double *thr,
                       unsigned long long *vsum, unsigned long long *nfb, hipStream_t s)
{
	if (size <= 0)
		return;
	const unsigned nb = (unsigned)((size + kBatchBlock - 1) / kBatchBlock);
	switch (d->T.grid_type) {
	case kUnionized:
		hipLaunchKernelGGL((k_batch_one<kUnionized, PREC>), dim3(nb), dim3(kBatchBlock), 0, s, d->T,
		                   size, seed, thr, vsum, nfb);
		break;
	case kNuclide:
		hipLaunchKernelGGL((k_batch_one<kNuclide, PREC>), dim3(nb), dim3(kBatchBlock), 0, s, d->T,
		                   size, seed, thr, vsum, nfb);
		break;
	default:
		hipLaunchKernelGGL((k_batch_one<kHash, PREC>), dim3(nb), dim3(kBatchBlock), 0, s, d->T,
		                   size, seed, thr, vsum, nfb);
		break;
	}
}


// ---- mode 0: sorted pipeline ----
constexpr int kLdsBatches = 1024;  // per-batch LDS sums when n_batches <= this
constexpr int kXcds = 8;
// Batch tables on the device.  off[b] = first global id of batch b,
// off[nb] = total; win[b] = sort window of batch b.
struct BatchView {
	const int64_t *sizes;
	const uint64_t *seeds;
	const double *thr;
	const uint32_t *off;
	const uint32_t *win;
	int32_t nb;
};

// last b with off[b] <= g (off[0] = 0 <= g < off[nb]): the batch holding g,
// skipping empty batches
__device__ __forceinline__ int batch_of(const uint32_t *off, int nb, uint32_t g)
{
	int lo = 0, hi = nb;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (off[mid] <= g)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

struct SortedQueue {
	const uint2 *pairs;        // sorted {key, global lookup id}: material in key
	                           // bits [ebits, ebits + 4)
	int64_t n;
	const uint32_t *wpos;      // [windows + 1]: window w holds sorted positions
	                           // [wpos[w], wpos[w + 1])
	int32_t windows;
	int32_t ebits;
	int32_t tile;
	int32_t wave;              // wave-uniform lookup path for material-uniform waves
	int32_t w_begin, w_end;    // the windows this launch works through
	uint32_t *heads;           // [windows x 8] items taken from each (window, XCD) slice
	unsigned long long *vsum;  // [nb]
	unsigned long long *cnt;   // [nb] lookups completed
	uint64_t *done_clock;      // [nb]
	unsigned long long *nfb;   // [1]
};

__global__ void k_breset(SortedQueue q, int nb, int windows)
{
	for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
		q.vsum[b] = 0;
		q.cnt[b] = 0;
		q.done_clock[b] = 0;
	}
	for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < windows * kXcds;
	     h += gridDim.x * blockDim.x)
		q.heads[h] = 0;
	if (blockIdx.x == 0 && threadIdx.x == 0)
		*q.nfb = 0;
}

__global__ __launch_bounds__(kSortBlock) void k_bkeys(Tables T, BatchView B, int64_t n, int ebits,
                                                      uint2 *pairs, uint32_t *hist, int ntiles)
{
	__shared__ uint32_t h[kBins];
	for (int d = threadIdx.x; d < kBins; d += kSortBlock)
		h[d] = 0;
	__syncthreads();
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tile = sort_tile();
	const int64_t base = (int64_t)tile * kTile;
	const uint32_t emax = (1u << ebits) - 1u;
	for (int it = 0; it < kSortItems; it++) {
		const int64_t e = base + tile_pos(wave, it, lane);
		if (e < n) {
			const uint32_t g = (uint32_t)e;
			const int b = batch_of(B.off, B.nb, g);
			double thr[kMaterials];
#pragma unroll
			for (int k = 0; k < kMaterials; k++)
				thr[k] = B.thr[b * kMaterials + k];
			double p;
			int mat;
			batch_sample(T, B.seeds[b], thr, (int64_t)(g - B.off[b]), p, mat);
			uint32_t eb = (uint32_t)(p * (double)(1u << ebits));  // exact scaling, truncation
			eb = eb < emax ? eb : emax;
			const uint32_t key = (B.win[b] << (4 + ebits)) | ((uint32_t)mat << ebits) | eb;
			pairs[e] = make_uint2(key, g);
			atomicAdd(&h[key & kDigitMask], 1u);
		}
	}
	__syncthreads();
	write_tile_hist(h, hist, ntiles, tile);
}

// Adds the block's LDS sums of batches [0, nb) to the device totals and
// stamps the completion clock of every batch whose count becomes complete.
__device__ __forceinline__ void flush_batches(const BatchView &B, const SortedQueue &q,
                                              uint32_t *s_v, uint32_t *s_c)
{
	for (int b = threadIdx.x; b < B.nb; b += kBSortedBlock) {
		const uint32_t c = s_c[b];
		if (c) {
			atomicAdd(q.vsum + b, (unsigned long long)s_v[b]);
			const unsigned long long before = atomicAdd(q.cnt + b, (unsigned long long)c);
			if (before + c == (unsigned long long)B.sizes[b])
				q.done_clock[b] = wall_clock64();
			s_v[b] = 0;
			s_c[b] = 0;
		}
	}
}

// At least four waves per SIMD (at most 128 VGPRs): left free, the register
// allocator lands on either side of 128 from one variant to the next, and
// three waves lose about a quarter of the throughput.  DENSE: eight waves
// (64 VGPRs, the per-lane path spilling), for windows dense enough that the
// wave-uniform path dominates (first_dense_window).
template <int GRID, int PREC, int CHUNK, bool SHARE3, bool DENSE>
__global__ __launch_bounds__(kBSortedBlock) __attribute__((amdgpu_waves_per_eu(DENSE ? 8 : 4))) void k_bsorted(
    Tables T, BatchView B, SortedQueue q)
{
	__shared__ uint32_t s_off[kLdsBatches + 1];
	__shared__ uint32_t s_v[kLdsBatches], s_c[kLdsBatches];
	__shared__ int s_item;
	const int nb = B.nb;
	const bool lds = nb <= kLdsBatches;
	if (lds) {
		for (int b = threadIdx.x; b <= nb; b += kBSortedBlock)
			s_off[b] = B.off[b];
		for (int b = threadIdx.x; b < nb; b += kBSortedBlock)
			s_v[b] = s_c[b] = 0;
	}
	__syncthreads();
	const uint32_t *off = lds ? s_off : B.off;
	uint32_t fb_total = 0, since_flush = 0;
	// the dispatcher places workgroup w on XCD w % 8
	const int home = (int)(blockIdx.x % kXcds);
	// window after window; each window's tiles are dealt in 8 contiguous
	// slices, one per XCD's queue; a workgroup drains its own XCD's queue,
	// then the others', then moves on to the next window (every workgroup
	// reaches the end: each queue hands out its items once).  The LDS sums
	// are flushed when the workgroup leaves a window: a batch's lookups
	// spread over its whole window in sorted order, so its last ones are in
	// the window's last items, after which their workgroups move on (a flush
	// after every item stamped the same completions and cost 5% in atomics).
	for (int w = q.w_begin, xq = home, tries = 0; w < q.w_end;) {
		if (tries == kXcds) {
			if (lds && w + 1 < q.w_end) {
				__syncthreads();
				flush_batches(B, q, s_v, s_c);
				since_flush = 0;
			}
			w++;
			xq = home;
			tries = 0;
			continue;
		}
		// the queue position is read once per workgroup through LDS and kept
		// wave-uniform (scalar registers) from there on
		const int64_t lo = q.wpos[w], hi = q.wpos[w + 1];
		const int64_t n_items = (hi - lo + q.tile - 1) / q.tile;
		if (threadIdx.x == 0)
			s_item = (int)atomicAdd(q.heads + w * kXcds + xq, 1u);
		__syncthreads();
		const int64_t item = n_items * xq / kXcds + __builtin_amdgcn_readfirstlane(s_item);
		const int64_t end = n_items * (xq + 1) / kXcds;
		__syncthreads();  // every thread has read s_item before it is rewritten
		if (item >= end) {  // this queue is drained: move on to the next one
			xq = (xq + 1) % kXcds;
			tries++;
			continue;
		}
		const int64_t p0 = lo + item * q.tile, p1 = min(p0 + (int64_t)q.tile, hi);
		// all threads run the same trip count; lanes past p1 repeat p1 - 1 and
		// contribute nothing, so waves stay whole for the wave-wide paths
		for (int64_t r = p0; r < p1; r += kBSortedBlock) {
			const int64_t pos = r + threadIdx.x;
			const bool valid = pos < p1;
			const int64_t pp = valid ? pos : p1 - 1;
			const uint2 kv = q.pairs[pp];
			const uint32_t key = kv.x, g = kv.y;
			const int mat = (int)((key >> q.ebits) & 15u);
			const int b = batch_of(off, nb, g);
			uint64_t s = ff_jump(B.seeds[b], 2ULL * (uint64_t)(g - off[b]), T.ff_tab);
			const double p = lcg_double(s);
			int64_t idx;
			double macro[5];
			uint32_t fb = 0;
			const int am = sorted_lookup<GRID, PREC, CHUNK, false, SHARE3>(T, p, mat, q.wave, idx,
			                                                               macro, fb, nullptr);
			if (valid) {
				fb_total += fb;
				if (lds) {
					atomicAdd(&s_v[b], (uint32_t)am + 1u);
					atomicAdd(&s_c[b], 1u);
				} else {
					atomicAdd(q.vsum + b, (unsigned long long)(am + 1));
					if (atomicAdd(q.cnt + b, 1ULL) + 1ULL == (unsigned long long)B.sizes[b])
						q.done_clock[b] = wall_clock64();
				}
			}
		}
		// LDS sums stay below 2^32: at most 5 per lookup, flushed every 2^26 lookups
		since_flush += (uint32_t)(p1 - p0);
		if (lds && since_flush >= (1u << 26)) {
			__syncthreads();
			flush_batches(B, q, s_v, s_c);
			since_flush = 0;
		}
	}
	if (lds) {
		__syncthreads();
		flush_batches(B, q, s_v, s_c);
	}
	if (PREC != XSB_XS_FP64)
		block_add_u64<kBSortedBlock>(fb_total, q.nfb);
}

template <int GRID, int PREC, int CHUNK, bool SHARE3, bool DENSE>
static int launch_or_size(Device *d, const BatchView &B, const SortedQueue &q, bool occupancy_only,
                          int *occ, hipStream_t s)
{
	if (occupancy_only) {
		XSB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
		    occ, k_bsorted<GRID, PREC, CHUNK, SHARE3, DENSE>, kBSortedBlock, 0));
		return XSB_OK;
	}
	hipLaunchKernelGGL((k_bsorted<GRID, PREC, CHUNK, SHARE3, DENSE>), dim3((unsigned)*occ),
	                   dim3(kBSortedBlock), 0, s, d->T, B, q);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

template <int GRID, int PREC, int CHUNK, bool SHARE3>
static int launch_or_size_dense(Device *d, const BatchView &B, const SortedQueue &q,
                                bool dense, bool occupancy_only, int *occ, hipStream_t s)
{
	if (GRID == kUnionized && dense)
		return launch_or_size<GRID, PREC, CHUNK, SHARE3, true>(d, B, q, occupancy_only, occ, s);
	return launch_or_size<GRID, PREC, CHUNK, SHARE3, false>(d, B, q, occupancy_only, occ, s);
}

// dispatch on (grid type, precision, staging depth, two-interval step, dense
// windows); the staging depth and the dense variant only matter on the
// unionized grid
template <int GRID, int PREC>
static int dispatch_bsorted(Device *d, const BatchView &B, const SortedQueue &q, int stage,
                            bool share3, bool dense, bool occupancy_only, int *occ, hipStream_t s)
{
	if (GRID != kUnionized || stage >= 16)
		return share3 ? launch_or_size_dense<GRID, PREC, 16, true>(d, B, q, dense,
		                                                           occupancy_only, occ, s)
		              : launch_or_size_dense<GRID, PREC, 16, false>(d, B, q, dense,
		                                                            occupancy_only, occ, s);
	return share3 ? launch_or_size_dense<GRID, PREC, 8, true>(d, B, q, dense, occupancy_only,
	                                                          occ, s)
	              : launch_or_size_dense<GRID, PREC, 8, false>(d, B, q, dense, occupancy_only,
	                                                           occ, s);
}

template <int GRID>
static int dispatch_bsorted_prec(Device *d, int prec, const BatchView &B, const SortedQueue &q,
                                 int stage, bool share3, bool dense, bool occupancy_only,
                                 int *occ, hipStream_t s)
{
	if (prec == XSB_XS_FP64)
		return dispatch_bsorted<GRID, XSB_XS_FP64>(d, B, q, stage, share3, dense, occupancy_only,
		                                            occ, s);
	if (prec == XSB_XS_FP32)
		return dispatch_bsorted<GRID, XSB_XS_FP32>(d, B, q, stage, share3, dense, occupancy_only,
		                                            occ, s);
	return dispatch_bsorted<GRID, XSB_XS_BF16>(d, B, q, stage, share3, dense, occupancy_only, occ,
	                                            s);
}

// occupancy_only: *occ receives resident workgroups per CU; else *occ is
// the grid size to launch.  dense: the eight-wave variant (unionized grid).
static int run_bsorted(Device *d, int prec, const BatchView &B, const SortedQueue &q, int stage,
                       bool share3, bool dense, bool occupancy_only, int *occ, hipStream_t s)
{
	switch (d->T.grid_type) {
	case kUnionized:
		return dispatch_bsorted_prec<kUnionized>(d, prec, B, q, stage, share3, dense,
		                                         occupancy_only, occ, s);
	case kNuclide:
		return dispatch_bsorted_prec<kNuclide>(d, prec, B, q, stage, share3, false,
		                                       occupancy_only, occ, s);
	default:
		return dispatch_bsorted_prec<kHash>(d, prec, B, q, stage, share3, false, occupancy_only,
		                                    occ, s);
	}
}

// ---- host side of mode 0 ----
// mapping: 1 one lane per lookup (the reference loop per lane), 2 the
// wave-uniform path where a wave's lanes share the material (and, on the
// hash grid, the bin), 3 the same without the two-interval shared step
struct BatchPlan {
	int tile = 0, stage = 0, sort = 0, mapping = 0;
	bool fits(const BatchPlan &fixed) const
	{
		return (!fixed.tile || tile == fixed.tile) && (!fixed.stage || stage == fixed.stage) &&
		       (!fixed.sort || sort == fixed.sort) && (!fixed.mapping || mapping == fixed.mapping);
	}
	bool complete() const { return tile && stage && sort && mapping; }
};

// What the tuner's choice is keyed on: the table set, the precision and a
// coarse histogram of the batch sizes (count, mean and largest, in powers
// of two), plus the sort windows.
struct TuneKey {
	int64_t n_iso;
	int grid, prec, nb_log, mean_log, max_log, windows;
	bool operator==(const TuneKey &o) const
	{
		return n_iso == o.n_iso && grid == o.grid && prec == o.prec && nb_log == o.nb_log &&
		       mean_log == o.mean_log && max_log == o.max_log && windows == o.windows;
	}
};

// Everything one enqueue of the pipeline needs; its device pointers and
// sizes are also the key of the captured graph.
struct PipeShape {
	int64_t n = 0;
	int nb = 0, windows = 1, prec = 0, grid_type = 0;
	BatchPlan plan;
	int ebits = 0, passes = 1, grid = 0, grid_dense = 0;
	int dense_from = 1;  // windows [dense_from, windows) run the eight-wave variant
	void *tables = nullptr, *sortbuf = nullptr;
	bool operator==(const PipeShape &o) const
	{
		return n == o.n && nb == o.nb && windows == o.windows && prec == o.prec &&
		       grid_type == o.grid_type && plan.tile == o.plan.tile &&
		       plan.stage == o.plan.stage && plan.sort == o.plan.sort &&
		       plan.mapping == o.plan.mapping && ebits == o.ebits &&
		       passes == o.passes && grid == o.grid && grid_dense == o.grid_dense &&
		       dense_from == o.dense_from && tables == o.tables && sortbuf == o.sortbuf;
	}
};

struct BatchState {
	std::vector<std::pair<TuneKey, BatchPlan>> tuned;
	hipGraphExec_t exec = nullptr;
	PipeShape shape;
};

void batch_state_free(Device *d)
{
	if (!d->batch)
		return;
	if (d->batch->exec)
		(void)hipGraphExecDestroy(d->batch->exec);
	delete d->batch;
	d->batch = nullptr;
}

static int ilog2_ceil(uint64_t x)
{
	int l = 0;
	while (l < 63 && ((uint64_t)1 << l) < x)
		l++;
	return l;
}

// Device layout of the batch tables in scratch slot 1.
// wpos: the windows' first sorted positions (at most one window per
// batch); wpos1: the single window of the tuner's runs; heads: one queue
// head per (window, XCD).
struct TableLayout {
	size_t sizes, seeds, thr, off, win, wpos, wpos1, vsum, cnt, done, heads, nfb, clk, bytes;
	explicit TableLayout(int nb)
	{
		auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
		const size_t nw = (size_t)std::max(nb, 1);
		sizes = 0;
		seeds = up(sizes + 8 * (size_t)nb);
		thr = up(seeds + 8 * (size_t)nb);
		off = up(thr + 8 * (size_t)kMaterials * nb);
		win = up(off + 4 * ((size_t)nb + 1));
		wpos = up(win + 4 * (size_t)nb);
		wpos1 = up(wpos + 4 * (nw + 1));
		vsum = up(wpos1 + 4 * 2);
		cnt = up(vsum + 8 * (size_t)nb);
		done = up(cnt + 8 * (size_t)nb);
		heads = up(done + 8 * (size_t)nb);
		nfb = up(heads + 4 * kXcds * nw);
		clk = up(nfb + 8);
		bytes = up(clk + 8 * 4);
	}
};

// Enqueues the whole mode-0 pipeline on s (kernels only: capturable).
static int enqueue_pipeline(Device *d, const PipeShape &ps, const BatchView &B, const SortedQueue &q0,
                            uint64_t *clk, hipStream_t s)
{
	const int64_t n2 = (ps.n + 1) & ~(int64_t)1;  // pair buffers start 16-byte aligned
	uint2 *pA = (uint2 *)ps.sortbuf, *pB = pA + n2;
	uint32_t *hist = (uint32_t *)(pB + n2), *part = hist + sort_hist_len(ps.n);
	const int ntiles = sort_tiles(ps.n);
	hipLaunchKernelGGL(k_breset, dim3((std::max(ps.nb, ps.windows * kXcds) + 255) / 256), dim3(256),
	                   0, s, q0, ps.nb, ps.windows);
	hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, clk);
	hipLaunchKernelGGL(k_bkeys, dim3(ntiles), dim3(kSortBlock), 0, s, d->T, B, ps.n, ps.ebits, pA,
	                   hist, ntiles);
	XSB_HIP(hipGetLastError());
	int rc = order_pairs(s, ps.n, ps.passes, true, &pA, &pB, hist, part);
	if (rc)
		return rc;
	hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, clk + 1);
	SortedQueue q = q0;
	q.pairs = pA;
	q.n = ps.n;
	q.windows = ps.windows;
	q.tile = ps.plan.tile;
	q.ebits = ps.ebits;
	q.wave = ps.plan.mapping >= 2;
	// the sparse windows first (four waves per SIMD), then the dense ones
	// (eight): each launch works through its own range of windows
	if (ps.dense_from > 0) {
		q.w_begin = 0;
		q.w_end = ps.dense_from;
		int grid = ps.grid;
		if ((rc = run_bsorted(d, ps.prec, B, q, ps.plan.stage, ps.plan.mapping == 2, false, false,
		                      &grid, s)))
			return rc;
	}
	if (ps.dense_from < ps.windows) {
		q.w_begin = ps.dense_from;
		q.w_end = ps.windows;
		int grid = ps.grid_dense;
		if ((rc = run_bsorted(d, ps.prec, B, q, ps.plan.stage, ps.plan.mapping == 2, true, false,
		                      &grid, s)))
			return rc;
	}
	hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, clk + 2);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

// Windows with at least 16 lookups per unionized energy are dense: their
// waves mostly take the wave-uniform path, and eight waves per SIMD hide
// its latency better (122M lookups in one window: 9.16 -> 8.17 ms); sparser
// windows, whose straddling waves take the per-lane path, lose by it (that
// path spills at 64 VGPRs: the small batches' window of the default plan
// took 2.1 -> 3.4 ms).  The eight-wave variant runs only when every window
// is dense: a launch of each variant for its own windows (sparse first) was
// slower than one four-wave launch for all (12.97 against 12.04 ms).
// Returns 0 (every window runs eight waves) or windows (none does).
static int first_dense_window(const Device *d, const std::vector<uint32_t> &wpos, int windows)
{
	if (d->T.grid_type != kUnionized)
		return windows;
	for (int w = 0; w < windows; w++)
		if ((int64_t)(wpos[w + 1] - wpos[w]) < 16 * (int64_t)d->T.n_union)
			return windows;
	return 0;
}

// Shape of one pipeline run over the first nb batches (total n lookups);
// windows [dense_from, windows) run the eight-wave variant.
static int make_shape(Device *d, int nb, int64_t n, int windows, int dense_from, int prec,
                      const BatchPlan &plan, void *tables, PipeShape *ps)
{
	ps->n = n;
	ps->nb = nb;
	ps->windows = windows;
	ps->dense_from = dense_from;
	ps->prec = prec;
	ps->grid_type = d->T.grid_type;
	ps->plan = plan;
	const int wbits = windows > 1 ? ilog2_ceil((uint64_t)windows) : 0;
	if (wbits > 26)
		return fail(XSB_ERR_INVALID, "%d sort windows is more than 2^26", windows);
	// two radix passes (20 key bits) while the energy keeps 14 bits or more;
	// beyond that a third pass, for 16 energy bits
	ps->ebits = plan.sort != 2 ? 0 : (wbits <= 2 ? 16 - wbits : std::min(16, 26 - wbits));
	const int bits = wbits + 4 + ps->ebits;
	ps->passes = (bits + kDigitBits - 1) / kDigitBits;
	for (int dense = 0; dense < 2; dense++) {
		int occ = 0;
		int rc = run_bsorted(d, prec, BatchView{}, SortedQueue{}, plan.stage, plan.mapping == 2,
		                     dense != 0, true, &occ, nullptr);
		if (rc)
			return rc;
		int &grid = dense ? ps->grid_dense : ps->grid;
		grid = std::max(1, d->cu_count) * std::max(1, occ);
		grid = (grid + kXcds - 1) / kXcds * kXcds;
	}
	ps->tables = tables;
	const size_t sort_bytes =
	    (size_t)(n + 1) * 16 + (size_t)sort_hist_len(n) * 4 + (size_t)sort_parts(n) * 4 + 256;
	return scratch_reserve(d, 2, sort_bytes, &ps->sortbuf);
}

static void make_views(const PipeShape &ps, char *base, const TableLayout &L, BatchView *B,
                       SortedQueue *q, bool tuning = false)
{
	B->sizes = (const int64_t *)(base + L.sizes);
	B->seeds = (const uint64_t *)(base + L.seeds);
	B->thr = (const double *)(base + L.thr);
	B->off = (const uint32_t *)(base + L.off);
	B->win = (const uint32_t *)(base + L.win);
	B->nb = ps.nb;
	memset(q, 0, sizeof(*q));
	q->wpos = (const uint32_t *)(base + (tuning ? L.wpos1 : L.wpos));
	q->heads = (uint32_t *)(base + L.heads);
	q->vsum = (unsigned long long *)(base + L.vsum);
	q->cnt = (unsigned long long *)(base + L.cnt);
	q->done_clock = (uint64_t *)(base + L.done);
	q->nfb = (unsigned long long *)(base + L.nfb);
}

// Runs the pipeline once, eagerly, over the first nb batches; returns the
// GPU time in *ms.
static int time_pipeline(Device *d, int nb, int64_t n, int prec, const BatchPlan &plan, char *base,
                         const TableLayout &L, float *ms)
{
	PipeShape ps;
	const std::vector<uint32_t> w1{0u, (uint32_t)n};
	int rc = make_shape(d, nb, n, 1, first_dense_window(d, w1, 1), prec, plan, base, &ps);
	if (rc)
		return rc;
	BatchView B;
	SortedQueue q;
	make_views(ps, base, L, &B, &q, true);
	hipStream_t s = d->stream;
	XSB_HIP(hipEventRecord(d->ev[4], s));
	if ((rc = enqueue_pipeline(d, ps, B, q, (uint64_t *)(base + L.clk), s)))
		return rc;
	XSB_HIP(hipEventRecord(d->ev[5], s));
	XSB_HIP(hipEventSynchronize(d->ev[5]));
	XSB_HIP(hipEventElapsedTime(ms, d->ev[4], d->ev[5]));
	return XSB_OK;
}

// Picks the plan fields left open in `fixed` by timing each candidate (twice,
// keeping the faster) on a prefix of the batches of about 16 M lookups in one
// sort window (wpos1 = {0, its lookups}); a prefix of a random set of
// batches keeps its size histogram.  The staging depth only matters on the
// unionized grid.
static int autotune(Device *d, const xsb_batched_args *a, const std::vector<uint32_t> &off,
                    char *base, const TableLayout &L, const BatchPlan &fixed, BatchPlan *plan)
{
	const int nb = a->n_batches;
	int nbt = nb;
	for (int b = 0; b < nb; b++)
		if (off[b + 1] >= (16u << 20)) {
			nbt = b + 1;
			break;
		}
	const int64_t nt = off[nbt];
	const uint32_t w1[2] = {0u, (uint32_t)nt};
	XSB_HIP(hipMemcpy(base + L.wpos1, w1, sizeof(w1), hipMemcpyHostToDevice));
	auto pick = [](int v, std::vector<int> all) { return v ? std::vector<int>{v} : all; };
	const bool staged = d->T.grid_type == kUnionized;
	const std::vector<int> tiles = pick(fixed.tile, {2048, 8192});
	const std::vector<int> sorts = pick(fixed.sort, {2, 1});
	const std::vector<int> maps = pick(fixed.mapping, {2, 3, 1});
	float best = 1e30f;
	BatchPlan best_plan;
	for (int so : sorts)
		for (int ti : tiles)
			for (int mp : maps) {
				const std::vector<int> stages =
				    (staged && mp >= 2) ? pick(fixed.stage, {16, 8}) : pick(fixed.stage, {16});
				for (int st : stages) {
					const BatchPlan c{ti, st, so, mp};
					float ms = 0, ms2 = 0;
					int rc = time_pipeline(d, nbt, nt, a->xs_precision, c, base, L, &ms);
					if (!rc)
						rc = time_pipeline(d, nbt, nt, a->xs_precision, c, base, L, &ms2);
					if (rc)
						return rc;
					ms = std::min(ms, ms2);
					if (ms < best) {
						best = ms;
						best_plan = c;
					}
				}
			}
	*plan = best_plan;
	return XSB_OK;
}

static int run_sorted_batches(Device *d, const xsb_batched_args *a, xsb_result *r)
{
	const int nb = a->n_batches;
	std::vector<uint32_t> off(nb + 1, 0), win(nb, 0);
	int64_t total = 0, max_size = 0;
	for (int b = 0; b < nb; b++) {
		off[b] = (uint32_t)total;
		total += a->sizes[b];
		max_size = std::max(max_size, a->sizes[b]);
	}
	if (total >= ((int64_t)1 << 32))
		return fail(XSB_ERR_INVALID, "mode 0 handles fewer than 2^32 lookups per call (%lld)",
		            (long long)total);
	off[nb] = (uint32_t)total;
	int windows = 1;
	if (a->window > 0) {
		// windows of about a->window lookups, in batch order
		int64_t last = -1;
		int w = -1;
		for (int b = 0; b < nb; b++) {
			const int64_t k = (int64_t)off[b] / a->window;
			if (k != last) {
				w++;
				last = k;
			}
			win[b] = (uint32_t)w;
		}
		windows = w + 1;
	} else if (a->window == 0 && nb > 1) {
		// the smallest batches first: at most 55% of them, holding at most 20%
		// of the lookups, in a window of their own
		std::vector<int> order(nb);
		for (int b = 0; b < nb; b++)
			order[b] = b;
		std::stable_sort(order.begin(), order.end(),
		                 [&](int x, int y) { return a->sizes[x] < a->sizes[y]; });
		int64_t small = 0;
		int taken = 0;
		for (int i = 0; i < nb && 100 * (i + 1) <= 55 * nb; i++) {
			if (5 * (small + a->sizes[order[i]]) > total)
				break;
			small += a->sizes[order[i]];
			taken++;
		}
		if (taken > 0 && taken < nb) {
			for (int b = 0; b < nb; b++)
				win[b] = 1;
			for (int i = 0; i < taken; i++)
				win[order[i]] = 0;
			windows = 2;
		}
	}
	// window w holds sorted positions [wpos[w], wpos[w + 1])
	std::vector<uint32_t> wpos(windows + 1, 0);
	for (int b = 0; b < nb; b++)
		wpos[win[b] + 1] += (uint32_t)a->sizes[b];
	for (int w = 0; w < windows; w++)
		wpos[w + 1] += wpos[w];
	hipStream_t s = d->stream;
	if (!d->batch)
		d->batch = new BatchState();
	BatchState &st = *d->batch;
	const TableLayout L(nb);
	void *mem = nullptr;
	int rc = scratch_reserve(d, 1, L.bytes, &mem);
	if (rc)
		return rc;
	char *base = (char *)mem;
	// the batch tables: plain copies before the pipeline, so that the graph
	// reads them from fixed addresses and replays for any batch contents
	XSB_HIP(hipMemcpyAsync(base + L.sizes, a->sizes, 8 * (size_t)nb, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + L.seeds, a->seeds, 8 * (size_t)nb, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + L.thr, a->thresholds, 8 * (size_t)kMaterials * nb,
	                       hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + L.off, off.data(), 4 * ((size_t)nb + 1), hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + L.win, win.data(), 4 * (size_t)nb, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + L.wpos, wpos.data(), 4 * ((size_t)windows + 1),
	                       hipMemcpyHostToDevice, s));
	XSB_HIP(hipStreamSynchronize(s));  // host vectors go out of scope with pending copies otherwise

	// plan: the caller's fields, the rest tuned (cached per key)
	BatchPlan fixed;
	fixed.tile = a->tile > 0 ? std::max(kBatchBlock, (a->tile + kBatchBlock - 1) / kBatchBlock * kBatchBlock) : 0;
	fixed.stage = a->stage > 0 ? (a->stage >= 16 ? 16 : 8) : 0;
	fixed.sort = a->sort > 0 ? (a->sort >= 2 ? 2 : 1) : 0;
	fixed.mapping = a->mapping > 0 ? std::min(a->mapping, 3) : 0;
	BatchPlan plan = fixed;
	int tuned = 0;
	double ms_tuning = 0;
	if (total > 0 && !fixed.complete()) {
		const TuneKey key{d->T.n_iso, d->T.grid_type, a->xs_precision, ilog2_ceil((uint64_t)nb),
		                  ilog2_ceil((uint64_t)std::max<int64_t>(1, total / nb)),
		                  ilog2_ceil((uint64_t)std::max<int64_t>(1, max_size)), windows};
		bool found = false;
		for (auto &e : st.tuned)
			if (e.first == key && e.second.fits(fixed)) {
				plan = e.second;
				found = true;
				break;
			}
		if (!found) {
			const auto t0 = std::chrono::steady_clock::now();
			if ((rc = autotune(d, a, off, base, L, fixed, &plan)))
				return rc;
			ms_tuning = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			tuned = 1;
			st.tuned.push_back({key, plan});
		}
	}
	if (!plan.tile) plan.tile = 4096;
	if (!plan.stage) plan.stage = 16;
	if (!plan.sort) plan.sort = 2;
	if (!plan.mapping) plan.mapping = 2;

	std::vector<unsigned long long> hv(nb, 0);
	std::vector<uint64_t> hd(nb, 0);
	uint64_t clk[4] = {0, 0, 0, 0};
	unsigned long long hfb = 0;
	bool used_graph = false;
	memset(r, 0, sizeof(*r));
	r->kernel_id = 0;
	if (total > 0) {
		PipeShape ps;
		if ((rc = make_shape(d, nb, total, windows, first_dense_window(d, wpos, windows),
		                     a->xs_precision, plan, base, &ps)))
			return rc;
		BatchView B;
		SortedQueue q;
		make_views(ps, base, L, &B, &q);
		uint64_t *dclk = (uint64_t *)(base + L.clk);
		if (!a->eager && !(st.exec && st.shape == ps)) {
			if (st.exec) {
				(void)hipGraphExecDestroy(st.exec);
				st.exec = nullptr;
			}
			XSB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
			rc = enqueue_pipeline(d, ps, B, q, dclk, s);
			hipGraph_t graph = nullptr;
			const hipError_t ec = hipStreamEndCapture(s, &graph);
			if (rc || ec != hipSuccess) {
				if (graph)
					(void)hipGraphDestroy(graph);
				return rc ? rc : fail(XSB_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
			}
			const hipError_t ei = hipGraphInstantiate(&st.exec, graph, nullptr, nullptr, 0);
			(void)hipGraphDestroy(graph);
			if (ei != hipSuccess) {
				st.exec = nullptr;
				return fail(XSB_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
			}
			st.shape = ps;
		}
		XSB_HIP(hipEventRecord(d->ev[0], s));
		if (a->eager) {
			if ((rc = enqueue_pipeline(d, ps, B, q, dclk, s)))
				return rc;
		} else {
			XSB_HIP(hipGraphLaunch(st.exec, s));
			used_graph = true;
		}
		XSB_HIP(hipEventRecord(d->ev[3], s));
		XSB_HIP(hipMemcpyAsync(hv.data(), q.vsum, 8 * (size_t)nb, hipMemcpyDeviceToHost, s));
		XSB_HIP(hipMemcpyAsync(hd.data(), q.done_clock, 8 * (size_t)nb, hipMemcpyDeviceToHost, s));
		XSB_HIP(hipMemcpyAsync(clk, dclk, 8 * 3, hipMemcpyDeviceToHost, s));
		XSB_HIP(hipMemcpyAsync(&hfb, q.nfb, 8, hipMemcpyDeviceToHost, s));
		XSB_HIP(hipStreamSynchronize(s));
		float ms = 0;
		if (hipEventElapsedTime(&ms, d->ev[0], d->ev[3]) == hipSuccess)
			r->ms_total = ms;
	}
	int rate_khz = 0;
	XSB_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, d->device));
	const double per_ms = rate_khz > 0 ? 1.0 / rate_khz : 0.0;
	r->ms_sample = (double)(int64_t)(clk[1] - clk[0]) * per_ms;
	r->ms_lookup = (double)(int64_t)(clk[2] - clk[1]) * per_ms;
	uint64_t vs = 0;
	for (int b = 0; b < nb; b++) {
		vs += hv[b];
		if (a->out_vsum)
			a->out_vsum[b] = hv[b];
		if (a->out_done_ms)
			a->out_done_ms[b] = (a->sizes[b] > 0) ? (double)(int64_t)(hd[b] - clk[0]) * per_ms : 0.0;
	}
	r->vsum = vs;
	r->n_fallback = (int64_t)hfb;
	if (a->out_plan) {
		a->out_plan->tile = plan.tile;
		a->out_plan->stage = plan.stage;
		a->out_plan->sort = plan.sort;
		a->out_plan->windows = windows;
		a->out_plan->mapping = plan.mapping;
		a->out_plan->tuned = tuned;
		a->out_plan->graph = used_graph ? 1 : 0;
		a->out_plan->ms_tuning = ms_tuning;
	}
	return XSB_OK;
}

// ---- modes 1 and 2 ----
static int run_unsorted_batches(Device *d, const xsb_batched_args *a, xsb_result *r)
{
	const int nb = a->n_batches;
	int tile = a->tile > 0 ? a->tile : 4096;
	tile = ((tile + kBatchBlock - 1) / kBatchBlock) * kBatchBlock;
	if (tile > kBatchBlock * 16)
		tile = kBatchBlock * 16;
	const int prec = a->xs_precision;
	hipStream_t s = d->stream;
	memset(r, 0, sizeof(*r));
	r->kernel_id = a->mode;

	std::vector<int64_t> item_off(nb + 1, 0);
	std::vector<uint32_t> tiles(nb);
	for (int b = 0; b < nb; b++) {
		tiles[b] = (uint32_t)((a->sizes[b] + tile - 1) / tile);
		item_off[b + 1] = item_off[b] + tiles[b];
	}
	const size_t off_sizes = 0, off_seeds = off_sizes + 8 * nb, off_thr = off_seeds + 8 * nb,
	             off_item = off_thr + 8 * kMaterials * nb, off_vsum = off_item + 8 * (nb + 1),
	             off_nfb = off_vsum + 8 * nb, off_rem = off_nfb + 64, off_done = off_rem + 4 * nb + 64,
	             off_head = off_done + 8 * nb, off_tiles = off_head + 64, off_clk = off_tiles + 4 * nb + 64,
	             bytes = off_clk + 64;
	void *mem = nullptr;
	int rc = scratch_reserve(d, 1, bytes, &mem);
	if (rc)
		return rc;
	char *base = (char *)mem;
	XSB_HIP(hipMemcpyAsync(base + off_sizes, a->sizes, 8 * nb, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + off_seeds, a->seeds, 8 * nb, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + off_thr, a->thresholds, 8 * kMaterials * nb, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + off_item, item_off.data(), 8 * (nb + 1), hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(base + off_tiles, tiles.data(), 4 * nb, hipMemcpyHostToDevice, s));
	unsigned long long *vsum = (unsigned long long *)(base + off_vsum);
	unsigned long long *nfb = (unsigned long long *)(base + off_nfb);
	uint64_t *done = (uint64_t *)(base + off_done);
	uint64_t *clk = (uint64_t *)(base + off_clk);

	BatchQueue q;
	q.sizes = (const int64_t *)(base + off_sizes);
	q.seeds = (const uint64_t *)(base + off_seeds);
	q.thr = (const double *)(base + off_thr);
	q.item_off = (const int64_t *)(base + off_item);
	q.vsum = vsum;
	q.nfb = nfb;
	q.remaining = (uint32_t *)(base + off_rem);
	q.done_clock = done;
	q.next_item = (uint32_t *)(base + off_head);
	q.n_items = item_off[nb];
	q.nb = nb;
	q.tile = tile;
	hipLaunchKernelGGL(k_queue_reset, dim3((nb + 255) / 256), dim3(256), 0, s, q,
	                   (const uint32_t *)(base + off_tiles));
	XSB_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, clk);
	XSB_HIP(hipEventRecord(d->ev[0], s));
	if (a->mode == 1) {
		for (int b = 0; b < nb; b++) {
			const double *thr = (const double *)(base + off_thr) + b * kMaterials;
			if (prec == XSB_XS_FP64)
				launch_one<XSB_XS_FP64>(d, a->sizes[b], a->seeds[b], thr, vsum + b, nfb, s);
			else if (prec == XSB_XS_FP32)
				launch_one<XSB_XS_FP32>(d, a->sizes[b], a->seeds[b], thr, vsum + b, nfb, s);
			else
				launch_one<XSB_XS_BF16>(d, a->sizes[b], a->seeds[b], thr, vsum + b, nfb, s);
			XSB_HIP(hipGetLastError());
			hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, done + b);
		}
	} else {
		const unsigned grid = (unsigned)std::max(1, d->cu_count) * 8u;
		if (prec == XSB_XS_FP64)
			launch_persistent<XSB_XS_FP64>(d, q, grid, s);
		else if (prec == XSB_XS_FP32)
			launch_persistent<XSB_XS_FP32>(d, q, grid, s);
		else
			launch_persistent<XSB_XS_BF16>(d, q, grid, s);
		XSB_HIP(hipGetLastError());
	}
	XSB_HIP(hipEventRecord(d->ev[3], s));
	std::vector<unsigned long long> hv(nb);
	std::vector<uint64_t> hd(nb);
	uint64_t c0 = 0;
	unsigned long long hfb = 0;
	XSB_HIP(hipMemcpyAsync(hv.data(), vsum, 8 * nb, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipMemcpyAsync(hd.data(), done, 8 * nb, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipMemcpyAsync(&c0, clk, 8, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipMemcpyAsync(&hfb, nfb, 8, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	int rate_khz = 0;
	XSB_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, d->device));
	uint64_t vs = 0;
	for (int b = 0; b < nb; b++) {
		vs += hv[b];
		if (a->out_vsum)
			a->out_vsum[b] = hv[b];
		if (a->out_done_ms)
			a->out_done_ms[b] = (rate_khz > 0 && a->sizes[b] > 0)
			                        ? (double)(int64_t)(hd[b] - c0) / rate_khz
			                        : 0.0;
	}
	r->vsum = vs;
	r->n_fallback = (int64_t)hfb;
	float ms = 0;
	if (hipEventElapsedTime(&ms, d->ev[0], d->ev[3]) == hipSuccess)
		r->ms_total = r->ms_lookup = ms;
	if (a->out_plan) {
		memset(a->out_plan, 0, sizeof(*a->out_plan));
		a->out_plan->tile = tile;
	}
	return XSB_OK;
}

int run_batched(Device *d, const xsb_batched_args *a, xsb_result *r)
{
	if (!a || !r || a->n_batches < 1 || !a->sizes || !a->seeds || !a->thresholds)
		return fail(XSB_ERR_INVALID, "n_batches >= 1 with sizes, seeds and thresholds required");
	if (a->n_batches > (1 << 20))
		return fail(XSB_ERR_INVALID, "n_batches %d is more than 2^20", a->n_batches);
	if (a->mode < 0 || a->mode > 2)
		return fail(XSB_ERR_INVALID, "mode %d is not 0, 1 or 2", a->mode);
	if (a->tile < 0 || a->stage < 0 || a->sort < 0 || a->sort > 2 || a->window < -1)
		return fail(XSB_ERR_INVALID, "tile, stage must be >= 0, window >= -1 and sort 0, 1 or 2");
	const int nb = a->n_batches;
	for (int b = 0; b < nb; b++) {
		if (a->sizes[b] < 0 || a->sizes[b] > ((int64_t)1 << 31))
			return fail(XSB_ERR_INVALID, "batch %d size %lld out of range", b,
			            (long long)a->sizes[b]);
		for (int m = 0; m < kMaterials; m++)
			if (!(a->thresholds[b * kMaterials + m] >= (m ? a->thresholds[b * kMaterials + m - 1] : 0.0)))
				return fail(XSB_ERR_INVALID, "batch %d thresholds are not non-decreasing", b);
		if (!(a->thresholds[b * kMaterials + kMaterials - 1] >= 1.0))
			return fail(XSB_ERR_INVALID, "batch %d: thresholds[11] must be >= 1", b);
	}
	return a->mode == 0 ? run_sorted_batches(d, a, r) : run_unsorted_batches(d, a, r);
}

}  // namespace xsb

extern "C" int xsb_hip_batched(void *handle, const xsb_batched_args *args, xsb_result *res)
{
	using namespace xsb;
	if (!handle)
		return fail(XSB_ERR_INVALID, "handle is NULL");
	Device *d = static_cast<Device *>(handle);
	XSB_HIP(hipSetDevice(d->device));
	if (args && args->xs_precision == XSB_XS_FP32 && !d->d_xs32)
		return fail(XSB_ERR_UNSUPPORTED, "FP32 tables were not built (XSB_BUILD_FP32)");
	if (args && args->xs_precision == XSB_XS_BF16 && !d->d_xs16)
		return fail(XSB_ERR_UNSUPPORTED, "BF16 tables were not built (XSB_BUILD_BF16)");
	if (args && (args->xs_precision < 0 || args->xs_precision > 2))
		return fail(XSB_ERR_INVALID, "xs_precision %d is not 0, 1 or 2", args->xs_precision);
	return run_batched(d, args, res);
}
