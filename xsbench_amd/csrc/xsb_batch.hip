// xsb_batch.hip -- many independent lookup batches (no reference
// counterpart): each batch has its own sample stream and material CDF.
//
// Batch b, lookup i:  s = FF(seed_b, 2i), E = LCG(s), roll = LCG(s),
//                     mat = first m with roll < thr_b[m]  (thr_b[11] >= 1).
//
// mode 1: one fused launch per batch, in batch order.
// mode 0: one persistent kernel fed by a device work queue of tiles; each
//         tile (a slice of one batch) is ordered by material inside the
//         workgroup before its lookups run, so waves run one material; the
//         queue reset and the kernel are captured once in a hipGraph and the
//         graph is replayed.  Each batch records the device clock when its
//         last tile completes.
#include <algorithm>
#include <cstring>
#include <vector>

#include "xsb_internal.h"
#include "xsb_mixed.h"

namespace xsb {

constexpr int kBatchBlock = 256;

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

// ---- mode 0: persistent kernel over a queue of tiles ----
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

int run_batched(Device *d, const xsb_batched_args *a, xsb_result *r)
{
	if (!a || !r || a->n_batches < 1 || !a->sizes || !a->seeds || !a->thresholds)
		return fail(XSB_ERR_INVALID, "n_batches >= 1 with sizes, seeds and thresholds required");
	const int nb = a->n_batches;
	int64_t total = 0;
	for (int b = 0; b < nb; b++) {
		if (a->sizes[b] < 0 || a->sizes[b] > ((int64_t)1 << 31))
			return fail(XSB_ERR_INVALID, "batch %d size %lld out of range", b,
			            (long long)a->sizes[b]);
		for (int m = 0; m < kMaterials; m++)
			if (!(a->thresholds[b * kMaterials + m] >= (m ? a->thresholds[b * kMaterials + m - 1] : 0.0)))
				return fail(XSB_ERR_INVALID, "batch %d thresholds are not non-decreasing", b);
		if (!(a->thresholds[b * kMaterials + kMaterials - 1] >= 1.0))
			return fail(XSB_ERR_INVALID, "batch %d: thresholds[11] must be >= 1", b);
		total += a->sizes[b];
	}
	int tile = a->tile > 0 ? a->tile : 4096;
	tile = ((tile + kBatchBlock - 1) / kBatchBlock) * kBatchBlock;
	if (tile > kBatchBlock * 16)
		tile = kBatchBlock * 16;
	const int prec = a->xs_precision;
	hipStream_t s = d->stream;
	memset(r, 0, sizeof(*r));
	r->kernel_id = a->mode;

	// device arrays: sizes, seeds, thr, item_off, vsum, nfb, remaining, done, head, tiles, clock
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
	// queue reset with a fill kernel (no memset node in the graph)
	hipLaunchKernelGGL(k_queue_reset, dim3((nb + 255) / 256), dim3(256), 0, s, q,
	                   (const uint32_t *)(base + off_tiles));
	XSB_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, clk);
	XSB_HIP(hipEventRecord(d->ev[0], s));
	if (a->mode == 1) {
		std::vector<hipEvent_t> evs;
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
			a->out_done_ms[b] = rate_khz > 0 ? (double)(int64_t)(hd[b] - c0) / rate_khz : 0.0;
	}
	r->vsum = vs;
	r->n_fallback = (int64_t)hfb;
	r->ms_total = r->ms_lookup = 0;
	float ms = 0;
	if (hipEventElapsedTime(&ms, d->ev[0], d->ev[3]) == hipSuccess)
		r->ms_total = r->ms_lookup = ms;
	(void)total;
	return XSB_OK;
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
