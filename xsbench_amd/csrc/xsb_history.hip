// xsb_history.hip -- history-based lookups in rounds (kernel_id 6 of
// xsb_hip_history; openmp-threading/Simulation.c:116-238).
//
// A particle's lookups are a dependent chain, but the particles are
// independent: round r performs lookup r of every particle.  Each round
// orders the particles by (material, energy of their next lookup) with the
// radix sort of the sorted event schedules, then looks them up in that
// order, so neighbouring lanes read neighbouring grid records, and advances
// each particle's LCG state exactly as the reference's inner loop does.
// Particle q's state between rounds: seed[q], the LCG state from which its
// next energy and material are drawn, and psum[q], its verification sum.
// The whole chain of rounds is captured once in a hipGraph and replayed.
#include <algorithm>
#include <cstring>
#include <vector>

#include "xsb_sorted.h"

namespace xsb {

constexpr int kHBlock = 256;
constexpr int kHEnergyBits = 16;
// Round kernel workgroup (512 threads: 30.0 -> 31.3 ms for the XL-45000
// run, 1024: 31.6 ms).
constexpr int kHRoundBlock = 256;
constexpr unsigned kHXcdRun = 4096 / kHRoundBlock;

__global__ __launch_bounds__(kHBlock) void k_hinit(Tables T, int64_t P, int lookups, uint64_t seed0,
                                                   uint64_t *seed, uint32_t *psum)
{
	const int64_t q = (int64_t)blockIdx.x * kHBlock + threadIdx.x;
	if (q < P) {
		seed[q] = ff_jump(seed0, history_offset(q, lookups), T.ff_tab);
		psum[q] = 0;
	}
}

// key = material << 16 | top 16 bits of the next energy; value = particle.
// Also empties the round's re-evaluation queue (the previous round's
// k_hfallback has finished with it).
__global__ __launch_bounds__(kSortBlock) void k_hkeys(Tables T, int64_t P,
                                                      const uint64_t *__restrict__ seed,
                                                      uint2 *pairs, uint32_t *hist, int ntiles,
                                                      uint32_t *queue)
{
	if (blockIdx.x == 0 && threadIdx.x == 0)
		queue[0] = 0;
	__shared__ uint32_t h[kBins];
	for (int d = threadIdx.x; d < kBins; d += kSortBlock)
		h[d] = 0;
	__syncthreads();
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tile = sort_tile();
	const int64_t base = (int64_t)tile * kTile;
	for (int it = 0; it < kSortItems; it++) {
		const int64_t e = base + tile_pos(wave, it, lane);
		if (e < P) {
			uint64_t s = seed[e];
			const double p = lcg_double(s);
			const int mat = pick_mat(s, T.thr);
			uint32_t eb = (uint32_t)(p * (double)(1u << kHEnergyBits));
			eb = eb < (1u << kHEnergyBits) ? eb : (1u << kHEnergyBits) - 1u;
			const uint32_t key = ((uint32_t)mat << kHEnergyBits) | eb;
			pairs[e] = make_uint2(key, (uint32_t)e);
			atomicAdd(&h[key & kDigitMask], 1u);
		}
	}
	__syncthreads();
	write_tile_hist(h, hist, ntiles, tile);
}

// One round in sorted order (Simulation.c:176-234 for one i): draw the
// energy and material, look up, add argmax + 1, fast-forward the state by
// the number of channels above 1.0.  Lanes past P repeat the last particle
// and write nothing.  ids: the sorted {key, particle} pairs.
// Mixed precision: a lookup the bounds cannot decide is not re-evaluated
// here -- its lane's FP64 loop would hold the wave (most often a fuel wave,
// the round's longest, as the undecided lookups are mostly fuel lookups
// with many terms) for another full chain of dependent loads -- but queued
// (queue[0] counts, queue[1 ..] particles) for k_hfallback, its particle's
// state left as it was.
template <int GRID, int PREC, int CHUNK>
__global__ __launch_bounds__(kHRoundBlock) void k_hround(Tables T, int64_t P,
                                                    const uint2 *__restrict__ ids,
                                                    uint64_t *__restrict__ seed,
                                                    uint32_t *__restrict__ psum,
                                                    unsigned long long *vsum,
                                                    unsigned long long *nfb,
                                                    uint32_t *__restrict__ queue)
{
	const int64_t i = (int64_t)xcd_block(kHXcdRun) * kHRoundBlock + threadIdx.x;
	const bool valid = i < P;
	const uint32_t q = ids[valid ? i : P - 1].y;
	uint64_t s = seed[q];
	const double p = lcg_double(s);
	const int mat = pick_mat(s, T.thr);
	int64_t idx;
	double macro[5];
	uint32_t fb = 0;
	const int am = sorted_lookup<GRID, PREC, CHUNK, true, true, PREC != XSB_XS_FP64>(
	    T, p, mat, T.variant & kVariantWave, idx, macro, fb, nullptr);
	const bool done = valid && !fb;
	if (valid && fb)
		queue[1 + atomicAdd(queue, 1u)] = q;
	uint32_t n_forward = 0;
#pragma unroll
	for (int c = 0; c < 5; c++)
		n_forward += (macro[c] > 1.0) ? 1u : 0u;
	if (n_forward)
		s = ff_jump(s, n_forward, T.ff_tab);
	if (done) {
		seed[q] = s;
		psum[q] += (uint32_t)am + 1u;
	}
	block_add_u64<kHRoundBlock>(done ? (uint32_t)am + 1u : 0u, vsum);
	if (PREC != XSB_XS_FP64)
		block_add_u64<kHRoundBlock>(valid ? fb : 0u, nfb);
}

// The queued lookups of a round in FP64, one workgroup each: thread t
// computes the terms xs_c * conc of nuclides t, t + kFbBlock, ...
// (Simulation.c:360-374) into LDS, threads 0 .. 4 add channel c's terms in
// list order -- the sums of macro_xs_fp64 bit for bit, after a dependent
// chain of one or two nuclides instead of the whole list -- and thread 0
// completes the round for the particle as k_hround does.  Nuclides go
// through LDS kFbChunk at a time.
constexpr int kFbBlock = 256;
constexpr int kFbChunk = 512;

template <int GRID>
__global__ __launch_bounds__(kFbBlock) void k_hfallback(Tables T, const uint32_t *__restrict__ queue,
                                                        uint64_t *__restrict__ seed,
                                                        uint32_t *__restrict__ psum,
                                                        unsigned long long *vsum)
{
	__shared__ double t[kFbChunk * 5];
	__shared__ double sums[5];
	const int tid = threadIdx.x;
	const uint32_t n = queue[0];
	for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
		const uint32_t q = queue[1 + w];
		uint64_t s = seed[q];
		const double p = lcg_double(s);
		const int mat = pick_mat(s, T.thr);
		const int64_t idx = lookup_index<GRID>(T, p);
		const int nn = T.num_nucs[mat];
		const int32_t *__restrict__ ml = T.mats + mat * T.max_num_nucs;
		const double *__restrict__ cl = T.concs + mat * T.max_num_nucs;
		const uint8_t *__restrict__ mrow = (GRID == kUnionized) ? ueg_row(T, mat, idx) : nullptr;
		double acc = 0.0;  // threads 0 .. 4: channel `tid`
		for (int j0 = 0; j0 < nn; j0 += kFbChunk) {
			const int nc = min(kFbChunk, nn - j0);
#pragma unroll 2
			for (int jj = tid; jj < nc; jj += kFbBlock) {
				const int j = j0 + jj;
				const int nuc = ml[j];
				const double *row = T.grid + (int64_t)nuc * T.n_gp * kPointDoubles;
				const int32_t lower = nuclide_lower<GRID, kPointDoubles, true>(T, row, nuc, mrow, j,
				                                                               idx, p);
				double tt[5];
				term_fp64(row + (int64_t)lower * kPointDoubles, p, cl[j], tt);
#pragma unroll
				for (int c = 0; c < 5; c++)
					t[jj * 5 + c] = tt[c];
			}
			__syncthreads();
			if (tid < 5)
				for (int jj = 0; jj < nc; jj++)
					acc += t[jj * 5 + tid];
			__syncthreads();
		}
		if (tid < 5)
			sums[tid] = acc;
		__syncthreads();
		if (tid == 0) {
			double macro[5];
#pragma unroll
			for (int c = 0; c < 5; c++)
				macro[c] = sums[c];
			const int am = argmax5(macro);
			uint32_t n_forward = 0;
#pragma unroll
			for (int c = 0; c < 5; c++)
				n_forward += (macro[c] > 1.0) ? 1u : 0u;
			if (n_forward)
				s = ff_jump(s, n_forward, T.ff_tab);
			seed[q] = s;
			psum[q] += (uint32_t)am + 1u;
			atomicAdd(vsum, (unsigned long long)(am + 1));
		}
		__syncthreads();
	}
}

// Per-particle outputs: the sum widened to 64 bits, and the final state
// after the draws of the energy and material that follow the last lookup
// (the reference draws them at the end of every iteration).
__global__ __launch_bounds__(kHBlock) void k_hout(int64_t P, const uint64_t *__restrict__ seed,
                                                  const uint32_t *__restrict__ psum,
                                                  uint64_t *out_sum, uint64_t *out_seed)
{
	const int64_t q = (int64_t)blockIdx.x * kHBlock + threadIdx.x;
	if (q < P) {
		uint64_t s = seed[q];
		(void)lcg_double(s);
		(void)lcg_double(s);
		if (out_sum)
			out_sum[q] = psum[q];
		if (out_seed)
			out_seed[q] = s;
	}
}

__global__ void k_hreset(unsigned long long *acc)
{
	if (threadIdx.x < 2)
		acc[threadIdx.x] = 0;
}

// k_hfallback's grid: the queue holds a few hundred lookups per million
// with FP32 tables, about a tenth of them with BF16's.
static unsigned fallback_blocks(int prec, int64_t P)
{
	const int64_t want = prec == XSB_XS_BF16 ? 4096 : 512;
	return (unsigned)std::max<int64_t>(1, std::min(want, P));
}

template <int GRID, int PREC>
static void launch_round(Device *d, hipStream_t s, int64_t P, const uint2 *ids, uint64_t *seed,
                         uint32_t *psum, unsigned long long *acc, uint32_t *queue)
{
	const unsigned nb = (unsigned)((P + kHRoundBlock - 1) / kHRoundBlock);
	constexpr int CHUNK = kIdxChunk;
	hipLaunchKernelGGL((k_hround<GRID, PREC, CHUNK>), dim3(nb), dim3(kHRoundBlock), 0, s, d->T, P, ids,
	                   seed, psum, acc, acc + 1, queue);
	if (PREC != XSB_XS_FP64)
		hipLaunchKernelGGL((k_hfallback<GRID>), dim3(fallback_blocks(PREC, P)), dim3(kFbBlock), 0, s,
		                   d->T, queue, seed, psum, acc);
}

template <int PREC>
static void launch_round_grid(Device *d, hipStream_t s, int64_t P, const uint2 *ids,
                              uint64_t *seed, uint32_t *psum, unsigned long long *acc,
                              uint32_t *queue)
{
	switch (d->T.grid_type) {
	case kUnionized: launch_round<kUnionized, PREC>(d, s, P, ids, seed, psum, acc, queue); break;
	case kNuclide: launch_round<kNuclide, PREC>(d, s, P, ids, seed, psum, acc, queue); break;
	default: launch_round<kHash, PREC>(d, s, P, ids, seed, psum, acc, queue); break;
	}
}

struct HistoryBuffers {
	uint64_t *seed;
	uint32_t *psum, *hist, *part;
	uint32_t *queue;  // [1 + P]: count, then the particles to re-evaluate in FP64
	uint2 *pA, *pB;  // {key, particle} pairs, two buffers
	uint64_t *out_sum, *out_seed;
	unsigned long long *acc;
};

// Enqueues the whole run (kernels only: capturable).
static int enqueue_rounds(Device *d, hipStream_t s, int64_t P, int lookups, uint64_t seed0, int prec,
                          const HistoryBuffers &B)
{
	const unsigned nb = (unsigned)((P + kHBlock - 1) / kHBlock);
	const int ntiles = sort_tiles(P);
	hipLaunchKernelGGL(k_hreset, dim3(1), dim3(64), 0, s, B.acc);
	hipLaunchKernelGGL(k_hinit, dim3(nb), dim3(kHBlock), 0, s, d->T, P, lookups, seed0, B.seed,
	                   B.psum);
	XSB_HIP(hipGetLastError());
	for (int r = 0; r < lookups; r++) {
		uint2 *pA = B.pA, *pB = B.pB;
		hipLaunchKernelGGL(k_hkeys, dim3(ntiles), dim3(kSortBlock), 0, s, d->T, P, B.seed, pA,
		                   B.hist, ntiles, B.queue);
		XSB_HIP(hipGetLastError());
		int rc = order_pairs(s, P, 2, true, &pA, &pB, B.hist, B.part);
		if (rc)
			return rc;
		if (prec == XSB_XS_FP64)
			launch_round_grid<XSB_XS_FP64>(d, s, P, pA, B.seed, B.psum, B.acc, B.queue);
		else if (prec == XSB_XS_FP32)
			launch_round_grid<XSB_XS_FP32>(d, s, P, pA, B.seed, B.psum, B.acc, B.queue);
		else
			launch_round_grid<XSB_XS_BF16>(d, s, P, pA, B.seed, B.psum, B.acc, B.queue);
		XSB_HIP(hipGetLastError());
	}
	if (B.out_sum || B.out_seed) {
		hipLaunchKernelGGL(k_hout, dim3(nb), dim3(kHBlock), 0, s, P, B.seed, B.psum, B.out_sum,
		                   B.out_seed);
		XSB_HIP(hipGetLastError());
	}
	return XSB_OK;
}

int run_history_rounds(Device *d, const xsb_history_args *a, xsb_result *r)
{
	const int64_t P = a->particles;
	const int L = a->lookups_per_particle;
	hipStream_t s = d->stream;
	const bool outs = a->out_particle_sum || a->out_final_seed;
	// the nine buffers below, each rounded up to 256 bytes by take()
	const size_t bytes = (size_t)P * (8 + 4 + 16) + (size_t)sort_hist_len(P) * 4 +
	                     (size_t)sort_parts(P) * 4 + (outs ? (size_t)P * 16 : 0) +
	                     ((size_t)P + 1) * 4 + 9 * 256;
	void *buf = nullptr;
	int rc = scratch_reserve(d, 3, bytes, &buf);
	if (rc)
		return rc;
	HistoryBuffers B;
	char *p = (char *)buf;
	auto take = [&p](size_t n) {
		char *q = p;
		p += (n + 255) & ~(size_t)255;
		return q;
	};
	B.seed = (uint64_t *)take((size_t)P * 8);
	B.psum = (uint32_t *)take((size_t)P * 4);
	B.pA = (uint2 *)take((size_t)P * 8);
	B.pB = (uint2 *)take((size_t)P * 8);
	B.hist = (uint32_t *)take((size_t)sort_hist_len(P) * 4);
	B.part = (uint32_t *)take((size_t)sort_parts(P) * 4);
	B.out_sum = a->out_particle_sum ? (uint64_t *)take((size_t)P * 8) : nullptr;
	B.out_seed = a->out_final_seed ? (uint64_t *)take((size_t)P * 8) : nullptr;
	B.queue = (uint32_t *)take(((size_t)P + 1) * 4);
	B.acc = d->d_acc + 8;

	// the graph is keyed on everything its nodes hold
	std::vector<int64_t> key = {P, L, (int64_t)a->seed, a->xs_precision, d->T.grid_type,
	                            (int64_t)(intptr_t)buf, (int64_t)(bool)B.out_sum,
	                            (int64_t)(bool)B.out_seed};
	GraphCache &gc = d->history_graph;
	if (!(gc.exec && gc.key == key)) {
		if (gc.exec) {
			(void)hipGraphExecDestroy(gc.exec);
			gc.exec = nullptr;
		}
		XSB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
		rc = enqueue_rounds(d, s, P, L, a->seed, a->xs_precision, B);
		hipGraph_t graph = nullptr;
		const hipError_t ec = hipStreamEndCapture(s, &graph);
		if (rc || ec != hipSuccess) {
			if (graph)
				(void)hipGraphDestroy(graph);
			return rc ? rc : fail(XSB_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
		}
		const hipError_t ei = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
		(void)hipGraphDestroy(graph);
		if (ei != hipSuccess) {
			gc.exec = nullptr;
			return fail(XSB_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
		}
		gc.key = key;
	}
	XSB_HIP(hipEventRecord(d->ev[0], s));
	XSB_HIP(hipGraphLaunch(gc.exec, s));
	XSB_HIP(hipEventRecord(d->ev[3], s));
	unsigned long long acc[2];
	XSB_HIP(hipMemcpyAsync(acc, B.acc, sizeof(acc), hipMemcpyDeviceToHost, s));
	if (a->out_particle_sum)
		XSB_HIP(hipMemcpyAsync(a->out_particle_sum, B.out_sum, (size_t)P * 8, hipMemcpyDeviceToHost, s));
	if (a->out_final_seed)
		XSB_HIP(hipMemcpyAsync(a->out_final_seed, B.out_seed, (size_t)P * 8, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	r->vsum = acc[0];
	r->n_fallback = (int64_t)acc[1];
	float ms = 0;
	if (hipEventElapsedTime(&ms, d->ev[0], d->ev[3]) == hipSuccess)
		r->ms_total = r->ms_lookup = ms;
	r->kernel_id = 6;
	return XSB_OK;
}

}  // namespace xsb
