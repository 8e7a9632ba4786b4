// xsb_sorted.h -- pieces shared by the sorted schedules of the event path
// (xsb_sched.hip) and of the batched path (xsb_batch.hip): the LSD radix sort
// of (key, id) pairs and the per-lane lookup body used in sorted order.
#pragma once

#include "xsb_internal.h"
#include "xsb_mixed.h"

namespace xsb {

// ---- radix sort of (key, value) pairs ----
constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kSortItems = 32;
constexpr int kTile = kSortBlock * kSortItems;  // 8192 elements per tile
constexpr int kDigitBits = 10;
constexpr int kBins = 1 << kDigitBits;
constexpr uint32_t kDigitMask = kBins - 1;
constexpr int kScanItems = 16;
constexpr int kScanChunk = 256 * kScanItems;  // 4096 entries per scan block

// Element e of tile t sits at position tile_pos(wave, item, lane) of the
// tile; the key kernels and the scatter agree on it.
__device__ __forceinline__ int tile_pos(int wave, int item, int lane)
{
	return wave * 64 * kSortItems + item * 64 + lane;
}

inline int sort_tiles(int64_t n) { return (int)((n + kTile - 1) / kTile); }

// The tile a sort kernel's workgroup takes: runs of kSortXcdRun neighbouring
// tiles per XCD (xcd_block), so that what neighbouring tiles write next to
// each other (histogram columns, a digit's output runs) merges into whole
// lines in one L2 before it reaches HBM.
constexpr unsigned kSortXcdRun = 16;
__device__ __forceinline__ int sort_tile() { return (int)xcd_block(kSortXcdRun); }

// Bytes of work memory the sorted event schedules take for n lookups: two
// buffers of {key, id} pairs (16-byte aligned), the per-tile digit histogram,
// the scan partials, and the last pass's digit starts.
constexpr size_t sorted_event_bytes(int64_t n)
{
	return (size_t)((n + 1) & ~(int64_t)1) * 16 +
	       (size_t)kBins * (size_t)((n + kTile - 1) / kTile) * 4 +
	       (size_t)(((int64_t)kBins * ((n + kTile - 1) / kTile) + kScanChunk - 1) / kScanChunk) *
	           4 +
	       64 + 4 * kBins;
}
inline int64_t sort_hist_len(int64_t n) { return (int64_t)kBins * sort_tiles(n); }
inline int sort_parts(int64_t n) { return (int)((sort_hist_len(n) + kScanChunk - 1) / kScanChunk); }

// Stable LSD sort of n pairs by the low kDigitBits * passes bits of the keys.
// hist holds sort_hist_len(n) words, part sort_parts(n) words.  When
// hist0_ready the caller's key kernel has already written the first pass's
// per-tile digit histogram (hist[digit * ntiles + tile]).  On return *k and
// *v point at the sorted pairs (either buffer of each pair).
int sort_pairs(hipStream_t s, int64_t n, int passes, bool hist0_ready, uint32_t **k, uint32_t **v,
               uint32_t **k2, uint32_t **v2, uint32_t *hist, uint32_t *part);
// The same for the lookup schedules, which need the order only for
// locality, on interleaved uint2 {key, value} pairs (one 8-byte word per
// element): sorted by key, but elements with equal keys may be permuted
// within a tile of kTile (faster: coalesced runs instead of a stable
// per-wave ranking).  *p points at the ordered pairs on return, or with
// ids_last at the ordered ids alone (uint32_t[n]).  starts, if given, gets
// the last pass's digit starts (kBins words).
int order_pairs(hipStream_t s, int64_t n, int passes, bool hist0_ready, uint2 **p, uint2 **p2,
                uint32_t *hist, uint32_t *part, bool ids_last = false,
                uint32_t *starts = nullptr);
// The same for 64-bit keys (no key-kernel histogram).
int sort_pairs64(hipStream_t s, int64_t n, int passes, uint64_t **k, uint32_t **v, uint64_t **k2,
                 uint32_t **v2, uint32_t *hist, uint32_t *part);

// Per-tile first-digit histogram from inside a key kernel: h is the block's
// LDS histogram (kBins words), zeroed by the caller.
__device__ __forceinline__ void write_tile_hist(const uint32_t *h, uint32_t *hist, int ntiles,
                                                int tile)
{
	for (int d = threadIdx.x; d < kBins; d += kSortBlock)
		hist[(int64_t)d * ntiles + tile] = h[d];
}

// ---- lookup body in sorted order ----
// One lane's lookup (p, mat).  Unionized waves whose lanes share the
// material take the wave-uniform path (macro_ueg_wave / _mixed, index words
// staged CHUNK deep) when `wave` is set; everything else runs the per-lane
// loop.  SHARE3 selects the wave paths' two-interval step (ueg_wave_nuclide).  Returns the argmax channel; idx gets the grid-level index,
// macro the sums, fb the fallback flag (mixed precision).
// HISTORY: the caller also tests macro_c > 1.0, so mixed precision falls
// back to FP64 near 1.0 too and then returns the FP64 sums.  DEFER: a mixed
// lookup that would fall back returns fb = 1 without its FP64 sums; the
// caller re-evaluates it.
template <int GRID, int PREC, int CHUNK, bool HISTORY = false, bool SHARE3 = true,
          bool DEFER = false>
__device__ __forceinline__ int sorted_lookup(const Tables &T, double p, int mat, bool wave,
                                             int64_t &idx, double macro[5], uint32_t &fb,
                                             double *bnd)
{
	idx = lookup_index<GRID>(T, p);
	if (PREC == XSB_XS_FP64 && GRID == kUnionized) {
		const int m0 = __builtin_amdgcn_readfirstlane(mat);
		if (wave && __ballot(mat != m0) == 0ULL)
			macro_ueg_wave<CHUNK, SHARE3>(T, p, m0, idx, macro);
		else
			macro_xs_fp64<GRID, HISTORY>(T, p, mat, idx, macro);
		return argmax5(macro);
	} else if (GRID == kHash) {
		// wave path when the lanes share the material and the hash bin
		const int m0 = __builtin_amdgcn_readfirstlane(mat);
		const int i0 = __builtin_amdgcn_readfirstlane((int)idx);
		if (wave && __ballot(mat != m0 || (int)idx != i0) == 0ULL) {
			if (PREC == XSB_XS_FP64) {
				macro_hash_wave<SHARE3>(T, p, m0, i0, macro);
				return argmax5(macro);
			}
			return macro_hash_wave_mixed<PREC, HISTORY, SHARE3, DEFER>(T, p, m0, i0, macro, &fb,
			                                                           bnd);
		}
	} else if (GRID == kUnionized && PREC != XSB_XS_FP64) {
		const int m0 = __builtin_amdgcn_readfirstlane(mat);
		if (wave && __ballot(mat != m0) == 0ULL)
			return macro_ueg_wave_mixed<PREC, CHUNK, HISTORY, SHARE3, DEFER>(T, p, m0, idx, macro,
			                                                                 &fb, bnd);
	}
	if (PREC == XSB_XS_FP64) {
		macro_xs_fp64<GRID, HISTORY>(T, p, mat, idx, macro);
		return argmax5(macro);
	}
	return macro_xs_mixed<GRID, PREC, HISTORY, DEFER>(T, p, mat, idx, macro, &fb, bnd);
}

}  // namespace xsb
