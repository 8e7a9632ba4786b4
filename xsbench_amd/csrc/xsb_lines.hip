// xsb_lines.hip -- compulsory bytes of a launch (a measurement helper, no
// reference counterpart): the distinct 128-byte lines of device memory that
// a workload's lookups read at least once, each line counted once however
// many lookups share it.  No kernel can move fewer bytes from HBM than this
// for the same launch on the same layout (short of lines still resident in
// the last-level cache from an earlier launch), so time against it gives a
// roofline fraction at most 1.
//
// A counting kernel replays each lookup's reads -- the unionized search's
// bracket words and probes, the material's index row, each nuclide's
// interval record or points (hash grid: the bin's index words, the energies
// of the bin's range and the bracket points) -- and marks their lines in a
// bitmap per table; a popcount of the bitmaps gives the lines.  The reads
// follow the lookup kernels of each workload:
//   event k >= 2, batched, history rounds: the sorted kernels' wave paths
//     (FP64 interval records on the unionized grid; the waves that straddle
//     a material boundary, at most 11 per sort, read points and are counted
//     as if they read records);
//   event k = 0 / 1: one lane per lookup (FP64 points);
//   FP32 / BF16 tables: their records; the few lookups re-evaluated in FP64
//     (under 1% with FP32 tables) re-read FP64 data that is not counted.
// The schedule's own buffers that the lookup kernel reads or writes (sorted
// pairs, sample arrays, particle states) and the small tables (materials,
// skip table) are added by size.
#include <cstring>
#include <vector>

#include "xsb_internal.h"
#include "xsb_mixed.h"
#include "xsb_sorted.h"

namespace xsb {

constexpr int kLineShift = 7;  // 128-byte lines

enum LineRegion : int {
	kRegUeg = 0,  // unionized energies
	kRegAcc,      // unionized search brackets
	kRegMidx,     // per-material index tables
	kRegIvl,      // FP64 interval records
	kRegIvl32,    // FP32 interval records
	kRegIndex,    // hash index grid
	kRegGrid,     // FP64 points
	kRegXs32,     // FP32 records
	kRegXs16,     // BF16 records
	kRegHwin,     // hash window records
	kRegEsoa,     // hash energy rows
	kRegCount
};

struct LineMap {
	uint32_t *bits;                 // the regions' bitmaps, one after another
	int64_t off[kRegCount];         // first bitmap word of each region
	const uint8_t *base[kRegCount]; // first byte of each region
};

__device__ __forceinline__ void touch(const LineMap &L, int r, const void *p, int64_t len)
{
	const int64_t o = (const uint8_t *)p - L.base[r];
	for (int64_t l = o >> kLineShift; l <= (o + len - 1) >> kLineShift; l++)
		atomicOr(&L.bits[L.off[r] + (l >> 5)], 1u << (l & 31));
}

// lookup_index<kUnionized> with every word it reads marked
__device__ __forceinline__ int64_t mark_ueg_search(const Tables &T, const LineMap &L, double p)
{
	int32_t lo, hi;
	const int32_t n = (int32_t)T.n_union;
	if (p >= 0.0 && p < 1.0) {
		const int b = (int)(p * (double)(1 << kUegAccBits));
		touch(L, kRegAcc, T.ueg_acc + b, 8);
		lo = T.ueg_acc[b] - 1;
		hi = T.ueg_acc[b + 1];
	} else {
		lo = 0;
		hi = n - 1;
	}
	while (hi - lo > 1) {
		const int32_t mid = lo + ((hi - lo) >> 1);
		touch(L, kRegUeg, T.ueg + mid, 8);
		if (T.ueg[mid] > p)
			hi = mid;
		else
			lo = mid;
	}
	if (p >= 0.0 && p < 1.0)
		return min(max(hi - 1, 0), n - 2);
	return lo;
}

// The reads of one lookup (p, mat) on precision PREC's tables.  WAVE: the
// sorted kernels' interval records (unionized FP64); else points.  HIST:
// the history rounds' hash bracket (hash_lower_window: the bin's window
// record, more of the energy row where the bin spans further); else the
// index grid and the bin's range of points.
template <int GRID, int PREC, bool WAVE, bool HIST = false>
__device__ __forceinline__ void mark_lookup(const Tables &T, const LineMap &L, double p, int mat)
{
	const int nn = T.num_nucs[mat];
	const int32_t *ml = T.mats + mat * T.max_num_nucs;
	const int32_t n_gp = T.n_gp;
	if (GRID == kUnionized) {
		const int64_t idx = mark_ueg_search(T, L, p);
		const uint8_t *row = ueg_row(T, mat, idx);
		if (nn > 0)
			touch(L, kRegMidx, row, (int64_t)nn << T.midx_shift);
		for (int j = 0; j < nn; j++) {
			const int64_t rec = (int64_t)ml[j] * n_gp + ueg_bracket(T, row, j);
			if (PREC == XSB_XS_FP64 && WAVE)
				touch(L, kRegIvl, T.ivl + rec * kIvlDoubles, kIvlDoubles * 8);
			else if (PREC == XSB_XS_FP64)
				touch(L, kRegGrid, T.grid + rec * kPointDoubles, 2 * kPointDoubles * 8);
			else if (PREC == XSB_XS_FP32)
				touch(L, kRegIvl32, T.ivl32 + rec * kIvl32Doubles, kIvl32Doubles * 8);
			else
				touch(L, kRegXs16, reinterpret_cast<const double *>(T.xs16) + rec * kRec16Doubles,
				      2 * kRec16Doubles * 8);
		}
	} else if (GRID == kHash) {
		const int64_t idx = lookup_index<kHash>(T, p);
		constexpr int REC = (PREC == XSB_XS_FP64)   ? kPointDoubles
		                    : (PREC == XSB_XS_FP32) ? kRec32Doubles
		                                            : kRec16Doubles;
		constexpr int REG = (PREC == XSB_XS_FP64) ? kRegGrid : (PREC == XSB_XS_FP32) ? kRegXs32 : kRegXs16;
		const double *base = (PREC == XSB_XS_FP64)   ? T.grid
		                     : (PREC == XSB_XS_FP32) ? reinterpret_cast<const double *>(T.xs32)
		                                             : reinterpret_cast<const double *>(T.xs16);
		const bool last = idx == T.hash_bins - 1;
		for (int j = 0; j < nn; j++) {
			const int nuc = ml[j];
			if (HIST) {
				const double *R = T.hwin + (idx * T.idx_pitch + nuc) * kHwinDoubles;
				touch(L, kRegHwin, R, kHwinDoubles * 8);
				const int32_t u_low = __double2loint(R[0]), u_high = __double2hiint(R[0]);
				const int32_t b = u_low & ~1, top = u_high - b;
				const double *E = T.esoa + (int64_t)nuc * T.e_pitch;
				if (!(T.variant & kVariantSortedRows) || top > 17) {
					touch(L, kRegEsoa, E + u_high, 8);
					if (E[u_low] < p && p < E[u_high]) {
						int32_t lo = u_low, hi = u_high;
						while (hi - lo > 1) {
							const int32_t mid = lo + ((hi - lo) >> 1);
							touch(L, kRegEsoa, E + mid, 8);
							if (E[mid] > p)
								hi = mid;
							else
								lo = mid;
						}
					}
				} else if (top > 9) {
					touch(L, kRegEsoa, E + b + 10, 8 * 8);
				}
				const int32_t lower = nuclide_lower<kHash, REC, true>(T, nullptr, nuc, nullptr, j, idx, p);
				touch(L, REG, base + ((int64_t)nuc * n_gp + lower) * REC, 2 * REC * 8);
				continue;
			}
			const int32_t *w = T.index_grid + idx * T.idx_pitch + nuc;
			touch(L, kRegIndex, w, 4);
			if (!last)
				touch(L, kRegIndex, w + T.idx_pitch, 4);
			const int32_t u_low = w[0];
			const int32_t u_high = last ? n_gp - 1 : w[T.idx_pitch] + 1;
			const double *row = base + (int64_t)nuc * n_gp * REC;
			// the energies of the bin's range (the bracket search reads them or a subset)
			touch(L, REG, row + (int64_t)u_low * REC, (int64_t)(u_high - u_low) * REC * 8 + 8);
			const int32_t lower = nuclide_lower<kHash, REC>(T, row, nuc, nullptr, j, idx, p);
			touch(L, REG, row + (int64_t)lower * REC, 2 * REC * 8);
		}
	}
}

// ---- the event workload: lookups [first, first + n) ----
template <int GRID, int PREC, bool WAVE>
__global__ __launch_bounds__(256) void k_lines_event(Tables T, LineMap L, int64_t first, int64_t n,
                                                     uint64_t seed0)
{
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= n)
		return;
	uint64_t s = ff_jump(seed0, 2ULL * (uint64_t)(first + i), T.ff_tab);
	const double p = lcg_double(s);
	const int mat = pick_mat(s, T.thr);
	mark_lookup<GRID, PREC, WAVE>(T, L, p, mat);
}

// ---- the batched workload: global lookup g of the batches (off[b] = first
// id of batch b, off[nb] = total) ----
template <int GRID, int PREC>
__global__ __launch_bounds__(256) void k_lines_batched(Tables T, LineMap L, const int64_t *off,
                                                       int nb, const uint64_t *seeds,
                                                       const double *thr)
{
	const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (g >= off[nb])
		return;
	int lo = 0, hi = nb;  // last b with off[b] <= g
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (off[mid] <= g)
			lo = mid;
		else
			hi = mid;
	}
	uint64_t s = ff_jump(seeds[lo], 2ULL * (uint64_t)(g - off[lo]), T.ff_tab);
	const double p = lcg_double(s);
	const double roll = lcg_double(s);
	int mat = 0;
	for (int k = 0; k < kMaterials - 1; k++)
		mat += (thr[lo * kMaterials + k] <= roll) ? 1 : 0;
	mark_lookup<GRID, PREC, true>(T, L, p, mat);
}

// ---- the history workload: one particle per lane replays its FP64 chain
// (openmp-threading/Simulation.c:141-237); lookup r marks bitmap set r ----
template <int GRID, int PREC>
__global__ __launch_bounds__(256) void k_lines_history(Tables T, LineMap L, int64_t set_words,
                                                       int64_t particles, int lookups,
                                                       uint64_t seed0)
{
	const int64_t pid = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (pid >= particles)
		return;
	uint64_t s = ff_jump(seed0, history_offset(pid, lookups), T.ff_tab);
	double p = lcg_double(s);
	int mat = pick_mat(s, T.thr);
	LineMap Lr = L;
	for (int r = 0; r < lookups; r++) {
		Lr.bits = L.bits + (int64_t)r * set_words;
		mark_lookup<GRID, PREC, true, true>(T, Lr, p, mat);
		const int64_t idx = lookup_index<GRID>(T, p);
		double macro[5];
		macro_xs_fp64<GRID>(T, p, mat, idx, macro);
		uint32_t n_forward = 0;
		for (int c = 0; c < 5; c++)
			n_forward += (macro[c] > 1.0) ? 1u : 0u;
		if (n_forward)
			s = ff_lcg(s, n_forward);
		p = lcg_double(s);
		mat = pick_mat(s, T.thr);
	}
}

__global__ void k_popcount(const uint32_t *__restrict__ bits, int64_t words,
                           unsigned long long *out)
{
	unsigned long long c = 0;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
	     i += (int64_t)gridDim.x * blockDim.x)
		c += (unsigned long long)__popc(bits[i]);
	for (int off = 32; off > 0; off >>= 1)
		c += __shfl_xor(c, off, 64);
	if ((threadIdx.x & 63) == 0 && c)
		atomicAdd(out, c);
}

__global__ void k_zero_words(uint32_t *bits, int64_t words)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
	     i += (int64_t)gridDim.x * blockDim.x)
		bits[i] = 0u;
}

// The regions of the handle's tables, their bitmaps laid out one after
// another (words per set).
static int64_t line_regions(const Device *d, LineMap *L)
{
	const Tables &T = d->T;
	const int64_t n = (int64_t)T.n_iso * T.n_gp;
	const int64_t idx_rows = T.grid_type == kUnionized ? 0 : (T.grid_type == kHash ? T.hash_bins : 0);
	const void *base[kRegCount] = {T.ueg,  T.ueg_acc, T.midx, T.ivl,  T.ivl32, T.index_grid,
	                               T.grid, T.xs32, T.xs16, T.hwin, T.esoa};
	const int64_t bytes[kRegCount] = {
	    T.ueg ? T.n_union * 8 : 0,
	    T.ueg_acc ? ((int64_t)(1 << kUegAccBits) + 1) * 4 : 0,
	    d->bytes_midx,
	    T.ivl ? n * kIvlDoubles * 8 : 0,
	    T.ivl32 ? n * kIvl32Doubles * 8 : 0,
	    T.index_grid ? idx_rows * T.idx_pitch * 4 : 0,
	    n * kPointDoubles * 8,
	    T.xs32 ? n * kRec32Doubles * 8 : 0,
	    T.xs16 ? n * kRec16Doubles * 8 : 0,
	    T.hwin ? idx_rows * T.idx_pitch * kHwinDoubles * 8 : 0,
	    T.esoa ? (int64_t)T.n_iso * T.e_pitch * 8 : 0};
	int64_t words = 0;
	for (int r = 0; r < kRegCount; r++) {
		L->base[r] = (const uint8_t *)base[r];
		L->off[r] = words;
		words += ((bytes[r] >> kLineShift) + 1 + 31) / 32;
	}
	return words;
}

// Counts the marked lines of `sets` consecutive bitmap sets, per region
// (summed over sets), into region_lines[kRegCount].
static int count_lines(Device *d, const LineMap &L, int64_t set_words, int sets,
                       int64_t *region_lines)
{
	hipStream_t s = d->stream;
	unsigned long long *acc = d->d_acc + 48;  // kRegCount words
	static_assert(48 + kRegCount <= 60, "accumulator words");
	static_assert(kRegCount == XSB_LINE_REGIONS, "one region_bytes entry per region");
	XSB_HIP(hipMemsetAsync(acc, 0, kRegCount * 8, s));
	for (int r = 0; r < kRegCount; r++) {
		const int64_t w = (r + 1 < kRegCount ? L.off[r + 1] : set_words) - L.off[r];
		for (int k = 0; k < sets; k++)
			hipLaunchKernelGGL(k_popcount, dim3(256), dim3(256), 0, s,
			                   L.bits + (int64_t)k * set_words + L.off[r], w, acc + r);
	}
	XSB_HIP(hipGetLastError());
	unsigned long long h[kRegCount];
	XSB_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	for (int r = 0; r < kRegCount; r++)
		region_lines[r] = (int64_t)h[r];
	return XSB_OK;
}

static int alloc_bits(Device *d, int64_t words, uint32_t **bits)
{
	if (hipMalloc((void **)bits, (size_t)words * 4) != hipSuccess) {
		(void)hipGetLastError();
		*bits = nullptr;
		return fail(XSB_ERR_OOM, "line bitmap of %lld words", (long long)words);
	}
	hipLaunchKernelGGL(k_zero_words, dim3(1024), dim3(256), 0, d->stream, *bits, words);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

static void fill_result(const int64_t *region_lines, int64_t schedule_bytes, int64_t small_bytes,
                        int64_t launches, xsb_lines *out)
{
	memset(out, 0, sizeof(*out));
	int64_t tables = 0;
	for (int r = 0; r < kRegCount; r++) {
		out->region_bytes[r] = region_lines[r] << kLineShift;
		tables += out->region_bytes[r];
	}
	out->bytes_tables = tables;
	out->bytes_schedule = schedule_bytes;
	out->bytes_small = small_bytes;
	out->bytes = tables + schedule_bytes + small_bytes;
	out->launches = launches;
}

static int64_t lines_of(int64_t bytes) { return ((bytes + 127) >> kLineShift) << kLineShift; }

// the small tables every lookup reads: skip table, material tables, FP32 constants
static int64_t small_table_bytes(const Device *d)
{
	const int64_t plane = (int64_t)kMaterials * d->T.max_num_nucs;
	return lines_of(8 * 256 * 16) + lines_of(kMaterials * 4) + lines_of(plane * 4) +
	       lines_of(plane * 8) + lines_of(plane * 4 * 4);
}

template <int GRID, int PREC>
static void launch_event_lines(Device *d, const LineMap &L, const xsb_event_args *a, bool wave)
{
	const unsigned nb = (unsigned)((a->lookups + 255) / 256);
	if (wave)
		hipLaunchKernelGGL((k_lines_event<GRID, PREC, true>), dim3(nb), dim3(256), 0, d->stream,
		                   d->T, L, a->first_lookup, a->lookups, a->seed);
	else
		hipLaunchKernelGGL((k_lines_event<GRID, PREC, false>), dim3(nb), dim3(256), 0, d->stream,
		                   d->T, L, a->first_lookup, a->lookups, a->seed);
}

template <int PREC>
static int dispatch_event_lines(Device *d, const LineMap &L, const xsb_event_args *a, bool wave)
{
	switch (d->T.grid_type) {
	case kUnionized: launch_event_lines<kUnionized, PREC>(d, L, a, wave); break;
	case kHash: launch_event_lines<kHash, PREC>(d, L, a, wave); break;
	default: return fail(XSB_ERR_UNSUPPORTED, "line counts cover the unionized and hash grids");
	}
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

static int check_counting(Device *d, int prec)
{
	if (d->T.grid_type == kNuclide)
		return fail(XSB_ERR_UNSUPPORTED, "line counts cover the unionized and hash grids");
	if ((prec == XSB_XS_FP32 && !d->d_xs32) || (prec == XSB_XS_BF16 && !d->d_xs16))
		return fail(XSB_ERR_UNSUPPORTED, "the precision's tables were not built");
	if (prec < 0 || prec > 2)
		return fail(XSB_ERR_INVALID, "xs_precision %d", prec);
	return XSB_OK;
}

int lines_event(Device *d, const xsb_event_args *a, xsb_lines *out)
{
	int rc = check_counting(d, a->xs_precision);
	if (rc)
		return rc;
	if (a->lookups < 0 || a->first_lookup < 0 || a->lookups > ((int64_t)1 << 32))
		return fail(XSB_ERR_INVALID, "lookups=%lld first=%lld", (long long)a->lookups,
		            (long long)a->first_lookup);
	int k = a->kernel_id == XSB_K_AUTO ? XSB_K_MAT_E_SORT : a->kernel_id;
	if (k == XSB_K_BASELINE && a->xs_precision != XSB_XS_FP64)
		k = XSB_K_SPLIT;
	const bool wave = k >= 2 && (d->T.variant & kVariantWave);
	LineMap L;
	const int64_t words = line_regions(d, &L);
	if ((rc = alloc_bits(d, words, &L.bits)))
		return rc;
	if (a->lookups > 0) {
		if (a->xs_precision == XSB_XS_FP64)
			rc = dispatch_event_lines<XSB_XS_FP64>(d, L, a, wave);
		else if (a->xs_precision == XSB_XS_FP32)
			rc = dispatch_event_lines<XSB_XS_FP32>(d, L, a, wave);
		else
			rc = dispatch_event_lines<XSB_XS_BF16>(d, L, a, wave);
	}
	int64_t reg[kRegCount] = {0};
	if (!rc)
		rc = count_lines(d, L, words, 1, reg);
	(void)hipFree(L.bits);
	if (rc)
		return rc;
	// what the lookup kernel reads besides the tables: the sorted pairs (k >= 2)
	// or the sample arrays (k = 1)
	const int64_t n = a->lookups;
	const int64_t sched = k >= 2 ? lines_of(n * 8) : (k == 1 ? lines_of(n * 8) + lines_of(n * 4) : 0);
	fill_result(reg, sched, small_table_bytes(d), 1, out);
	return XSB_OK;
}

int lines_batched(Device *d, const xsb_batched_args *a, xsb_lines *out)
{
	int rc = check_counting(d, a->xs_precision);
	if (rc)
		return rc;
	const int nb = a->n_batches;
	if (nb < 0 || (nb > 0 && (!a->sizes || !a->seeds || !a->thresholds)))
		return fail(XSB_ERR_INVALID, "batch arrays");
	std::vector<int64_t> off((size_t)nb + 1, 0);
	for (int b = 0; b < nb; b++) {
		if (a->sizes[b] < 0)
			return fail(XSB_ERR_INVALID, "sizes[%d] < 0", b);
		off[b + 1] = off[b] + a->sizes[b];
	}
	const int64_t total = off[nb];
	LineMap L;
	const int64_t words = line_regions(d, &L);
	if ((rc = alloc_bits(d, words, &L.bits)))
		return rc;
	void *tab = nullptr;
	const size_t tab_bytes = (size_t)(nb + 1) * 8 + (size_t)nb * 8 + (size_t)nb * kMaterials * 8;
	if (hipMalloc(&tab, tab_bytes + 64) != hipSuccess) {
		(void)hipGetLastError();
		(void)hipFree(L.bits);
		return fail(XSB_ERR_OOM, "batch tables for the line count");
	}
	int64_t *doff = (int64_t *)tab;
	uint64_t *dseed = (uint64_t *)(doff + nb + 1);
	double *dthr = (double *)(dseed + nb);
	hipStream_t s = d->stream;
	int64_t reg[kRegCount] = {0};
	if (hipMemcpyAsync(doff, off.data(), (size_t)(nb + 1) * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
	    (nb > 0 && (hipMemcpyAsync(dseed, a->seeds, (size_t)nb * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
	                hipMemcpyAsync(dthr, a->thresholds, (size_t)nb * kMaterials * 8,
	                               hipMemcpyHostToDevice, s) != hipSuccess))) {
		rc = fail(XSB_ERR_HIP, "batch table copy failed");
	} else if (total > 0) {
		const unsigned g = (unsigned)((total + 255) / 256);
		const int p = a->xs_precision;
		if (d->T.grid_type == kUnionized) {
			if (p == XSB_XS_FP64)
				hipLaunchKernelGGL((k_lines_batched<kUnionized, XSB_XS_FP64>), dim3(g), dim3(256), 0, s, d->T, L, doff, nb, dseed, dthr);
			else if (p == XSB_XS_FP32)
				hipLaunchKernelGGL((k_lines_batched<kUnionized, XSB_XS_FP32>), dim3(g), dim3(256), 0, s, d->T, L, doff, nb, dseed, dthr);
			else
				hipLaunchKernelGGL((k_lines_batched<kUnionized, XSB_XS_BF16>), dim3(g), dim3(256), 0, s, d->T, L, doff, nb, dseed, dthr);
		} else {
			if (p == XSB_XS_FP64)
				hipLaunchKernelGGL((k_lines_batched<kHash, XSB_XS_FP64>), dim3(g), dim3(256), 0, s, d->T, L, doff, nb, dseed, dthr);
			else if (p == XSB_XS_FP32)
				hipLaunchKernelGGL((k_lines_batched<kHash, XSB_XS_FP32>), dim3(g), dim3(256), 0, s, d->T, L, doff, nb, dseed, dthr);
			else
				hipLaunchKernelGGL((k_lines_batched<kHash, XSB_XS_BF16>), dim3(g), dim3(256), 0, s, d->T, L, doff, nb, dseed, dthr);
		}
		if (hipGetLastError() != hipSuccess)
			rc = fail(XSB_ERR_HIP, "line count launch failed");
	}
	if (!rc)
		rc = count_lines(d, L, words, 1, reg);
	(void)hipStreamSynchronize(s);
	(void)hipFree(tab);
	(void)hipFree(L.bits);
	if (rc)
		return rc;
	// the persistent kernel reads the sorted pairs and the batch tables, and
	// adds to per-batch sums and counts
	const int64_t sched = lines_of(total * 8) + lines_of((int64_t)nb * (8 + 8 + 4 + 4 + 8 * kMaterials)) +
	                      lines_of((int64_t)nb * 24);
	fill_result(reg, sched, small_table_bytes(d), 1, out);
	return XSB_OK;
}

int lines_history(Device *d, const xsb_history_args *a, xsb_lines *out)
{
	int rc = check_counting(d, a->xs_precision);
	if (rc)
		return rc;
	const int64_t P = a->particles;
	const int R = a->lookups_per_particle;
	if (P < 0 || P > INT32_MAX || R < 0 || R > 4096)
		return fail(XSB_ERR_INVALID, "particles=%lld lookups=%d", (long long)P, R);
	LineMap L;
	const int64_t words = line_regions(d, &L);
	const int sets = R > 0 ? R : 1;
	if ((rc = alloc_bits(d, words * sets, &L.bits)))
		return rc;
	int64_t reg[kRegCount] = {0};
	if (P > 0 && R > 0) {
		const unsigned g = (unsigned)((P + 255) / 256);
		hipStream_t s = d->stream;
		const int p = a->xs_precision;
#define XSB_HL(G, PR) hipLaunchKernelGGL((k_lines_history<G, PR>), dim3(g), dim3(256), 0, s, d->T, L, words, P, R, a->seed)
		if (d->T.grid_type == kUnionized) {
			if (p == XSB_XS_FP64) XSB_HL(kUnionized, XSB_XS_FP64);
			else if (p == XSB_XS_FP32) XSB_HL(kUnionized, XSB_XS_FP32);
			else XSB_HL(kUnionized, XSB_XS_BF16);
		} else {
			if (p == XSB_XS_FP64) XSB_HL(kHash, XSB_XS_FP64);
			else if (p == XSB_XS_FP32) XSB_HL(kHash, XSB_XS_FP32);
			else XSB_HL(kHash, XSB_XS_BF16);
		}
#undef XSB_HL
		if (hipGetLastError() != hipSuccess)
			rc = fail(XSB_ERR_HIP, "line count launch failed");
	}
	if (!rc)
		rc = count_lines(d, L, words, sets, reg);
	(void)hipFree(L.bits);
	if (rc)
		return rc;
	// per round the lookup kernel reads the sorted pairs and reads and writes
	// each particle's state (LCG state 8 bytes, sum 4 bytes)
	const int64_t sched = (int64_t)R * (lines_of(P * 8) + 2 * lines_of(P * 8) + 2 * lines_of(P * 4));
	fill_result(reg, sched, (int64_t)R * small_table_bytes(d), R, out);
	return XSB_OK;
}

}  // namespace xsb

using namespace xsb;

static int lines_handle(void *h, Device **d)
{
	if (!h)
		return fail(XSB_ERR_INVALID, "handle is NULL");
	*d = static_cast<Device *>(h);
	XSB_HIP(hipSetDevice((*d)->device));
	return XSB_OK;
}

extern "C" {

int xsb_hip_lines_event(void *handle, const xsb_event_args *args, xsb_lines *out)
{
	Device *d;
	int rc = lines_handle(handle, &d);
	if (rc)
		return rc;
	if (!args || !out)
		return fail(XSB_ERR_INVALID, "args and out must be non-NULL");
	return lines_event(d, args, out);
}

int xsb_hip_lines_history(void *handle, const xsb_history_args *args, xsb_lines *out)
{
	Device *d;
	int rc = lines_handle(handle, &d);
	if (rc)
		return rc;
	if (!args || !out)
		return fail(XSB_ERR_INVALID, "args and out must be non-NULL");
	return lines_history(d, args, out);
}

int xsb_hip_lines_batched(void *handle, const xsb_batched_args *args, xsb_lines *out)
{
	Device *d;
	int rc = lines_handle(handle, &d);
	if (rc)
		return rc;
	if (!args || !out)
		return fail(XSB_ERR_INVALID, "args and out must be non-NULL");
	return lines_batched(d, args, out);
}

}  // extern "C"
