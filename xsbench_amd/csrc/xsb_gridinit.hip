// xsb_gridinit.hip -- grid_init_do_not_profile on the device
// (openmp-threading/GridInit.c:3-151): the nuclide grid, the unionized
// energy array and its index grid, or the hash index grid, built in device
// memory and identical to what the reference builds on the host.
//
//   points   point g of the flattened grid takes LCG draws 6g .. 6g+5 of the
//            stream seeded with 42 (GridInit.c:12, :39-47), in field order;
//            each lane jumps to its draw with the skip-ahead table
//   sort     all points by energy (64-bit LSD radix sort of the energies'
//            bit patterns, which order like the non-negative doubles), then
//            stably by nuclide: each nuclide's points in ascending energy,
//            ties in generation order, as a stable sort of each nuclide's
//            segment gives (GridInit.c:50-51); the energy-sorted keys are
//            the unionized energy array (GridInit.c:84-89)
//   index    the reference's sweep (GridInit.c:102-119) split into row
//            chunks: a chunk restarts from the sweep's state at the end of
//            the run of equal energies before it, which is
//            min(#{k >= 1 : E_k <= U[e]}, n_gp - 2) (every nuclide energy is
//            itself in U, so the sweep has caught up by the end of each run),
//            then follows the reference's per-row rule; swept a few hundred
//            MB of rows at a time into the per-material tables, never whole
//   hash     index_grid[b][i] = grid_search_nuclide(b * du) (GridInit.c:128-151)
#include <vector>

#include "xsb_sorted.h"

namespace xsb {

constexpr uint64_t kGridSeed = 42;

__global__ void k_gen_points(int64_t n, const uint64_t *__restrict__ ff_tab, double *__restrict__ raw,
                             uint64_t *__restrict__ ekey, uint32_t *__restrict__ id)
{
	for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n;
	     g += (int64_t)gridDim.x * blockDim.x) {
		uint64_t s = ff_jump(kGridSeed, 6ULL * (uint64_t)g, ff_tab);
		double v[kPointDoubles];
#pragma unroll
		for (int w = 0; w < kPointDoubles; w++)
			v[w] = lcg_double(s);
		double *p = raw + g * kPointDoubles;
#pragma unroll
		for (int w = 0; w < kPointDoubles; w++)
			p[w] = v[w];
		ekey[g] = (uint64_t)__double_as_longlong(v[0]);
		id[g] = (uint32_t)g;
	}
}

// key = nuclide of the point (for the stable nuclide pass)
__global__ void k_nuclide_keys(int64_t n, int32_t n_gp, const uint32_t *__restrict__ id,
                               uint32_t *__restrict__ key)
{
	for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
	     j += (int64_t)gridDim.x * blockDim.x)
		key[j] = id[j] / (uint32_t)n_gp;
}

__global__ void k_gather_points(int64_t n, const uint32_t *__restrict__ id,
                                const double *__restrict__ raw, double *__restrict__ grid)
{
	for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
	     j += (int64_t)gridDim.x * blockDim.x) {
		const double *src = raw + (int64_t)id[j] * kPointDoubles;
		double *dst = grid + j * kPointDoubles;
#pragma unroll
		for (int w = 0; w < kPointDoubles; w++)
			dst[w] = src[w];
	}
}

__global__ void k_keys_to_energy(int64_t n, const uint64_t *__restrict__ key, double *__restrict__ ueg)
{
	for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
	     j += (int64_t)gridDim.x * blockDim.x)
		ueg[j] = __longlong_as_double((long long)key[j]);
}

// Index grid rows [row0 + c R, row0 + (c+1) R) within [row0, row1), for
// nuclides blockIdx.y * 256 + threadIdx.x; row e is written at
// index + (e - row0) * pitch.
__global__ __launch_bounds__(256) void k_index_sweep(const double *__restrict__ U, int64_t row0,
                                                     int64_t row1,
                                                     const double *__restrict__ grid, int32_t n_iso,
                                                     int32_t n_gp, int64_t R, int32_t pitch,
                                                     int32_t *__restrict__ index)
{
	const int i = blockIdx.y * 256 + threadIdx.x;
	if (i >= n_iso)
		return;
	const int64_t e0 = row0 + (int64_t)blockIdx.x * R;
	const int64_t e1 = min(e0 + R, row1);
	const double *E = grid + (int64_t)i * n_gp * kPointDoubles;  // energy k at E[6k]
	int64_t r = e0;
	while (r > 0 && U[r - 1] == U[r])
		r--;
	int32_t idx = 0;
	if (r > 0) {
		// #{k >= 1 : E_k <= U[r-1]}: first k in [1, n_gp) with E_k > x, minus 1
		const double x = U[r - 1];
		int32_t lo = 1, hi = n_gp;
		while (lo < hi) {
			const int32_t mid = lo + ((hi - lo) >> 1);
			if (E[(int64_t)mid * kPointDoubles] > x)
				hi = mid;
			else
				lo = mid + 1;
		}
		idx = min(lo - 1, n_gp - 2);
	}
	double e_high = E[(int64_t)(idx + 1) * kPointDoubles];
	for (int64_t e = r; e < e1; e++) {
		const double u = U[e];
		if (!(u < e_high) && idx != n_gp - 2) {
			idx++;
			e_high = E[(int64_t)(idx + 1) * kPointDoubles];
		}
		if (e >= e0)
			index[(e - row0) * pitch + i] = idx;
	}
}

// Rows [row0, row0 + rows) of the unionized index grid (GridInit.c:92-122)
// from the device's nuclide grid and unionized energies, into out with a
// row pitch of `pitch` ints.
int sweep_index_rows(Device *d, int64_t row0, int64_t rows, int32_t pitch, int32_t *out,
                     hipStream_t s)
{
	const int64_t R = 2048;
	const int32_t n_iso = d->T.n_iso;
	const dim3 grid((unsigned)((rows + R - 1) / R), (unsigned)((n_iso + 255) / 256));
	hipLaunchKernelGGL(k_index_sweep, grid, dim3(256), 0, s, d->d_ueg, row0, row0 + rows,
	                   d->d_grid, n_iso, d->T.n_gp, R, pitch, out);
	XSB_HIP(hipGetLastError());
	return XSB_OK;
}

__global__ void k_hash_grid(const double *__restrict__ grid, int32_t n_iso, int32_t n_gp,
                            int32_t bins, double du, int32_t pitch, int32_t *__restrict__ index)
{
	const int64_t n = (int64_t)bins * n_iso;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	     t += (int64_t)gridDim.x * blockDim.x) {
		const int64_t b = t / n_iso;
		const int i = (int)(t % n_iso);
		const double energy = (double)b * du;
		index[b * pitch + i] = grid_search_nuclide<kPointDoubles>(
		    energy, grid + (int64_t)i * n_gp * kPointDoubles, 0, n_gp - 1);
	}
}

// max |xs| over the points and channels of each nuclide (block per nuclide)
__global__ __launch_bounds__(256) void k_nuclide_xs_max(const double *__restrict__ grid, int32_t n_gp,
                                                        double *__restrict__ out)
{
	__shared__ double part[4];
	const double *p = grid + (int64_t)blockIdx.x * n_gp * kPointDoubles;
	double m = 0.0;
	for (int64_t k = threadIdx.x; k < n_gp; k += 256)
		for (int c = 1; c < kPointDoubles; c++)
			m = fmax(m, fabs(p[k * kPointDoubles + c]));
	for (int off = 32; off > 0; off >>= 1)
		m = fmax(m, __shfl_xor(m, off, 64));
	if ((threadIdx.x & 63) == 0)
		part[threadIdx.x >> 6] = m;
	__syncthreads();
	if (threadIdx.x == 0)
		out[blockIdx.x] = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
}

static int device_alloc(void **p, size_t bytes, const char *what)
{
	if (hipMalloc(p, bytes) != hipSuccess) {
		(void)hipGetLastError();
		*p = nullptr;
		return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes (%s) failed", bytes, what);
	}
	return XSB_OK;
}

int build_grids_on_device(Device *d, const xsb_tables *t, int64_t pitch, std::vector<double> *xs_max)
{
	hipStream_t s = d->stream;
	const int64_t n_iso = t->n_isotopes, n_gp = t->n_gridpoints, n = n_iso * n_gp;
	const unsigned nb = 4096;
	int rc;
	// work buffers: raw points, two key/id pairs, sort histogram, the
	// per-nuclide maxima
	const size_t raw_bytes = (size_t)n * kPointDoubles * 8;
	const size_t pair_bytes = (size_t)n * (8 + 4);
	const size_t hist_bytes = (size_t)sort_hist_len(n) * 4 + (size_t)sort_parts(n) * 4;
	const size_t max_bytes = (size_t)n_iso * 8;
	void *work = nullptr;
	// the eight buffers below, each rounded up to 256 bytes by take()
	if ((rc = device_alloc(&work, raw_bytes + 2 * pair_bytes + hist_bytes + max_bytes + 8 * 256,
	                       "grid build work")))
		return rc;
	char *w = (char *)work;
	auto take = [&w](size_t b) {
		char *p = w;
		w += (b + 255) & ~(size_t)255;
		return p;
	};
	double *raw = (double *)take(raw_bytes);
	uint64_t *kA = (uint64_t *)take((size_t)n * 8), *kB = (uint64_t *)take((size_t)n * 8);
	uint32_t *vA = (uint32_t *)take((size_t)n * 4), *vB = (uint32_t *)take((size_t)n * 4);
	uint32_t *hist = (uint32_t *)take((size_t)sort_hist_len(n) * 4);
	uint32_t *part = (uint32_t *)take((size_t)sort_parts(n) * 4);
	double *dmax = (double *)take(max_bytes);

	auto done = [&](int code) {
		(void)hipStreamSynchronize(s);
		(void)hipFree(work);
		return code;
	};
	hipLaunchKernelGGL(k_gen_points, dim3(nb), dim3(256), 0, s, n, d->T.ff_tab, raw, kA, vA);
	if (hipGetLastError() != hipSuccess)
		return done(fail(XSB_ERR_HIP, "point generation launch failed"));
	// energies in [0, 1): the bit patterns' top bit is clear, 63 bits sort
	if ((rc = sort_pairs64(s, n, (63 + kDigitBits - 1) / kDigitBits, &kA, &vA, &kB, &vB, hist, part)))
		return done(rc);
	if (t->grid_type == XSB_UNIONIZED) {
		if ((rc = device_alloc((void **)&d->d_ueg, (size_t)n * 8, "unionized energies")))
			return done(rc);
		d->bytes_union = n * 8;
		hipLaunchKernelGGL(k_keys_to_energy, dim3(nb), dim3(256), 0, s, n, kA, d->d_ueg);
	}
	// stable pass on the nuclide: u32 keys in the spare 64-bit key buffer
	uint32_t *nk = (uint32_t *)kB, *nk2 = nk + n;
	hipLaunchKernelGGL(k_nuclide_keys, dim3(nb), dim3(256), 0, s, n, (int32_t)n_gp, vA, nk);
	int npass = 1;
	while ((int64_t)1 << (kDigitBits * npass) < n_iso)
		npass++;
	if ((rc = sort_pairs(s, n, npass, false, &nk, &vA, &nk2, &vB, hist, part)))
		return done(rc);
	if ((rc = device_alloc((void **)&d->d_grid, raw_bytes, "nuclide grid")))
		return done(rc);
	d->bytes_fp64 = (int64_t)raw_bytes;
	hipLaunchKernelGGL(k_gather_points, dim3(nb), dim3(256), 0, s, n, vA, raw, d->d_grid);
	if (hipGetLastError() != hipSuccess)
		return done(fail(XSB_ERR_HIP, "nuclide grid gather launch failed"));

	// the unionized index grid is never held whole: the per-material tables
	// are built from swept row chunks (sweep_index_rows) after this returns
	if (t->grid_type == XSB_HASH) {
		const int64_t rows = t->hash_bins;
		const size_t bytes = (size_t)rows * pitch * sizeof(int32_t);
		if ((rc = device_alloc((void **)&d->d_index, bytes, "index grid")))
			return done(rc);
		d->bytes_index = (int64_t)bytes;
		if (hipMemsetAsync(d->d_index, 0, bytes, s) != hipSuccess)
			return done(fail(XSB_ERR_HIP, "index grid clear failed"));
		hipLaunchKernelGGL(k_hash_grid, dim3(nb), dim3(256), 0, s, d->d_grid, (int32_t)n_iso,
		                   (int32_t)n_gp, t->hash_bins, 1.0 / t->hash_bins, (int32_t)pitch,
		                   d->d_index);
		if (hipGetLastError() != hipSuccess)
			return done(fail(XSB_ERR_HIP, "index grid launch failed"));
	}
	// per-nuclide max |xs| for the compact tables' scales and bounds
	hipLaunchKernelGGL(k_nuclide_xs_max, dim3((unsigned)n_iso), dim3(256), 0, s, d->d_grid,
	                   (int32_t)n_gp, dmax);
	xs_max->assign((size_t)n_iso, 0.0);
	if (hipMemcpyAsync(xs_max->data(), dmax, (size_t)n_iso * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
		return done(fail(XSB_ERR_HIP, "copy of the nuclide maxima failed"));
	if (hipStreamSynchronize(s) != hipSuccess)
		return done(fail(XSB_ERR_HIP, "grid construction failed on the device"));
	return done(XSB_OK);
}

}  // namespace xsb
