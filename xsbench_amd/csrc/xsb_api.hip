// xsb_api.hip -- the extern "C" entry points of libxsb_hip.so (xsb_hip.h).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "xsb_internal.h"
#include "xsb_mixed.h"
#include "xsb_sorted.h"

namespace xsb {

static thread_local char g_msg[1024];

int fail(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
	va_end(ap);
	return code;
}

int scratch_reserve(Device *d, int slot, size_t bytes, void **out)
{
	Scratch &s = d->scratch[slot];
	if (s.bytes < bytes) {
		if (s.ptr) {
			XSB_HIP(hipStreamSynchronize(d->stream));
			XSB_HIP(hipFree(s.ptr));
			s.ptr = nullptr;
			s.bytes = 0;
		}
		const size_t want = (bytes + 0xFFFFF) & ~(size_t)0xFFFFF;
		if (hipMalloc(&s.ptr, want) != hipSuccess) {
			(void)hipGetLastError();
			return fail(XSB_ERR_OOM, "device scratch allocation of %zu bytes failed", want);
		}
		s.bytes = want;
	}
	*out = s.ptr;
	return XSB_OK;
}

// Volume fractions of the 12 materials (openmp-threading/Simulation.c:433-445)
// and pick_mat's thresholds: dist[i] + dist[i-1] + ... + dist[1], summed in
// that order (Simulation.c:450-457).
static void pick_thresholds(double thr[kMaterials])
{
	static const double dist[kMaterials] = {0.140, 0.052, 0.275, 0.134, 0.154, 0.064,
	                                        0.066, 0.055, 0.008, 0.015, 0.025, 0.013};
	for (int i = 0; i < kMaterials; i++) {
		double running = 0;
		for (int j = i; j > 0; j--)
			running += dist[j];
		thr[i] = running;
	}
}

// {A, C} of the LCG skip by d * 256^k for k < 8, d < 256, computed by the
// reference's square-and-multiply (openmp-threading/Simulation.c:472-499)
static void ff_table(uint64_t *tab)
{
	const uint64_t A = 2806196910506780709ULL;
	for (int k = 0; k < 8; k++) {
		for (int dgt = 0; dgt < 256; dgt++) {
			uint64_t n = (uint64_t)dgt << (8 * k);
			uint64_t a = A, c = 1, an = 1, cn = 0;
			while (n) {
				if (n & 1) {
					an *= a;
					cn = cn * a + c;
				}
				c *= (a + 1);
				a *= a;
				n >>= 1;
			}
			tab[2 * (256 * k + dgt)] = an;
			tab[2 * (256 * k + dgt) + 1] = cn;
		}
	}
}

template <typename T>
static int upload_array(T **dst, const T *src, int64_t count, int64_t *bytes_acc)
{
	*dst = nullptr;
	if (count <= 0)
		return XSB_OK;
	const size_t bytes = (size_t)count * sizeof(T);
	if (hipMalloc((void **)dst, bytes) != hipSuccess) {
		(void)hipGetLastError();
		*dst = nullptr;
		return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes failed", bytes);
	}
	XSB_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
	if (bytes_acc)
		*bytes_acc += (int64_t)bytes;
	return XSB_OK;
}

static void release(Device *d)
{
	if (!d)
		return;
	// best effort: a failure here leaves nothing the caller could act on
	(void)hipSetDevice(d->device);
	if (d->stream)
		(void)hipStreamSynchronize(d->stream);
	batch_state_free(d);
	if (d->history_graph.exec)
		(void)hipGraphExecDestroy(d->history_graph.exec);
	void *ptrs[] = {d->d_ueg,  d->d_index, d->d_hwin, d->d_esoa, d->d_midx, d->d_grid,  d->d_ivl,
	                d->d_ivl32, d->d_ueg_acc, d->d_xs32, d->d_xs16, d->d_xs16_scale, d->d_num_nucs,
	                d->d_mats, d->d_concs, d->d_concs32, d->d_acc, d->d_fftab};
	for (void *p : ptrs)
		if (p)
			(void)hipFree(p);
	for (auto &s : d->scratch)
		if (s.ptr)
			(void)hipFree(s.ptr);
	if (d->h_out)
		(void)hipHostFree(d->h_out);
	for (auto &e : d->ev)
		if (e)
			(void)hipEventDestroy(e);
	if (d->stream)
		(void)hipStreamDestroy(d->stream);
	(void)hipGetLastError();
	delete d;
}

// Device-side range check of the index grid: every entry must lie in
// [0, n_gp - 1] (grid construction never writes more than n_gp - 2).
__global__ void k_index_range(const int32_t *__restrict__ idx, int64_t n, int32_t hi,
                              unsigned long long *bad)
{
	uint32_t b = 0;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int32_t v = idx[i];
		b += (v < 0 || v > hi) ? 1u : 0u;
	}
	b = wave_sum_u32(b);
	if ((threadIdx.x & 63) == 0 && b)
		atomicAdd(bad, (unsigned long long)b);
}

// Unionized energies must be non-decreasing for the bisection to be exact.
__global__ void k_sorted_check(const double *__restrict__ a, int64_t n, unsigned long long *bad)
{
	uint32_t b = 0;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		b += (a[i] > a[i + 1]) ? 1u : 0u;
	b = wave_sum_u32(b);
	if ((threadIdx.x & 63) == 0 && b)
		atomicAdd(bad, (unsigned long long)b);
}

// XS values that are negative or NaN: with none (and no negative
// concentration) every mixed-precision term of an interpolating lookup is
// non-negative, which gives its channels relative error bounds.
__global__ void k_negative_xs(const double *__restrict__ grid, int64_t n, unsigned long long *bad)
{
	uint32_t b = 0;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		for (int c = 1; c < kPointDoubles; c++)
			b += (grid[i * kPointDoubles + c] >= 0.0) ? 0u : 1u;
	b = wave_sum_u32(b);
	if ((threadIdx.x & 63) == 0 && b)
		atomicAdd(bad, (unsigned long long)b);
}

// Descents (or NaN) inside the nuclides' energy rows (points [nuc][k], stride
// kPointDoubles): with none, a bracket search over a short range may count
// the energies <= p instead of bisecting (the same index on a sorted row).
__global__ void k_rows_sorted_check(const double *__restrict__ grid, int64_t n_iso, int64_t n_gp,
                                    unsigned long long *bad)
{
	uint32_t b = 0;
	const int64_t n = n_iso * n_gp;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n;
	     i += (int64_t)gridDim.x * blockDim.x)
		if ((i + 1) % n_gp != 0)
			b += (grid[i * kPointDoubles] <= grid[(i + 1) * kPointDoubles]) ? 0u : 1u;  // NaN too
	b = wave_sum_u32(b);
	if ((threadIdx.x & 63) == 0 && b)
		atomicAdd(bad, (unsigned long long)b);
}

static int validate_tables(const xsb_tables *t, uint32_t flags)
{
	const bool gen = (flags & XSB_BUILD_GRIDS) != 0;
	if (!t)
		return fail(XSB_ERR_INVALID, "tables is NULL");
	if (t->n_isotopes < 1 || t->n_gridpoints < 2)
		return fail(XSB_ERR_INVALID, "n_isotopes=%lld n_gridpoints=%lld out of range",
		            (long long)t->n_isotopes, (long long)t->n_gridpoints);
	if ((int64_t)t->n_isotopes * t->n_gridpoints > INT32_MAX)
		return fail(XSB_ERR_INVALID, "n_isotopes * n_gridpoints exceeds 2^31-1");
	if (t->grid_type < 0 || t->grid_type > 2)
		return fail(XSB_ERR_INVALID, "grid_type %d is not 0, 1 or 2", t->grid_type);
	if (!t->num_nucs || !t->mats || !t->concs || (!gen && !t->nuclide_grid))
		return fail(XSB_ERR_INVALID, "num_nucs, mats, concs and nuclide_grid are required");
	if (!gen && t->length_nuclide_grid != t->n_isotopes * t->n_gridpoints)
		return fail(XSB_ERR_INVALID, "length_nuclide_grid %lld != n_isotopes * n_gridpoints",
		            (long long)t->length_nuclide_grid);
	if (t->max_num_nucs < 1)
		return fail(XSB_ERR_INVALID, "max_num_nucs must be positive");
	for (int m = 0; m < kMaterials; m++) {
		const int nn = t->num_nucs[m];
		if (nn < 0 || nn > t->max_num_nucs)
			return fail(XSB_ERR_INVALID, "num_nucs[%d]=%d outside [0, max_num_nucs]", m, nn);
		for (int j = 0; j < nn; j++) {
			const int nuc = t->mats[m * t->max_num_nucs + j];
			if (nuc < 0 || nuc >= t->n_isotopes)
				return fail(XSB_ERR_INVALID, "mats[%d][%d]=%d is not a nuclide id", m, j, nuc);
		}
	}
	if (gen) {
		if (t->grid_type == XSB_HASH && t->hash_bins < 1)
			return fail(XSB_ERR_INVALID, "hash grid needs hash_bins >= 1");
		if (t->grid_type == XSB_UNIONIZED &&
		    (double)t->n_isotopes * t->n_gridpoints * t->n_isotopes > 9.0e18 / 4)
			return fail(XSB_ERR_INVALID, "unionized index grid too large");
	} else if (t->grid_type == XSB_UNIONIZED) {
		if (!t->unionized_energy_array || !t->index_grid)
			return fail(XSB_ERR_INVALID, "unionized grid needs unionized_energy_array and index_grid");
		if (t->length_unionized_energy_array != t->n_isotopes * t->n_gridpoints)
			return fail(XSB_ERR_INVALID, "unionized grid length %lld != n_isotopes * n_gridpoints",
			            (long long)t->length_unionized_energy_array);
		if (t->length_index_grid != t->length_unionized_energy_array * t->n_isotopes)
			return fail(XSB_ERR_INVALID, "index grid length %lld != n_union * n_isotopes",
			            (long long)t->length_index_grid);
	} else if (t->grid_type == XSB_HASH) {
		if (!t->index_grid || t->hash_bins < 1)
			return fail(XSB_ERR_INVALID, "hash grid needs index_grid and hash_bins >= 1");
		if (t->length_index_grid != (int64_t)t->hash_bins * t->n_isotopes)
			return fail(XSB_ERR_INVALID, "hash index grid length %lld != hash_bins * n_isotopes",
			            (long long)t->length_index_grid);
	}
	return XSB_OK;
}

// float rounded up (towards +inf) for bound constants
static float f32_up(double x)
{
	float f = (float)x;
	if ((double)f < x)
		f = std::nextafter(f, INFINITY);
	return f;
}

__global__ void k_build_fp32(const double *__restrict__ g, int64_t n, double *__restrict__ rec)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const double *p = g + i * kPointDoubles;
		float4 x = make_float4((float)p[1], (float)p[2], (float)p[3], (float)p[4]);
		double *r = rec + i * kRec32Doubles;
		r[0] = p[0];
		*reinterpret_cast<float4 *>(r + 1) = x;
		*reinterpret_cast<float2 *>(r + 3) = make_float2((float)p[5], 0.f);
	}
}

// round-to-nearest-even float -> bf16 (finite inputs)
__device__ __forceinline__ uint32_t f32_to_bf16(float f)
{
	uint32_t u = __float_as_uint(f);
	u += 0x7FFFu + ((u >> 16) & 1u);
	return u >> 16;
}

__global__ void k_build_bf16(const double *__restrict__ g, int64_t n, int32_t n_gp,
                             const float *__restrict__ inv_scale, double *__restrict__ rec)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const double *p = g + i * kPointDoubles;
		const double is = (double)inv_scale[i / n_gp];  // power of two: exact
		uint32_t h[5];
		for (int c = 0; c < 5; c++)
			h[c] = f32_to_bf16((float)(p[1 + c] * is));
		double *r = rec + i * kRec16Doubles;
		r[0] = p[0];
		*reinterpret_cast<uint2 *>(r + 1) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
		*reinterpret_cast<uint2 *>(r + 2) = make_uint2(h[4], 0u);
	}
}

// Hash grid search tables (Tables::esoa, Tables::hwin).
__global__ void k_energy_rows(const double *__restrict__ grid, int32_t n_iso, int32_t n_gp,
                              int32_t e_pitch, double *__restrict__ esoa)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)n_iso * e_pitch;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t nuc = i / e_pitch;
		const int32_t k = (int32_t)(i - nuc * e_pitch);
		esoa[i] = k < n_gp ? grid[(nuc * n_gp + k) * kPointDoubles] : (double)INFINITY;
	}
}

// Index entries are in [0, n_gp - 2] (checked before), so u_high <= n_gp - 1
// and the window E[b .. b + 9] lies inside the energy row's pitch.
__global__ void k_hash_windows(const int32_t *__restrict__ index, int64_t bins, int32_t n_iso,
                               int32_t pitch, int32_t n_gp, const double *__restrict__ esoa,
                               int32_t e_pitch, double *__restrict__ hwin)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < bins * pitch;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t b = i / pitch;
		const int32_t nuc = (int32_t)(i - b * pitch);
		double *r = hwin + i * kHwinDoubles;
		if (nuc >= n_iso) {
			for (int k = 0; k < kHwinDoubles; k++)
				r[k] = 0.0;
			continue;
		}
		const int32_t u_low = index[i];
		const int32_t u_high = (b == bins - 1) ? n_gp - 1 : index[i + pitch] + 1;
		r[0] = __hiloint2double(u_high, u_low);
		r[1] = 0.0;
		const double *E = esoa + (int64_t)nuc * e_pitch + (u_low & ~1);
		for (int k = 0; k < 10; k++)
			r[2 + k] = E[k];
	}
}

// Row pitch of Tables::esoa: room for an eighteen-energy window from any
// bin's u_low <= n_gp - 2, rounded down to even, and 64-byte aligned rows.
static int64_t energy_pitch(int64_t n_gp)
{
	return (n_gp + 16 + 7) & ~(int64_t)7;
}

static int build_hash_search(Device *d, int64_t n_iso, int64_t n_gp, int64_t bins, int64_t pitch)
{
	const int64_t e_pitch = energy_pitch(n_gp);
	const size_t wb = (size_t)bins * pitch * kHwinDoubles * 8, eb = (size_t)n_iso * e_pitch * 8;
	if (hipMalloc((void **)&d->d_esoa, eb) != hipSuccess) {
		(void)hipGetLastError();
		d->d_esoa = nullptr;
		return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes (hash energy rows) failed", eb);
	}
	if (hipMalloc((void **)&d->d_hwin, wb) != hipSuccess) {
		(void)hipGetLastError();
		d->d_hwin = nullptr;
		return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes (hash window records) failed", wb);
	}
	hipLaunchKernelGGL(k_energy_rows, dim3(4096), dim3(256), 0, d->stream, d->d_grid, (int32_t)n_iso,
	                   (int32_t)n_gp, (int32_t)e_pitch, d->d_esoa);
	XSB_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_hash_windows, dim3(4096), dim3(256), 0, d->stream, d->d_index, bins,
	                   (int32_t)n_iso, (int32_t)pitch, (int32_t)n_gp, d->d_esoa, (int32_t)e_pitch,
	                   d->d_hwin);
	XSB_HIP(hipGetLastError());
	d->bytes_index += (int64_t)(wb + eb);
	d->T.hwin = d->d_hwin;
	d->T.esoa = d->d_esoa;
	d->T.e_pitch = (int32_t)e_pitch;
	return XSB_OK;
}

// Per-nuclide max |xs| over the points and channels of host tables.
static int host_xs_max(const xsb_tables *t, std::vector<double> *mx)
{
	const int64_t n_iso = t->n_isotopes, n_gp = t->n_gridpoints;
	mx->assign((size_t)n_iso, 0.0);
	for (int64_t nuc = 0; nuc < n_iso; nuc++) {
		double m = 0.0;
		const double *p = t->nuclide_grid + nuc * n_gp * kPointDoubles;
		for (int64_t k = 0; k < n_gp; k++)
			for (int c = 1; c < kPointDoubles; c++)
				m = std::fmax(m, std::fabs(p[k * kPointDoubles + c]));
		(*mx)[nuc] = m;
	}
	return XSB_OK;
}

static int build_compact(Device *d, const xsb_tables *t, uint32_t flags, const std::vector<double> &mx)
{
	const int64_t n_iso = t->n_isotopes, n_gp = t->n_gridpoints, n = n_iso * n_gp;
	const int max_n = t->max_num_nucs;
	const int plane = kMaterials * max_n;
	for (int64_t nuc = 0; nuc < n_iso; nuc++)
		if (!std::isfinite(mx[nuc]))
			return fail(XSB_ERR_INVALID, "nuclide %lld holds non-finite XS", (long long)nuc);
	// power-of-two scale per nuclide with max|xs| <= scale (BF16 table)
	std::vector<float> scale(n_iso), inv_scale(n_iso);
	for (int64_t nuc = 0; nuc < n_iso; nuc++) {
		int e = 0;
		if (mx[nuc] > 0) {
			std::frexp(mx[nuc], &e);  // mx = f * 2^e, f in [0.5, 1)
		}
		scale[nuc] = std::ldexp(1.0f, e);
		inv_scale[nuc] = std::ldexp(1.0f, -e);
	}
	// constant planes: FP32 c_eff, FP32 c_bound, BF16 c_eff, BF16 c_bound
	std::vector<float> planes(4 * (size_t)plane, 0.f);
	for (int m = 0; m < kMaterials; m++) {
		for (int j = 0; j < t->num_nucs[m]; j++) {
			const int nuc = t->mats[m * max_n + j];
			const double conc = t->concs[m * max_n + j];
			const size_t o = (size_t)m * max_n + j;
			planes[o] = (float)conc;
			planes[plane + o] = f32_up(std::fabs(conc) * mx[nuc] * (1.0 + 0x1p-20));
			planes[2 * plane + o] = (float)(conc * (double)scale[nuc]);
			planes[3 * plane + o] = f32_up(std::fabs(conc) * (double)scale[nuc] * (1.0 + 0x1p-7));
		}
	}
	int rc = upload_array(&d->d_concs32, planes.data(), (int64_t)planes.size(), nullptr);
	if (rc)
		return rc;
	const unsigned nb = 2048;
	if (flags & XSB_BUILD_FP32) {
		const size_t bytes = (size_t)n * kRec32Doubles * sizeof(double);
		if (hipMalloc((void **)&d->d_xs32, bytes) != hipSuccess) {
			(void)hipGetLastError();
			d->d_xs32 = nullptr;
			return fail(XSB_ERR_OOM, "FP32 table allocation (%zu bytes) failed", bytes);
		}
		hipLaunchKernelGGL(k_build_fp32, dim3(nb), dim3(256), 0, d->stream, d->d_grid, n,
		                   (double *)d->d_xs32);
		XSB_HIP(hipGetLastError());
		d->bytes_fp32 = (int64_t)bytes;
	}
	if (flags & XSB_BUILD_BF16) {
		const size_t bytes = (size_t)n * kRec16Doubles * sizeof(double);
		if (hipMalloc((void **)&d->d_xs16, bytes) != hipSuccess) {
			(void)hipGetLastError();
			d->d_xs16 = nullptr;
			return fail(XSB_ERR_OOM, "BF16 table allocation (%zu bytes) failed", bytes);
		}
		rc = upload_array(&d->d_xs16_scale, inv_scale.data(), n_iso, nullptr);
		if (rc)
			return rc;
		hipLaunchKernelGGL(k_build_bf16, dim3(nb), dim3(256), 0, d->stream, d->d_grid, n,
		                   (int32_t)n_gp, d->d_xs16_scale, (double *)d->d_xs16);
		XSB_HIP(hipGetLastError());
		d->bytes_bf16 = (int64_t)bytes;
	}
	XSB_HIP(hipStreamSynchronize(d->stream));
	return XSB_OK;
}

// rows of n ints -> rows of `pitch` ints (padding already zeroed)
__global__ void k_repitch(const int32_t *__restrict__ src, int64_t rows, int32_t n, int32_t pitch,
                          int32_t *__restrict__ dst)
{
	const int64_t total = rows * n;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t r = i / n;
		dst[r * pitch + (i - r * n)] = src[i];
	}
}

// Index-grid rows staged at a time (host copies, device sweeps).
constexpr int64_t kIndexChunkBytes = (int64_t)64 << 20;

// Copies a rows x n int grid into the device's pitched layout through a
// bounded staging buffer (contiguous host-to-device copies, re-pitched on
// the device).
static int upload_pitched(Device *d, const int32_t *host, int64_t rows, int64_t n, int64_t pitch)
{
	const int64_t stage_rows = std::max<int64_t>(1, kIndexChunkBytes / (n * 4));
	void *stage = nullptr;
	int rc = scratch_reserve(d, 0, (size_t)std::min(rows, stage_rows) * n * 4, &stage);
	if (rc)
		return rc;
	for (int64_t r0 = 0; r0 < rows; r0 += stage_rows) {
		const int64_t nr = std::min(stage_rows, rows - r0);
		XSB_HIP(hipMemcpy(stage, host + r0 * n, (size_t)nr * n * 4, hipMemcpyHostToDevice));
		hipLaunchKernelGGL(k_repitch, dim3(2048), dim3(256), 0, nullptr, (const int32_t *)stage, nr,
		                   (int32_t)n, (int32_t)pitch, d->d_index + r0 * pitch);
		XSB_HIP(hipGetLastError());
		XSB_HIP(hipDeviceSynchronize());
	}
	return XSB_OK;
}

// rows of `pitch` ints -> rows of n ints
__global__ void k_unpitch(const int32_t *__restrict__ src, int64_t rows, int32_t n, int32_t pitch,
                          int32_t *__restrict__ dst)
{
	const int64_t total = rows * n;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t r = i / n;
		dst[i] = src[r * pitch + (i - r * n)];
	}
}

static int download(Device *d, double *grid, double *ueg, int32_t *index)
{
	const Tables &T = d->T;
	const int64_t n = (int64_t)T.n_iso * T.n_gp;
	if (grid)
		XSB_HIP(hipMemcpy(grid, d->d_grid, (size_t)n * kPointDoubles * 8, hipMemcpyDeviceToHost));
	if (ueg && d->d_ueg)
		XSB_HIP(hipMemcpy(ueg, d->d_ueg, (size_t)T.n_union * 8, hipMemcpyDeviceToHost));
	// the hash grid's index grid is resident; the unionized one is swept again
	// from the device's nuclide grid and unionized energies (GridInit.c:92-122)
	if (index && (d->d_index || T.grid_type == kUnionized)) {
		const int64_t rows = (T.grid_type == kUnionized) ? T.n_union : T.hash_bins;
		const int64_t stage_rows = std::max<int64_t>(1, kIndexChunkBytes / ((int64_t)T.n_iso * 4));
		void *stage = nullptr;
		int rc = scratch_reserve(d, 0, (size_t)std::min(rows, stage_rows) * T.n_iso * 4, &stage);
		if (rc)
			return rc;
		for (int64_t r0 = 0; r0 < rows; r0 += stage_rows) {
			const int64_t nr = std::min(stage_rows, rows - r0);
			if (T.grid_type == kUnionized) {
				if ((rc = sweep_index_rows(d, r0, nr, T.n_iso, (int32_t *)stage, d->stream)))
					return rc;
			} else {
				hipLaunchKernelGGL(k_unpitch, dim3(2048), dim3(256), 0, d->stream,
				                   d->d_index + r0 * T.idx_pitch, nr, T.n_iso, T.idx_pitch,
				                   (int32_t *)stage);
			}
			XSB_HIP(hipGetLastError());
			XSB_HIP(hipStreamSynchronize(d->stream));
			XSB_HIP(hipMemcpy(index + r0 * T.n_iso, stage, (size_t)nr * T.n_iso * 4,
			                  hipMemcpyDeviceToHost));
		}
	}
	return XSB_OK;
}

// Material m's unionized index table: row e holds index_grid[e][mats[m][j]]
// for j < nn, then zeros up to the pitch mp; IDX is the element type.  The
// brackets are stored with the lookup's lower == n_gp - 1 -> n_gp - 2 rule
// (Simulation.c:246-247) already applied.
template <class IDX>
__global__ void k_material_index(const int32_t *__restrict__ index, int64_t rows, int32_t pitch,
                                 const int32_t *__restrict__ ml, int32_t nn, int32_t mp,
                                 int32_t n_gp, IDX *__restrict__ dst)
{
	const int64_t total = rows * mp;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
	     t += (int64_t)gridDim.x * blockDim.x) {
		const int64_t e = t / mp;
		const int j = (int)(t - e * mp);
		int32_t v = (j < nn) ? index[e * pitch + ml[j]] : 0;
		if (v == n_gp - 1)
			v--;
		dst[t] = (IDX)v;
	}
}

// Interval records (Tables::ivl): record i = nuc * n_gp + k holds point k+1
// and then point k+1 - point k, channel by channel; the last record of each
// nuclide (k = n_gp - 1) is zero and never read.
__global__ void k_interval_records(const double *__restrict__ grid, int64_t total, int32_t n_gp,
                                   double *__restrict__ ivl)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t nuc = i / n_gp;
		const bool last = (i - nuc * n_gp) == n_gp - 1;
		const double *lo = grid + i * kPointDoubles;
		double *r = ivl + i * kIvlDoubles;
#pragma unroll
		for (int c = 0; c < kPointDoubles; c++) {
			const double hi = last ? 0.0 : lo[kPointDoubles + c];
			r[c] = hi;
			r[kPointDoubles + c] = last ? 0.0 : hi - lo[c];
		}
	}
}

// FP32 interval records (Tables::ivl32): {E_{k+1}, 1 / (E_{k+1} - E_k),
// float(xs_{k+1})[5], float(xs_k)[5], pad}; zero for the last point of each
// nuclide.  A reciprocal that is not a normal number (equal energies, or a
// width above 2^1022) is stored as NaN: the lookup's error bound is then NaN
// and the lookup is evaluated on the FP64 path.
__global__ void k_interval_records32(const double *__restrict__ grid, int64_t total, int32_t n_gp,
                                     double *__restrict__ ivl)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
	     i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t nuc = i / n_gp;
		const bool last = (i - nuc * n_gp) == n_gp - 1;
		const double *lo = grid + i * kPointDoubles;
		double *r = ivl + i * kIvl32Doubles;
		float *f = reinterpret_cast<float *>(r + 2);
		r[0] = last ? 0.0 : lo[kPointDoubles];
		const double rw = 1.0 / (lo[kPointDoubles] - lo[0]);
		r[1] = last ? 0.0 : (__builtin_isnormal(rw) ? rw : __builtin_nan(""));
#pragma unroll
		for (int c = 0; c < 5; c++) {
			f[c] = last ? 0.f : (float)lo[kPointDoubles + 1 + c];
			f[5 + c] = last ? 0.f : (float)lo[1 + c];
		}
		f[10] = 0.f;
		f[11] = 0.f;
	}
}

// Tables::ueg_acc: acc[b] = #{u : ueg[u] <= b / 2^16} (upper_bound by
// bisection), b = 0..2^16.
__global__ void k_ueg_acc(const double *__restrict__ ueg, int32_t n, int32_t *__restrict__ acc)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b > (1 << kUegAccBits))
		return;
	const double q = (double)b / (double)(1 << kUegAccBits);
	int32_t lo = 0, hi = n;  // first index with ueg > q lies in [lo, hi]
	while (lo < hi) {
		const int32_t mid = lo + ((hi - lo) >> 1);
		if (ueg[mid] > q)
			hi = mid;
		else
			lo = mid + 1;
	}
	acc[b] = lo;
}

static int build_ueg_acc(Device *d)
{
	const int nb = (1 << kUegAccBits) + 1;
	if (hipMalloc((void **)&d->d_ueg_acc, (size_t)nb * sizeof(int32_t)) != hipSuccess) {
		(void)hipGetLastError();
		d->d_ueg_acc = nullptr;
		return fail(XSB_ERR_OOM, "hipMalloc of the unionized search brackets failed");
	}
	hipLaunchKernelGGL(k_ueg_acc, dim3((nb + 255) / 256), dim3(256), 0, d->stream, d->d_ueg,
	                   (int32_t)d->T.n_union, d->d_ueg_acc);
	XSB_HIP(hipGetLastError());
	XSB_HIP(hipStreamSynchronize(d->stream));
	d->T.ueg_acc = d->d_ueg_acc;
	return XSB_OK;
}

static int build_interval_records(Device *d)
{
	Tables &T = d->T;
	const int64_t total = (int64_t)T.n_iso * T.n_gp;
	const size_t bytes = (size_t)total * kIvlDoubles * sizeof(double);
	if (hipMalloc((void **)&d->d_ivl, bytes) != hipSuccess) {
		(void)hipGetLastError();
		d->d_ivl = nullptr;
		return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes (interval records) failed", bytes);
	}
	d->bytes_ivl = (int64_t)bytes;
	hipLaunchKernelGGL(k_interval_records, dim3(4096), dim3(256), 0, d->stream, d->d_grid, total,
	                   T.n_gp, d->d_ivl);
	XSB_HIP(hipGetLastError());
	if (d->d_xs32) {
		const size_t bytes32 = (size_t)total * kIvl32Doubles * sizeof(double);
		if (hipMalloc((void **)&d->d_ivl32, bytes32) != hipSuccess) {
			(void)hipGetLastError();
			d->d_ivl32 = nullptr;
			return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes (FP32 interval records) failed",
			            bytes32);
		}
		d->bytes_ivl += (int64_t)bytes32;
		hipLaunchKernelGGL(k_interval_records32, dim3(4096), dim3(256), 0, d->stream, d->d_grid,
		                   total, T.n_gp, d->d_ivl32);
		XSB_HIP(hipGetLastError());
	}
	XSB_HIP(hipStreamSynchronize(d->stream));
	T.ivl = d->d_ivl;
	T.ivl32 = d->d_ivl32;
	return XSB_OK;
}

// Per-material copies of the unionized index grid (Tables::midx): each
// lookup then reads its material's brackets from one contiguous row.
// Brackets are at most n_gp - 1, so they are stored as uint16 when
// n_gp <= 65536 unless `wide` asks for int32.  The full grid (n_union rows of
// n_iso ints, 5.7 GB for the large problem) is never held on the device: rows
// come a chunk at a time, copied from the host's index grid (and range-checked)
// or, for grids built on the device, swept from its nuclide grid.
static int build_material_index(Device *d, const xsb_tables *t, bool from_host, bool wide)
{
	Tables &T = d->T;
	const int64_t rows = T.n_union, n_iso = T.n_iso;
	const bool narrow = !wide && T.n_gp <= 65536;
	T.midx_shift = narrow ? 1 : 2;
	const int align = narrow ? 8 : 4;  // entries per 16 bytes
	int64_t total = 0;
	for (int m = 0; m < kMaterials; m++) {
		T.midx_off[m] = total;
		T.midx_pitch[m] = (t->num_nucs[m] + align - 1) & ~(align - 1);
		total += rows * T.midx_pitch[m];
	}
	const size_t bytes = (size_t)total << T.midx_shift;
	if (hipMalloc((void **)&d->d_midx, bytes) != hipSuccess) {
		(void)hipGetLastError();
		d->d_midx = nullptr;
		return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes (material index tables) failed", bytes);
	}
	d->bytes_midx = (int64_t)bytes;
	// device sweeps take larger chunks: a sweep block walks 2048 rows, so a
	// chunk of 64 MB would leave most of the device idle
	const int64_t chunk_bytes = from_host ? kIndexChunkBytes : 8 * kIndexChunkBytes;
	const int64_t chunk = std::max<int64_t>(1, chunk_bytes / (n_iso * 4));
	void *stage = nullptr;
	int rc = scratch_reserve(d, 0, (size_t)std::min(rows, chunk) * n_iso * 4, &stage);
	if (rc)
		return rc;
	int32_t *rows_buf = (int32_t *)stage;
	hipStream_t s = d->stream;
	unsigned long long *bad = d->d_acc + 60;
	XSB_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned long long), s));
	for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
		const int64_t nr = std::min(chunk, rows - r0);
		if (from_host) {
			// the staging buffer is reused by the next chunk: the copy waits
			// for the previous chunk's kernels (same stream, synchronous copy)
			XSB_HIP(hipStreamSynchronize(s));
			XSB_HIP(hipMemcpy(rows_buf, t->index_grid + r0 * n_iso, (size_t)nr * n_iso * 4,
			                  hipMemcpyHostToDevice));
			hipLaunchKernelGGL(k_index_range, dim3(2048), dim3(256), 0, s, rows_buf, nr * n_iso,
			                   (int32_t)(T.n_gp - 1), bad);
			XSB_HIP(hipGetLastError());
		} else if ((rc = sweep_index_rows(d, r0, nr, (int32_t)n_iso, rows_buf, s))) {
			return rc;
		}
		for (int m = 0; m < kMaterials; m++) {
			if (!T.midx_pitch[m])
				continue;
			const int32_t *ml = d->d_mats + (int64_t)m * t->max_num_nucs;
			const int64_t off = T.midx_off[m] + r0 * T.midx_pitch[m];
			if (narrow)
				hipLaunchKernelGGL(k_material_index<uint16_t>, dim3(2048), dim3(256), 0, s,
				                   rows_buf, nr, (int32_t)n_iso, ml, t->num_nucs[m],
				                   T.midx_pitch[m], T.n_gp, (uint16_t *)d->d_midx + off);
			else
				hipLaunchKernelGGL(k_material_index<int32_t>, dim3(2048), dim3(256), 0, s,
				                   rows_buf, nr, (int32_t)n_iso, ml, t->num_nucs[m],
				                   T.midx_pitch[m], T.n_gp, (int32_t *)d->d_midx + off);
			XSB_HIP(hipGetLastError());
		}
	}
	unsigned long long nbad = 0;
	XSB_HIP(hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	if (nbad)
		return fail(XSB_ERR_INVALID, "index grid holds %llu entries outside [0, %d]", nbad,
		            T.n_gp - 1);
	T.midx = d->d_midx;
	return XSB_OK;
}

// ---- the allocation-free event entry point's work memory ----
// One fixed array per GPU (each device holds its own copy of a __device__
// variable), large enough for the sorted schedules' buffers at
// XSB_NOALLOC_MAX_LOOKUPS lookups; held until the process ends.  In host code
// `scratch` only stands in for the symbol: its device address comes from
// hipGetSymbolAddress on each device, and only that address is used.
constexpr size_t kNoallocBytes = (sorted_event_bytes(XSB_NOALLOC_MAX_LOOKUPS) + 255) & ~(size_t)255;
static_assert(kNoallocBytes >= (size_t)XSB_NOALLOC_MAX_LOOKUPS * 12 + 256,
              "the split schedule's sample arrays fit too");
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_lock;
static void *g_noalloc_ptr[kMaxDevices];
static int g_noalloc_rc[kMaxDevices];  // 0: not looked up yet, else an XSB_* code + 1

// Looks up `scratch`'s address on `device` (current) once; called by
// xsb_hip_upload, so the entry point itself makes no runtime call for it.
static int noalloc_setup(int device)
{
	if (device < 0 || device >= kMaxDevices)
		return fail(XSB_ERR_UNSUPPORTED, "device %d beyond the allocation-free table", device);
	std::lock_guard<std::mutex> g(g_noalloc_lock);
	if (g_noalloc_rc[device] == 0) {
		void *ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr) {
			(void)hipGetLastError();
			g_noalloc_rc[device] = XSB_ERR_HIP + 1;
			return fail(XSB_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s",
			            device, hipGetErrorString(e));
		}
		g_noalloc_ptr[device] = ptr;
		g_noalloc_rc[device] = XSB_OK + 1;
	}
	return g_noalloc_rc[device] - 1;
}

// Bytes of each resident table an upload of t with `flags` allocates (the
// same sizes the upload records in xsb_info), computed on the host.
static void table_bytes(const xsb_tables *t, uint32_t flags, xsb_info *o)
{
	memset(o, 0, sizeof(*o));
	const int64_t n_iso = t->n_isotopes, n_gp = t->n_gridpoints, n = n_iso * n_gp;
	const bool uni = t->grid_type == XSB_UNIONIZED;
	o->grid_type = t->grid_type;
	o->n_isotopes = n_iso;
	o->n_gridpoints = n_gp;
	o->n_union = uni ? n : 0;
	o->bytes_fp64 = n * kPointDoubles * 8;
	o->bytes_fp32 = (flags & XSB_BUILD_FP32) ? n * kRec32Doubles * 8 : 0;
	o->bytes_bf16 = (flags & XSB_BUILD_BF16) ? n * kRec16Doubles * 8 : 0;
	o->bytes_union = uni ? n * 8 : 0;
	// hash grid: the index grid (int32, rows padded to 16 bytes), its window
	// records (same rows) and the energy rows (Tables::hwin, ::esoa)
	o->bytes_index = (t->grid_type == XSB_HASH) ? (int64_t)t->hash_bins * ((n_iso + 3) & ~3) *
	                                                      (4 + kHwinDoubles * 8) +
	                                                  n_iso * energy_pitch(n_gp) * 8
	                                            : 0;
	if (uni) {
		const bool narrow = !(flags & XSB_BUILD_IDX32) && n_gp <= 65536;
		const int align = narrow ? 8 : 4, width = narrow ? 2 : 4;
		for (int m = 0; m < kMaterials; m++)
			o->bytes_midx += n * (int64_t)((t->num_nucs[m] + align - 1) & ~(align - 1)) * width;
		o->bytes_ivl = n * kIvlDoubles * 8 + ((flags & XSB_BUILD_FP32) ? n * kIvl32Doubles * 8 : 0);
	}
}

static int upload(int device, const xsb_tables *t, uint32_t flags, Device **out)
{
	int rc = validate_tables(t, flags);
	if (rc)
		return rc;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
		(void)hipGetLastError();
		return fail(XSB_ERR_NO_DEVICE, "no HIP device is available");
	}
	if (device < 0 || device >= ndev)
		return fail(XSB_ERR_NO_DEVICE, "device %d not in [0, %d)", device, ndev);
	XSB_HIP(hipSetDevice(device));
	Device *d = new Device();
	*out = d;
	d->device = device;
	if ((rc = noalloc_setup(device)))
		return rc;
	hipDeviceProp_t prop;
	XSB_HIP(hipGetDeviceProperties(&prop, device));
	d->cu_count = prop.multiProcessorCount;
	snprintf(d->arch, sizeof(d->arch), "%s", prop.gcnArchName);
	XSB_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
	for (auto &e : d->ev)
		XSB_HIP(hipEventCreate(&e));

	const int64_t n_iso = t->n_isotopes, n_gp = t->n_gridpoints;
	const int64_t pitch = (n_iso + 3) & ~(int64_t)3;  // 16-byte aligned rows
	const bool gen = (flags & XSB_BUILD_GRIDS) != 0;
	const int64_t n_union = (t->grid_type == XSB_UNIONIZED) ? n_iso * n_gp : 0;
	const int64_t index_rows = (t->grid_type == XSB_UNIONIZED) ? n_union
	                           : (t->grid_type == XSB_HASH)    ? t->hash_bins
	                                                           : 0;
	{
		std::vector<uint64_t> tab(8 * 256 * 2);
		ff_table(tab.data());
		if ((rc = upload_array(&d->d_fftab, tab.data(), (int64_t)tab.size(), nullptr)))
			return rc;
		d->T.ff_tab = d->d_fftab;
	}
	if ((rc = upload_array(&d->d_num_nucs, t->num_nucs, kMaterials, nullptr)))
		return rc;
	if ((rc = upload_array(&d->d_mats, t->mats, (int64_t)kMaterials * t->max_num_nucs, nullptr)))
		return rc;
	if ((rc = upload_array(&d->d_concs, t->concs, (int64_t)kMaterials * t->max_num_nucs, nullptr)))
		return rc;
	std::vector<double> xs_max;
	if (gen) {
		if ((rc = build_grids_on_device(d, t, pitch, &xs_max)))
			return rc;
	} else {
		if ((rc = upload_array(&d->d_grid, t->nuclide_grid, n_iso * n_gp * kPointDoubles,
		                       &d->bytes_fp64)))
			return rc;
		if (t->grid_type == XSB_UNIONIZED) {
			if ((rc = upload_array(&d->d_ueg, t->unionized_energy_array, n_union, &d->bytes_union)))
				return rc;
		}
		if (t->grid_type == XSB_HASH) {
			const size_t bytes = (size_t)index_rows * pitch * sizeof(int32_t);
			if (hipMalloc((void **)&d->d_index, bytes) != hipSuccess) {
				(void)hipGetLastError();
				d->d_index = nullptr;
				return fail(XSB_ERR_OOM, "hipMalloc of %zu bytes (index grid) failed", bytes);
			}
			XSB_HIP(hipMemset(d->d_index, 0, bytes));
			if ((rc = upload_pitched(d, t->index_grid, index_rows, n_iso, pitch)))
				return rc;
			d->bytes_index = (int64_t)bytes;
		}
		if ((rc = host_xs_max(t, &xs_max)))
			return rc;
	}
	if (hipMalloc((void **)&d->d_acc, 64 * sizeof(unsigned long long)) != hipSuccess) {
		(void)hipGetLastError();
		d->d_acc = nullptr;
		return fail(XSB_ERR_OOM, "accumulator allocation failed");
	}
	XSB_HIP(hipMemset(d->d_acc, 0, 64 * sizeof(unsigned long long)));
	if (hipHostMalloc((void **)&d->h_out, 2 * sizeof(unsigned long long) + 4 * 1024,
	                  hipHostMallocDefault) != hipSuccess) {
		(void)hipGetLastError();
		d->h_out = nullptr;
		return fail(XSB_ERR_OOM, "pinned host result buffer allocation failed");
	}

	// integrity checks that protect the kernels' memory accesses
	unsigned long long bad = 0;
	if (d->d_index) {
		const int32_t hi = (t->grid_type == XSB_HASH) ? (int32_t)(n_gp - 2) : (int32_t)(n_gp - 1);
		hipLaunchKernelGGL(k_index_range, dim3(4096), dim3(256), 0, d->stream, d->d_index,
		                   index_rows * pitch, hi, d->d_acc + 60);
		XSB_HIP(hipGetLastError());
	}
	if (d->d_ueg) {
		hipLaunchKernelGGL(k_sorted_check, dim3(2048), dim3(256), 0, d->stream, d->d_ueg, n_union,
		                   d->d_acc + 61);
		XSB_HIP(hipGetLastError());
	}
	hipLaunchKernelGGL(k_rows_sorted_check, dim3(2048), dim3(256), 0, d->stream, d->d_grid,
	                   n_iso, n_gp, d->d_acc + 62);
	XSB_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_negative_xs, dim3(2048), dim3(256), 0, d->stream, d->d_grid, n_iso * n_gp,
	                   d->d_acc + 63);
	XSB_HIP(hipGetLastError());
	unsigned long long chk[4] = {0, 0, 0, 0};
	XSB_HIP(hipMemcpyAsync(chk, d->d_acc + 60, sizeof(chk), hipMemcpyDeviceToHost, d->stream));
	XSB_HIP(hipStreamSynchronize(d->stream));
	bad = chk[0];
	if (bad)
		return fail(XSB_ERR_INVALID, "index grid holds %llu entries outside [0, %lld]", bad,
		            (long long)(t->grid_type == XSB_HASH ? n_gp - 2 : n_gp - 1));
	if (chk[1])
		return fail(XSB_ERR_INVALID, "unionized energy array is not sorted (%llu descents)", chk[1]);

	Tables &T = d->T;
	T.ueg = d->d_ueg;
	T.index_grid = d->d_index;
	T.grid = d->d_grid;
	T.num_nucs = d->d_num_nucs;
	T.mats = d->d_mats;
	T.concs = d->d_concs;
	T.n_union = n_union;
	T.n_iso = (int32_t)n_iso;
	T.idx_pitch = (int32_t)pitch;
	bool conc_nonneg = true;
	for (int m = 0; m < kMaterials; m++)
		for (int j = 0; j < t->num_nucs[m]; j++)
			conc_nonneg = conc_nonneg && t->concs[m * t->max_num_nucs + j] >= 0.0;
	T.variant = kVariantDefault | (chk[2] == 0 ? kVariantSortedRows : 0) |
	            (chk[3] == 0 && conc_nonneg ? kVariantNonNeg : 0);
	T.n_gp = (int32_t)n_gp;
	T.hash_bins = t->hash_bins;
	T.max_num_nucs = t->max_num_nucs;
	T.grid_type = t->grid_type;
	if (t->grid_type == XSB_UNIONIZED &&
	    (rc = build_material_index(d, t, !gen, (flags & XSB_BUILD_IDX32) != 0)))
		return rc;
	if (t->grid_type == XSB_HASH && (rc = build_hash_search(d, n_iso, n_gp, index_rows, pitch)))
		return rc;
	T.hash_du = (t->hash_bins > 0) ? 1.0 / t->hash_bins : 0.0;
	pick_thresholds(T.thr);
	for (int i = 1; i < kMaterials; i++)
		if (!(T.thr[i] > T.thr[i - 1]))
			return fail(XSB_ERR_INVALID, "material thresholds are not increasing");
	for (int m = 0; m < kMaterials; m++)
		d->num_nucs_host[m] = t->num_nucs[m];

	if ((rc = build_compact(d, t, flags, xs_max)))
		return rc;
	if (t->grid_type == XSB_UNIONIZED && ((rc = build_interval_records(d)) || (rc = build_ueg_acc(d))))
		return rc;
	T.xs32 = reinterpret_cast<const float *>(d->d_xs32);
	T.xs16 = d->d_xs16;
	T.xs16_scale = d->d_xs16_scale;
	T.concs32 = d->d_concs32;
	return XSB_OK;
}

static int check_handle(void *h, Device **d)
{
	if (!h)
		return fail(XSB_ERR_INVALID, "handle is NULL");
	*d = static_cast<Device *>(h);
	XSB_HIP(hipSetDevice((*d)->device));
	return XSB_OK;
}

static float elapsed(hipEvent_t a, hipEvent_t b)
{
	float ms = 0.f;
	if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
		(void)hipGetLastError();
		return 0.f;
	}
	return ms;
}

static int precision_ok(Device *d, int prec)
{
	if (prec == XSB_XS_FP64)
		return XSB_OK;
	if (prec == XSB_XS_FP32)
		return d->d_xs32 ? XSB_OK
		                 : fail(XSB_ERR_UNSUPPORTED, "FP32 tables were not built (XSB_BUILD_FP32)");
	if (prec == XSB_XS_BF16)
		return d->d_xs16 ? XSB_OK
		                 : fail(XSB_ERR_UNSUPPORTED, "BF16 tables were not built (XSB_BUILD_BF16)");
	return fail(XSB_ERR_INVALID, "xs_precision %d is not 0, 1 or 2", prec);
}

// Per-lookup detail buffers on the device, copied back after the run.
struct DetailBufs {
	DetailOut dev{};
	bool any = false;
};

static int detail_alloc(Device *d, int64_t n, const xsb_event_args *a, DetailBufs *b)
{
	b->any = a->out_energy || a->out_mat || a->out_index || a->out_macro || a->out_argmax ||
	         a->out_bound || a->out_fallback;
	if (!b->any)
		return XSB_OK;
	const size_t per = 8 + 4 + 8 + 40 + 4 + 4 + 8 + 4;
	void *base = nullptr;
	int rc = scratch_reserve(d, 3, (size_t)n * per + 4096, &base);
	if (rc)
		return rc;
	char *p = (char *)base;
	b->dev.macro = a->out_macro ? (double *)p : nullptr;
	p += (size_t)n * 40;
	b->dev.energy = a->out_energy ? (double *)p : nullptr;
	p += (size_t)n * 8;
	b->dev.index = a->out_index ? (int64_t *)p : nullptr;
	p += (size_t)n * 8;
	b->dev.mat = a->out_mat ? (int32_t *)p : nullptr;
	p += (size_t)n * 4;
	b->dev.argmax = a->out_argmax ? (int32_t *)p : nullptr;
	p += (size_t)n * 4;
	b->dev.bound = a->out_bound ? (double *)p : nullptr;
	p += (size_t)n * 8;
	b->dev.flag = a->out_fallback ? (int32_t *)p : nullptr;
	return XSB_OK;
}

static int detail_copy_back(Device *d, int64_t n, const xsb_event_args *a, const DetailBufs *b)
{
	if (!b->any)
		return XSB_OK;
	hipStream_t s = d->stream;
	if (a->out_macro)
		XSB_HIP(hipMemcpyAsync(a->out_macro, b->dev.macro, n * 40, hipMemcpyDeviceToHost, s));
	if (a->out_energy)
		XSB_HIP(hipMemcpyAsync(a->out_energy, b->dev.energy, n * 8, hipMemcpyDeviceToHost, s));
	if (a->out_index)
		XSB_HIP(hipMemcpyAsync(a->out_index, b->dev.index, n * 8, hipMemcpyDeviceToHost, s));
	if (a->out_mat)
		XSB_HIP(hipMemcpyAsync(a->out_mat, b->dev.mat, n * 4, hipMemcpyDeviceToHost, s));
	if (a->out_argmax)
		XSB_HIP(hipMemcpyAsync(a->out_argmax, b->dev.argmax, n * 4, hipMemcpyDeviceToHost, s));
	if (a->out_bound) {
		if (a->xs_precision == XSB_XS_FP64)
			XSB_HIP(hipMemsetAsync(b->dev.bound, 0, n * 8, s));
		XSB_HIP(hipMemcpyAsync(a->out_bound, b->dev.bound, n * 8, hipMemcpyDeviceToHost, s));
	}
	if (a->out_fallback) {
		if (a->xs_precision == XSB_XS_FP64)
			XSB_HIP(hipMemsetAsync(b->dev.flag, 0, n * 4, s));
		XSB_HIP(hipMemcpyAsync(a->out_fallback, b->dev.flag, n * 4, hipMemcpyDeviceToHost, s));
	}
	XSB_HIP(hipStreamSynchronize(s));
	return XSB_OK;
}

// The event run.  Work buffers (sample arrays, sort buffers) come from
// `work` (work_bytes, 256-byte aligned) when it is given, else from the
// handle's scratch; per-lookup outputs always take the handle's scratch.
static int event_impl(Device *d, const xsb_event_args *a, xsb_result *r, void *work = nullptr,
                      size_t work_bytes = 0)
{
	if (!a || !r)
		return fail(XSB_ERR_INVALID, "args and result must be non-NULL");
	if (a->lookups < 0 || a->first_lookup < 0 || a->lookups > ((int64_t)1 << 32) ||
	    a->first_lookup > ((int64_t)1 << 61))
		return fail(XSB_ERR_INVALID, "lookups=%lld first_lookup=%lld out of range",
		            (long long)a->lookups, (long long)a->first_lookup);
	int rc = precision_ok(d, a->xs_precision);
	if (rc)
		return rc;
	int k = a->kernel_id;
	if (k == XSB_K_AUTO)
		k = XSB_K_MAT_E_SORT;
	if (k < 0 || k > 6)
		return fail(XSB_ERR_INVALID, "kernel_id %d is not in [-1, 6]", a->kernel_id);
	if (k == XSB_K_BASELINE && a->xs_precision != XSB_XS_FP64)
		k = XSB_K_SPLIT;  // the fused kernel is the FP64 baseline only
	memset(r, 0, sizeof(*r));
	r->kernel_id = k;
	const int64_t n = a->lookups;
	hipStream_t s = d->stream;
	DetailBufs db;
	if ((rc = detail_alloc(d, n, a, &db)))
		return rc;
	unsigned long long *vsum = d->d_acc, *nfb = d->d_acc + 1;
	XSB_HIP(hipMemsetAsync(d->d_acc, 0, 2 * sizeof(unsigned long long), s));
	XSB_HIP(hipEventRecord(d->ev[0], s));
	if (k == XSB_K_BASELINE) {
		XSB_HIP(hipEventRecord(d->ev[1], s));
		if ((rc = launch_event_fused(d, a->first_lookup, n, a->seed, db.any ? &db.dev : nullptr,
		                             vsum, s)))
			return rc;
		XSB_HIP(hipEventRecord(d->ev[2], s));
	} else if (k == XSB_K_SPLIT) {
		void *buf = work;
		if (work && work_bytes < (size_t)n * 12 + 256)
			return fail(XSB_ERR_INVALID, "work buffer of %zu bytes for %lld lookups", work_bytes,
			            (long long)n);
		if (!work && (rc = scratch_reserve(d, 0, (size_t)n * 12 + 256, &buf)))
			return rc;
		double *E = (double *)buf;
		int32_t *M = (int32_t *)((char *)buf + (size_t)n * 8);
		if ((rc = launch_sample(d, a->first_lookup, n, a->seed, E, M, s)))
			return rc;
		XSB_HIP(hipEventRecord(d->ev[1], s));
		if ((rc = launch_lookup_list(d, n, E, M, nullptr, a->xs_precision,
		                             db.any ? &db.dev : nullptr, vsum, nfb, s)))
			return rc;
		XSB_HIP(hipEventRecord(d->ev[2], s));
	} else {
		if ((rc = run_sorted_event(d, a, db.any ? &db.dev : nullptr, r, vsum, nfb, work,
		                           work_bytes)))
			return rc;
	}
	XSB_HIP(hipEventRecord(d->ev[3], s));
	XSB_HIP(hipMemcpyAsync(d->h_out, d->d_acc, 2 * sizeof(unsigned long long),
	                       hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	r->vsum = d->h_out[0];
	r->n_fallback = (int64_t)d->h_out[1];
	if (k != XSB_K_BASELINE && k != XSB_K_SPLIT)
		finish_sorted_event(d, a, r);
	r->ms_total = elapsed(d->ev[0], d->ev[3]);
	r->ms_sample = elapsed(d->ev[0], d->ev[1]);
	r->ms_lookup = elapsed(d->ev[1], d->ev[2]);
	return detail_copy_back(d, n, a, &db);
}

}  // namespace xsb

using namespace xsb;

extern "C" {

int xsb_hip_upload(int device, const xsb_tables *t, uint32_t flags, void **handle)
{
	if (!handle)
		return fail(XSB_ERR_INVALID, "handle output pointer is NULL");
	*handle = nullptr;
	Device *d = nullptr;
	int rc = upload(device, t, flags, &d);
	if (rc) {
		release(d);
		return rc;
	}
	*handle = d;
	return XSB_OK;
}

int xsb_hip_release(void *handle)
{
	if (!handle)
		return XSB_OK;
	release(static_cast<Device *>(handle));
	return XSB_OK;
}

int xsb_hip_event(void *handle, const xsb_event_args *args, xsb_result *res)
{
	Device *d;
	int rc = check_handle(handle, &d);
	return rc ? rc : event_impl(d, args, res);
}

int xsb_hip_event_noalloc(void *handle, const xsb_event_args *args, xsb_result *res)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	if (!args || !res)
		return fail(XSB_ERR_INVALID, "args and result must be non-NULL");
	if (args->out_energy || args->out_mat || args->out_index || args->out_macro ||
	    args->out_argmax || args->out_bound || args->out_fallback)
		return fail(XSB_ERR_INVALID, "xsb_hip_event_noalloc takes no per-lookup outputs");
	if (args->lookups > XSB_NOALLOC_MAX_LOOKUPS)
		return fail(XSB_ERR_INVALID, "lookups=%lld exceeds the allocation-free maximum %d",
		            (long long)args->lookups, XSB_NOALLOC_MAX_LOOKUPS);
	// set up by xsb_hip_upload on this device; only read here
	void *work = nullptr;
	{
		std::lock_guard<std::mutex> g(g_noalloc_lock);
		if (g_noalloc_rc[d->device] != XSB_OK + 1)
			return fail(XSB_ERR_HIP, "no allocation-free work memory on device %d", d->device);
		work = g_noalloc_ptr[d->device];
	}
	return event_impl(d, args, res, work, kNoallocBytes);
}

int xsb_hip_history(void *handle, const xsb_history_args *a, xsb_result *r)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	if (!a || !r)
		return fail(XSB_ERR_INVALID, "args and result must be non-NULL");
	if (a->particles < 0 || a->particles > INT32_MAX || a->lookups_per_particle < 0)
		return fail(XSB_ERR_INVALID, "particles=%lld lookups=%d out of range",
		            (long long)a->particles, a->lookups_per_particle);
	if ((rc = precision_ok(d, a->xs_precision)))
		return rc;
	if (a->kernel_id != 0 && a->kernel_id != 6 && a->kernel_id != XSB_K_AUTO)
		return fail(XSB_ERR_INVALID, "history kernel_id %d is not 0, 6 or -1", a->kernel_id);
	memset(r, 0, sizeof(*r));
	// rounds need a lookup to sort on; the automatic choice takes them from
	// 64K particles on, where a round fills the device
	const bool rounds = a->particles > 0 && a->lookups_per_particle > 0 &&
	                    (a->kernel_id == 6 || (a->kernel_id == XSB_K_AUTO && a->particles >= 65536));
	if (rounds)
		return run_history_rounds(d, a, r);
	r->kernel_id = 0;
	hipStream_t s = d->stream;
	const int64_t n = a->particles;
	uint64_t *dsum = nullptr, *dseed = nullptr;
	if (a->out_particle_sum || a->out_final_seed) {
		void *buf = nullptr;
		if ((rc = scratch_reserve(d, 3, (size_t)n * 16 + 256, &buf)))
			return rc;
		dsum = (uint64_t *)buf;
		dseed = dsum + n;
	}
	XSB_HIP(hipMemsetAsync(d->d_acc, 0, 2 * sizeof(unsigned long long), s));
	XSB_HIP(hipEventRecord(d->ev[0], s));
	if ((rc = launch_history(d, n, a->lookups_per_particle, a->seed, a->xs_precision, dsum, dseed,
	                         d->d_acc, d->d_acc + 1, s)))
		return rc;
	XSB_HIP(hipEventRecord(d->ev[3], s));
	unsigned long long acc[2];
	XSB_HIP(hipMemcpyAsync(acc, d->d_acc, sizeof(acc), hipMemcpyDeviceToHost, s));
	if (a->out_particle_sum)
		XSB_HIP(hipMemcpyAsync(a->out_particle_sum, dsum, n * 8, hipMemcpyDeviceToHost, s));
	if (a->out_final_seed)
		XSB_HIP(hipMemcpyAsync(a->out_final_seed, dseed, n * 8, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	r->vsum = acc[0];
	r->n_fallback = (int64_t)acc[1];
	r->ms_total = r->ms_lookup = elapsed(d->ev[0], d->ev[3]);
	return XSB_OK;
}

int xsb_hip_lookup(void *handle, int64_t n, const double *d_energy, const int32_t *d_mat,
                   int32_t xs_precision, double *d_macro, int32_t *d_argmax,
                   unsigned long long *d_vsum, void *stream)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	if (n < 0 || n > ((int64_t)1 << 32) || (n > 0 && (!d_energy || !d_mat)))
		return fail(XSB_ERR_INVALID, "n=%lld or sample pointers invalid", (long long)n);
	if ((rc = precision_ok(d, xs_precision)))
		return rc;
	hipStream_t s = stream ? (hipStream_t)stream : d->stream;
	DetailOut det{};
	det.macro = d_macro;
	det.argmax = d_argmax;
	const bool detail = d_macro || d_argmax;
	return launch_lookup_list(d, n, d_energy, d_mat, nullptr, xs_precision, detail ? &det : nullptr,
	                          d_vsum ? d_vsum : d->d_acc + 40, d->d_acc + 41, s);
}

int xsb_hip_macro_host(void *handle, int64_t n, const double *energy, const int32_t *mat,
                       int32_t xs_precision, double *macro)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	if (n < 0 || n > ((int64_t)1 << 32) || (n > 0 && (!energy || !mat || !macro)))
		return fail(XSB_ERR_INVALID, "n=%lld or sample pointers invalid", (long long)n);
	if ((rc = precision_ok(d, xs_precision)))
		return rc;
	for (int64_t i = 0; i < n; i++)
		if ((uint32_t)mat[i] >= (uint32_t)kMaterials)
			return fail(XSB_ERR_INVALID, "mat[%lld]=%d is not a material id", (long long)i, mat[i]);
	if (n == 0)
		return XSB_OK;
	void *buf = nullptr;
	if ((rc = scratch_reserve(d, 3, (size_t)n * (8 + 4 + 40) + 512, &buf)))
		return rc;
	double *dE = (double *)buf;
	double *dM = dE + n;
	int32_t *dmat = (int32_t *)(dM + 5 * n);
	hipStream_t s = d->stream;
	XSB_HIP(hipMemcpyAsync(dE, energy, (size_t)n * 8, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(dmat, mat, (size_t)n * 4, hipMemcpyHostToDevice, s));
	DetailOut det{};
	det.macro = dM;
	if ((rc = launch_lookup_list(d, n, dE, dmat, nullptr, xs_precision, &det, d->d_acc + 40,
	                             d->d_acc + 41, s)))
		return rc;
	XSB_HIP(hipMemcpyAsync(macro, dM, (size_t)n * 40, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	return XSB_OK;
}

int xsb_hip_micro_host(void *handle, int64_t n, const double *energy, const int32_t *nuc,
                       const int32_t *lo, const int32_t *hi, double *xs)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	const Tables &T = d->T;
	const bool need_lo = T.grid_type != kNuclide, need_hi = T.grid_type == kHash;
	if (n < 0 || n > ((int64_t)1 << 32) ||
	    (n > 0 && (!energy || !nuc || !xs || (need_lo && !lo) || (need_hi && !hi))))
		return fail(XSB_ERR_INVALID, "n=%lld or sample pointers invalid", (long long)n);
	// every index the kernel reads with must lie in the nuclide's row
	for (int64_t i = 0; i < n; i++) {
		if (nuc[i] < 0 || nuc[i] >= T.n_iso)
			return fail(XSB_ERR_INVALID, "nuc[%lld]=%d is not a nuclide id", (long long)i, nuc[i]);
		if (need_lo && (lo[i] < 0 || lo[i] > T.n_gp - 1))
			return fail(XSB_ERR_INVALID, "lo[%lld]=%d outside [0, %d]", (long long)i, lo[i],
			            T.n_gp - 1);
		if (need_hi && (hi[i] < 0 || hi[i] > T.n_gp - 1))
			return fail(XSB_ERR_INVALID, "hi[%lld]=%d outside [0, %d]", (long long)i, hi[i],
			            T.n_gp - 1);
	}
	if (n == 0)
		return XSB_OK;
	void *buf = nullptr;
	if ((rc = scratch_reserve(d, 3, (size_t)n * (8 + 40 + 3 * 4) + 512, &buf)))
		return rc;
	double *dE = (double *)buf;
	double *dX = dE + n;
	int32_t *dnuc = (int32_t *)(dX + 5 * n), *dlo = dnuc + n, *dhi = dlo + n;
	hipStream_t s = d->stream;
	XSB_HIP(hipMemcpyAsync(dE, energy, (size_t)n * 8, hipMemcpyHostToDevice, s));
	XSB_HIP(hipMemcpyAsync(dnuc, nuc, (size_t)n * 4, hipMemcpyHostToDevice, s));
	if (need_lo)
		XSB_HIP(hipMemcpyAsync(dlo, lo, (size_t)n * 4, hipMemcpyHostToDevice, s));
	if (need_hi)
		XSB_HIP(hipMemcpyAsync(dhi, hi, (size_t)n * 4, hipMemcpyHostToDevice, s));
	if ((rc = launch_micro(d, n, dE, dnuc, dlo, dhi, dX, s)))
		return rc;
	XSB_HIP(hipMemcpyAsync(xs, dX, (size_t)n * 40, hipMemcpyDeviceToHost, s));
	XSB_HIP(hipStreamSynchronize(s));
	return XSB_OK;
}

int xsb_hip_balg(void *handle, const xsb_event_args *a, double *bytes_per_lookup)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	if (!a || !bytes_per_lookup || a->lookups < 0)
		return fail(XSB_ERR_INVALID, "bad arguments");
	return launch_balg(d, a->first_lookup, a->lookups, a->seed, a->xs_precision, bytes_per_lookup,
	                   d->stream);
}

int xsb_hip_copy_probe(void *handle, int64_t bytes, int32_t reps, double *ms_per_copy)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	if (bytes < 16 || reps < 1 || !ms_per_copy)
		return fail(XSB_ERR_INVALID, "bytes >= 16, reps >= 1 and an output pointer are required");
	bytes &= ~(int64_t)15;
	void *buf = nullptr;
	if ((rc = scratch_reserve(d, 2, (size_t)bytes * 2, &buf)))
		return rc;
	char *src = (char *)buf, *dst = src + bytes;
	hipStream_t s = d->stream;
	XSB_HIP(hipMemsetAsync(src, 1, bytes, s));
	if ((rc = launch_copy(src, dst, bytes, s)))
		return rc;
	XSB_HIP(hipEventRecord(d->ev[4], s));
	for (int i = 0; i < reps; i++)
		if ((rc = launch_copy(src, dst, bytes, s)))
			return rc;
	XSB_HIP(hipEventRecord(d->ev[5], s));
	XSB_HIP(hipStreamSynchronize(s));
	*ms_per_copy = elapsed(d->ev[4], d->ev[5]) / reps;
	return XSB_OK;
}

int xsb_hip_table_bytes(const xsb_tables *t, uint32_t flags, xsb_info *info)
{
	if (!info)
		return fail(XSB_ERR_INVALID, "info is NULL");
	int rc = validate_tables(t, flags | XSB_BUILD_GRIDS);  // sizes and materials only
	if (rc)
		return rc;
	table_bytes(t, flags, info);
	info->device = -1;
	return XSB_OK;
}

int xsb_hip_download(void *handle, double *nuclide_grid, double *unionized_energy_array,
                     int32_t *index_grid)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	return download(d, nuclide_grid, unionized_energy_array, index_grid);
}

int xsb_hip_info(void *handle, xsb_info *info)
{
	Device *d;
	int rc = check_handle(handle, &d);
	if (rc)
		return rc;
	if (!info)
		return fail(XSB_ERR_INVALID, "info is NULL");
	memset(info, 0, sizeof(*info));
	info->device = d->device;
	info->grid_type = d->T.grid_type;
	info->n_isotopes = d->T.n_iso;
	info->n_gridpoints = d->T.n_gp;
	info->n_union = d->T.n_union;
	info->bytes_fp64 = d->bytes_fp64;
	info->bytes_fp32 = d->bytes_fp32;
	info->bytes_bf16 = d->bytes_bf16;
	info->bytes_index = d->bytes_index;
	info->bytes_midx = d->bytes_midx;
	info->bytes_ivl = d->bytes_ivl;
	info->bytes_union = d->bytes_union;
	info->cu_count = d->cu_count;
	info->xcd_count = d->xcd_count;
	snprintf(info->arch, sizeof(info->arch), "%s", d->arch);
	return XSB_OK;
}

const char *xsb_hip_strerror(int code)
{
	switch (code) {
	case XSB_OK: return "success";
	case XSB_ERR_INVALID: return "invalid argument";
	case XSB_ERR_NO_DEVICE: return "no usable HIP device";
	case XSB_ERR_OOM: return "device out of memory";
	case XSB_ERR_HIP: return "HIP runtime error";
	case XSB_ERR_UNSUPPORTED: return "unsupported option";
	default: return "unknown error";
	}
}

const char *xsb_hip_last_error(void)
{
	return g_msg;
}

}  // extern "C"
