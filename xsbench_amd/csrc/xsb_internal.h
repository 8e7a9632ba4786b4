// xsb_internal.h -- host-side internals shared by the HIP translation units
// of libxsb_hip.so (device handle, error plumbing, kernel launchers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "xsb_device.h"
#include "xsb_hip.h"

namespace xsb {

// Records a message for xsb_hip_last_error() and returns `code`.
int fail(int code, const char *fmt, ...);

#define XSB_HIP(call)                                                                  \
	do {                                                                               \
		hipError_t e_ = (call);                                                        \
		if (e_ != hipSuccess)                                                          \
			return ::xsb::fail(XSB_ERR_HIP, "%s failed: %s (%s:%d)", #call,            \
			                   hipGetErrorString(e_), __FILE__, __LINE__);             \
	} while (0)

// Growable device scratch owned by a handle.
struct Scratch {
	void *ptr = nullptr;
	size_t bytes = 0;
};

// An instantiated hipGraph and the key it was captured for (the sizes,
// arguments and buffer addresses its nodes hold).
struct GraphCache {
	hipGraphExec_t exec = nullptr;
	std::vector<int64_t> key;
};

struct Device {
	int device = 0;
	hipStream_t stream = nullptr;
	Tables T{};
	int cu_count = 0;
	int xcd_count = 8;
	char arch[32] = {0};
	// owned allocations
	double *d_ueg = nullptr;
	int32_t *d_index = nullptr;
	double *d_hwin = nullptr;   // hash grid: window records per bin and nuclide (Tables::hwin)
	double *d_esoa = nullptr;   // hash grid: energies by nuclide (Tables::esoa)
	uint8_t *d_midx = nullptr;  // per-material unionized index tables (uint16 or int32)
	double *d_grid = nullptr;
	double *d_ivl = nullptr;    // FP64 interval records (Tables::ivl)
	int32_t *d_ueg_acc = nullptr;  // unionized search brackets (Tables::ueg_acc)
	double *d_ivl32 = nullptr;  // FP32 interval records (Tables::ivl32)
	float *d_xs32 = nullptr;
	uint16_t *d_xs16 = nullptr;
	float *d_xs16_scale = nullptr;
	int32_t *d_num_nucs = nullptr;
	int32_t *d_mats = nullptr;
	double *d_concs = nullptr;
	float *d_concs32 = nullptr;
	unsigned long long *d_acc = nullptr;  // 64 accumulators
	// pinned host words the event path copies its results into (2
	// accumulators, then kBins digit starts), so one synchronisation ends a call
	unsigned long long *h_out = nullptr;
	uint64_t *d_fftab = nullptr;          // LCG skip-ahead table
	int64_t bytes_fp64 = 0, bytes_fp32 = 0, bytes_bf16 = 0, bytes_index = 0, bytes_union = 0;
	int64_t bytes_midx = 0, bytes_ivl = 0;
	int32_t num_nucs_host[kMaterials] = {0};
	Scratch scratch[4];
	hipEvent_t ev[16] = {};
	struct BatchState *batch = nullptr;  // batched path: graph cache, tuned plans
	GraphCache history_graph;            // history rounds
};

// Frees what the batched path keeps on a handle (xsb_batch.hip).
void batch_state_free(Device *d);

// Grows scratch slot `slot` of `d` to at least `bytes`.
int scratch_reserve(Device *d, int slot, size_t bytes, void **out);

// Per-lookup optional outputs of the lookup kernels (device pointers).
struct DetailOut {
	double *energy;
	int32_t *mat;
	int64_t *index;
	double *macro;
	int32_t *argmax;
	int32_t *flag;  // mixed precision: 1 where the lookup was re-evaluated in FP64
	double *bound;  // mixed precision: the error bound B of each lookup
};

// ---- launchers (xsb_lookup.hip) ----
int launch_event_fused(Device *d, int64_t first, int64_t n, uint64_t seed, const DetailOut *det,
                       unsigned long long *vsum, hipStream_t s);
int launch_sample(Device *d, int64_t first, int64_t n, uint64_t seed, double *E, int32_t *mat,
                  hipStream_t s);
int launch_lookup_list(Device *d, int64_t n, const double *E, const int32_t *mat,
                       const uint32_t *perm, int precision, const DetailOut *det,
                       unsigned long long *vsum, unsigned long long *nfallback, hipStream_t s);
int launch_micro(Device *d, int64_t n, const double *E, const int32_t *nuc, const int32_t *lo,
                 const int32_t *hi, double *xs, hipStream_t s);
int launch_history(Device *d, int64_t particles, int lookups, uint64_t seed, int precision,
                   uint64_t *out_sum, uint64_t *out_seed, unsigned long long *vsum,
                   unsigned long long *nfallback, hipStream_t s);
int launch_balg(Device *d, int64_t first, int64_t n, uint64_t seed, int precision, double *out_sum,
                hipStream_t s);
int launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t s);

// ---- sorted schedules (xsb_sched.hip) ----
// Enqueues the sorted schedule (no synchronisation; the digit starts are
// copied to h_out + 2), then, once the stream has been synchronised,
// finish_sorted_event fills r->material_counts from them.  The work buffers
// (sorted_event_bytes(lookups), 16-byte aligned) come from `work` when it is
// given, else from the handle's scratch.
int run_sorted_event(Device *d, const xsb_event_args *a, const DetailOut *det, xsb_result *r,
                     unsigned long long *vsum, unsigned long long *nfb, void *work = nullptr,
                     size_t work_bytes = 0);
void finish_sorted_event(const Device *d, const xsb_event_args *a, xsb_result *r);

// ---- grid construction on the device (xsb_gridinit.hip): allocates and
// fills d_grid, d_ueg (unionized) or d_index (hash grid, rows padded to
// `pitch`), and returns each nuclide's max |xs| ----
int build_grids_on_device(Device *d, const xsb_tables *t, int64_t pitch, std::vector<double> *xs_max);
// Rows [row0, row0 + rows) of the unionized index grid, swept from the
// device's nuclide grid and unionized energies, into out (row pitch `pitch`).
int sweep_index_rows(Device *d, int64_t row0, int64_t rows, int32_t pitch, int32_t *out,
                     hipStream_t s);

// ---- history in material-energy sorted rounds (xsb_history.hip) ----
int run_history_rounds(Device *d, const xsb_history_args *a, xsb_result *r);

}  // namespace xsb
