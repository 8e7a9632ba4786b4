/*
 * simulation.c -- the reference's lookup drivers, executed on the GPU
 * through libxsb_hip.so (openmp-threading/Simulation.c:15-238, :698-871;
 * cuda/GridInit.cu:4-88).
 *
 * move_simulation_data_to_device() uploads a SimulationData and registers
 * the device copy under the address of its nuclide grid; the run_*
 * functions look the copy up (or upload a temporary one when the caller
 * never staged SD, as openmp-threading's main does) and return the
 * un-modded verification sum, like the reference.  Failures print the
 * library's message and exit(1), as the reference's gpuErrchk does.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xsb_hip.h"
#include "xsbench.h"

#define LCG_M 9223372036854775808ULL
#define LCG_A 2806196910506780709ULL

/* openmp-threading/Simulation.c:462-470 */
double LCG_random_double(uint64_t *seed)
{
	*seed = (LCG_A * (*seed) + 1ULL) % LCG_M;
	return (double)(*seed) / (double)LCG_M;
}

/* openmp-threading/Simulation.c:472-499 */
uint64_t fast_forward_LCG(uint64_t seed, uint64_t n)
{
	uint64_t a = LCG_A, c = 1ULL, an = 1ULL, cn = 0ULL;
	n %= LCG_M;
	while (n > 0) {
		if (n & 1) {
			an *= a;
			cn = cn * a + c;
		}
		c *= (a + 1);
		a *= a;
		n >>= 1;
	}
	return (an * seed + cn) % LCG_M;
}

static xsbench_options g_opts = {0, 0, 0, 0.0, 0.0, 0, 0};

xsbench_options *xsbench_get_options(void)
{
	return &g_opts;
}

/* ---- registry of device copies ---- */
#define MAX_STAGED 16
static struct {
	const void *key;
	void *handle;
	int precision_flags;
} g_staged[MAX_STAGED];
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

static void die(int rc, const char *what)
{
	fprintf(stderr, "XSBench GPU error in %s: %s: %s\n", what, xsb_hip_strerror(rc),
	        xsb_hip_last_error());
	exit(1);
}

static uint32_t flags_for_precision(int prec)
{
	if (prec == 1)
		return XSB_BUILD_FP32;
	if (prec == 2)
		return XSB_BUILD_BF16;
	return 0;
}

static void fill_tables(Inputs in, SimulationData SD, xsb_tables *t)
{
	memset(t, 0, sizeof(*t));
	t->n_isotopes = in.n_isotopes;
	t->n_gridpoints = in.n_gridpoints;
	t->grid_type = in.grid_type;
	t->hash_bins = in.hash_bins;
	t->num_nucs = SD.num_nucs;
	t->mats = SD.mats;
	t->concs = SD.concs;
	t->max_num_nucs = SD.max_num_nucs;
	t->unionized_energy_array = (in.grid_type == UNIONIZED) ? SD.unionized_energy_array : NULL;
	t->length_unionized_energy_array =
	    (in.grid_type == UNIONIZED) ? SD.length_unionized_energy_array : 0;
	t->index_grid = (in.grid_type == NUCLIDE) ? NULL : SD.index_grid;
	t->length_index_grid = (in.grid_type == NUCLIDE) ? 0 : SD.length_index_grid;
	t->nuclide_grid = (const double *)SD.nuclide_grid;
	t->length_nuclide_grid = SD.length_nuclide_grid;
}

static void *upload(Inputs in, SimulationData SD)
{
	xsb_tables t;
	fill_tables(in, SD, &t);
	void *h = NULL;
	int rc = xsb_hip_upload(g_opts.device, &t, flags_for_precision(g_opts.xs_precision), &h);
	if (rc)
		die(rc, "xsb_hip_upload");
	return h;
}

static void *find_staged(SimulationData SD)
{
	void *h = NULL;
	pthread_mutex_lock(&g_lock);
	for (int i = 0; i < MAX_STAGED; i++)
		if (g_staged[i].key == SD.nuclide_grid && g_staged[i].handle &&
		    (g_staged[i].precision_flags & flags_for_precision(g_opts.xs_precision)) ==
		        flags_for_precision(g_opts.xs_precision))
			h = g_staged[i].handle;
	pthread_mutex_unlock(&g_lock);
	return h;
}

/* cuda/GridInit.cu:4-78 */
SimulationData move_simulation_data_to_device(Inputs in, int mype, SimulationData SD)
{
	if (mype == 0)
		printf("Allocating and moving simulation data to GPU memory space...\n");
	release_device_memory(SD);
	void *h = upload(in, SD);
	pthread_mutex_lock(&g_lock);
	int slot = -1;
	for (int i = 0; i < MAX_STAGED && slot < 0; i++)
		if (!g_staged[i].handle)
			slot = i;
	if (slot >= 0) {
		g_staged[slot].key = SD.nuclide_grid;
		g_staged[slot].handle = h;
		g_staged[slot].precision_flags = (int)flags_for_precision(g_opts.xs_precision);
	}
	pthread_mutex_unlock(&g_lock);
	if (slot < 0) {
		xsb_hip_release(h);
		fprintf(stderr, "move_simulation_data_to_device: more than %d staged copies\n", MAX_STAGED);
		exit(1);
	}
	if (mype == 0) {
		xsb_info info;
		if (xsb_hip_info(h, &info) == XSB_OK)
			printf("GPU %s: %d CUs, %.0lf MB of tables resident\n", info.arch, info.cu_count,
			       (info.bytes_fp64 + info.bytes_fp32 + info.bytes_bf16 + info.bytes_index +
			        info.bytes_midx + info.bytes_union + info.bytes_ivl) / 1048576.0);
	}
	return SD;
}

/* cuda/GridInit.cu:81-88 */
void release_device_memory(SimulationData SD)
{
	pthread_mutex_lock(&g_lock);
	for (int i = 0; i < MAX_STAGED; i++) {
		if (g_staged[i].handle && g_staged[i].key == SD.nuclide_grid) {
			xsb_hip_release(g_staged[i].handle);
			g_staged[i].handle = NULL;
			g_staged[i].key = NULL;
		}
	}
	pthread_mutex_unlock(&g_lock);
}

static unsigned long long run_event(Inputs in, SimulationData SD, int kernel_id)
{
	void *h = find_staged(SD);
	int temporary = 0;
	if (!h) {
		h = upload(in, SD);
		temporary = 1;
	}
	xsb_event_args a;
	memset(&a, 0, sizeof(a));
	a.lookups = in.lookups;
	a.first_lookup = 0;
	a.seed = STARTING_SEED;
	a.kernel_id = kernel_id;
	a.xs_precision = g_opts.xs_precision;
	xsb_result r;
	int rc = xsb_hip_event(h, &a, &r);
	if (rc)
		die(rc, "xsb_hip_event");
	g_opts.last_kernel_ms = r.ms_total;
	g_opts.last_lookup_ms = r.ms_lookup;
	g_opts.last_fallbacks = r.n_fallback;
	if (g_opts.verbose)
		printf("GPU kernels: %.3f ms total, %.3f ms sampling/sort, %.3f ms lookups, %lld FP64 "
		       "re-evaluations\n",
		       r.ms_total, r.ms_sample, r.ms_lookup, (long long)r.n_fallback);
	if (temporary)
		xsb_hip_release(h);
	return r.vsum;
}

/* openmp-threading/Simulation.c:15-114.  The GPU schedule follows
 * in.kernel_id with the GPU port's numbering (cuda/Main.cu:62-83):
 * 0 baseline, 1 split, 2/4 material segments, 3/5 fuel first, 6 material +
 * energy sort. */
unsigned long long run_event_based_simulation(Inputs in, SimulationData SD, int mype)
{
	if (mype == 0)
		printf("Beginning event based simulation...\n");
	return run_event(in, SD, in.kernel_id);
}

/* openmp-threading/Simulation.c:698-871: sample, sort by material and then
 * energy, look up material by material. */
unsigned long long run_event_based_simulation_optimization_1(Inputs in, SimulationData SD, int mype)
{
	if (mype == 0)
		printf("Simulation Kernel:\"Optimization 1 - Kernel splitting + full material & energy "
		       "sort\"\n");
	return run_event(in, SD, XSB_K_MAT_E_SORT);
}

/* openmp-threading/Simulation.c:116-238 */
unsigned long long run_history_based_simulation(Inputs in, SimulationData SD, int mype)
{
	if (mype == 0)
		printf("Beginning history based simulation...\n");
	void *h = find_staged(SD);
	int temporary = 0;
	if (!h) {
		h = upload(in, SD);
		temporary = 1;
	}
	xsb_history_args a;
	memset(&a, 0, sizeof(a));
	a.particles = in.particles;
	a.lookups_per_particle = in.lookups;
	a.kernel_id = in.kernel_id;
	a.xs_precision = g_opts.xs_precision;
	a.seed = STARTING_SEED;
	xsb_result r;
	int rc = xsb_hip_history(h, &a, &r);
	if (rc)
		die(rc, "xsb_hip_history");
	g_opts.last_kernel_ms = r.ms_total;
	g_opts.last_lookup_ms = r.ms_lookup;
	g_opts.last_fallbacks = r.n_fallback;
	if (temporary)
		xsb_hip_release(h);
	return r.vsum;
}

/* ---- batched workload (no reference counterpart) ---- */

/* Batch b: u = LCG draw, size = round(exp(ln lo + u (ln hi - ln lo)))
 * clamped to [lo, hi]; one draw is skipped and the generator state becomes
 * the batch's stream seed; then 12 weights 0.05 + U(0,1), normalised into
 * cumulative thresholds, with thresholds[11] = 2 so that every roll < 1
 * picks a material. */
void xsbench_make_batches(uint64_t seed, int n_batches, long lo, long hi, long *sizes,
                          uint64_t *batch_seeds, double *thresholds)
{
	uint64_t s = seed;
	const double llo = log((double)lo), lhi = log((double)hi);
	for (int b = 0; b < n_batches; b++) {
		long size = lround(exp(llo + LCG_random_double(&s) * (lhi - llo)));
		sizes[b] = size < lo ? lo : (size > hi ? hi : size);
		(void)LCG_random_double(&s);
		batch_seeds[b] = s;
		double w[12], total = 0.0;
		for (int m = 0; m < 12; m++) {
			w[m] = 0.05 + LCG_random_double(&s);
			total += w[m];
		}
		double acc = 0.0;
		for (int m = 0; m < 11; m++) {
			acc += w[m];
			thresholds[12 * b + m] = acc / total;
		}
		thresholds[12 * b + 11] = 2.0;
	}
}

/* Lookup i of a batch: s = FF(batch_seed, 2i), energy = LCG(s), roll = LCG(s),
 * material = first m with roll < thresholds[m]. */
void xsbench_batch_sample(uint64_t batch_seed, const double *thresholds, long i, double *energy,
                          int *mat)
{
	uint64_t s = fast_forward_LCG(batch_seed, 2ULL * (uint64_t)i);
	*energy = LCG_random_double(&s);
	const double roll = LCG_random_double(&s);
	int m = 0;
	while (m < 11 && !(roll < thresholds[m]))
		m++;
	*mat = m;
}

/* All batches on the GPU copy of SD (staged, or a temporary upload) with the
 * sorted pipeline and the autotuned plan; returns the un-modded sum over all
 * batches, per-batch sums in out_vsum and completion times in out_done_ms
 * (both optional). */
unsigned long long run_batched_simulation(Inputs in, SimulationData SD, int n_batches,
                                          const long *sizes, const uint64_t *batch_seeds,
                                          const double *thresholds, unsigned long long *out_vsum,
                                          double *out_done_ms)
{
	void *h = find_staged(SD);
	int temporary = 0;
	if (!h) {
		h = upload(in, SD);
		temporary = 1;
	}
	int64_t *sz = malloc(sizeof(int64_t) * (size_t)(n_batches > 0 ? n_batches : 1));
	if (!sz) {
		fprintf(stderr, "out of host memory\n");
		exit(1);
	}
	for (int b = 0; b < n_batches; b++)
		sz[b] = sizes[b];
	xsb_batched_args a;
	memset(&a, 0, sizeof(a));
	a.n_batches = n_batches;
	a.xs_precision = g_opts.xs_precision;
	a.sizes = sz;
	a.seeds = batch_seeds;
	a.thresholds = thresholds;
	a.mode = 0;
	a.out_vsum = (uint64_t *)out_vsum;
	a.out_done_ms = out_done_ms;
	xsb_result r;
	int rc = xsb_hip_batched(h, &a, &r);
	free(sz);
	if (rc)
		die(rc, "xsb_hip_batched");
	g_opts.last_kernel_ms = r.ms_total;
	g_opts.last_lookup_ms = r.ms_lookup;
	g_opts.last_fallbacks = r.n_fallback;
	if (temporary)
		xsb_hip_release(h);
	return r.vsum;
}
