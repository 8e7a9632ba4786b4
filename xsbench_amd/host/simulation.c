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

static xsbench_options g_opts = {0, 0, 0, 0.0, 0.0, 0, 0, -1, 0};

/* XSBENCH_SCHEDULE=auto in the environment selects the fastest exact
 * schedule for every run_* call, so that a program built against the
 * reference's own Main.c reaches it without a source change. */
static void options_from_env(void)
{
	if (g_opts.schedule >= 0)
		return;
	const char *e = getenv("XSBENCH_SCHEDULE");
	g_opts.schedule = (e && strcmp(e, "auto") == 0) ? XSBENCH_SCHEDULE_AUTO : XSBENCH_SCHEDULE_ASKED;
}

xsbench_options *xsbench_get_options(void)
{
	options_from_env();
	return &g_opts;
}

/* ---- registry of device copies ----
 * An entry is keyed by the address of the host nuclide grid it was uploaded
 * from, and remembers the material arrays it holds.  Entries come from
 * move_simulation_data_to_device, or are made on first use by
 * calculate_micro_xs / calculate_macro_xs (kept until release_device_memory
 * of an SD with that nuclide grid, or the end of the process). */
#define MAX_STAGED 16
static struct {
	const void *key;
	void *handle;
	int precision_flags;
	const int *num_nucs;
	const int *mats;
	const double *concs;
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
		    g_staged[i].mats == SD.mats && g_staged[i].num_nucs == SD.num_nucs &&
		    g_staged[i].concs == SD.concs &&
		    (g_staged[i].precision_flags & flags_for_precision(g_opts.xs_precision)) ==
		        flags_for_precision(g_opts.xs_precision))
			h = g_staged[i].handle;
	pthread_mutex_unlock(&g_lock);
	return h;
}

/* Registers handle h for SD; exits when the registry is full. */
static void register_staged(SimulationData SD, void *h, int precision_flags)
{
	pthread_mutex_lock(&g_lock);
	int slot = -1;
	for (int i = 0; i < MAX_STAGED && slot < 0; i++)
		if (!g_staged[i].handle)
			slot = i;
	if (slot >= 0) {
		g_staged[slot].key = SD.nuclide_grid;
		g_staged[slot].handle = h;
		g_staged[slot].precision_flags = precision_flags;
		g_staged[slot].num_nucs = SD.num_nucs;
		g_staged[slot].mats = SD.mats;
		g_staged[slot].concs = SD.concs;
	}
	pthread_mutex_unlock(&g_lock);
	if (slot < 0) {
		xsb_hip_release(h);
		fprintf(stderr, "XSBench: more than %d device copies staged\n", MAX_STAGED);
		exit(1);
	}
}

/* cuda/GridInit.cu:4-78 */
SimulationData move_simulation_data_to_device(Inputs in, int mype, SimulationData SD)
{
	if (mype == 0)
		printf("Allocating and moving simulation data to GPU memory space...\n");
	release_device_memory(SD);
	void *h = upload(in, SD);
	register_staged(SD, h, (int)flags_for_precision(g_opts.xs_precision));
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
	g_opts.last_kernel_id = r.kernel_id;
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
	options_from_env();
	return run_event(in, SD, g_opts.schedule == XSBENCH_SCHEDULE_AUTO ? XSB_K_AUTO : in.kernel_id);
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
	/* the reference runs history mode whatever -k says (Main.c:80-93); here
	 * -k 0 is one particle per lane and -k 6 the sorted rounds, and any other
	 * id (or XSBENCH_SCHEDULE=auto) takes the automatic choice */
	options_from_env();
	a.kernel_id = (g_opts.schedule == XSBENCH_SCHEDULE_AUTO)       ? XSB_K_AUTO
	              : (in.kernel_id == 0 || in.kernel_id == XSB_K_MAT_E_SORT) ? in.kernel_id
	                                                                       : XSB_K_AUTO;
	a.xs_precision = g_opts.xs_precision;
	a.seed = STARTING_SEED;
	xsb_result r;
	int rc = xsb_hip_history(h, &a, &r);
	if (rc)
		die(rc, "xsb_hip_history");
	g_opts.last_kernel_id = r.kernel_id;
	g_opts.last_kernel_ms = r.ms_total;
	g_opts.last_lookup_ms = r.ms_lookup;
	g_opts.last_fallbacks = r.n_fallback;
	if (temporary)
		xsb_hip_release(h);
	return r.vsum;
}

/* ---- per-lookup functions of the reference's interface ---- */

/* openmp-threading/Simulation.c:380-400: the last index in [0, n - 1) whose
 * value is <= quarry (clamped to [0, n - 2]), by bisection. */
long grid_search(long n, double quarry, double *restrict A)
{
	long lo = 0, hi = n - 1;
	while (hi - lo > 1) {
		const long mid = lo + (hi - lo) / 2;
		if (A[mid] > quarry)
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

/* openmp-threading/Simulation.c:403-423: the same on A[low..high].energy. */
long grid_search_nuclide(long n, double quarry, NuclideGridPoint *A, long low, long high)
{
	(void)n;
	long lo = low, hi = high;
	while (hi - lo > 1) {
		const long mid = lo + (hi - lo) / 2;
		if (A[mid].energy > quarry)
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

/* openmp-threading/Simulation.c:426-460: one LCG draw against the volume
 * fractions' reverse-order partial sums (dist[i] + dist[i-1] + ... + dist[1],
 * summed in that order, so the thresholds are the reference's bit for bit);
 * the first threshold above the roll names the material, else material 0. */
int pick_mat(uint64_t *seed)
{
	static const double dist[12] = {0.140, 0.052, 0.275, 0.134, 0.154, 0.064,
	                                0.066, 0.055, 0.008, 0.015, 0.025, 0.013};
	const double roll = LCG_random_double(seed);
	for (int i = 0; i < 12; i++) {
		double running = 0.0;
		for (int j = i; j > 0; j--)
			running += dist[j];
		if (roll < running)
			return i;
	}
	return 0;
}

/* The device copy of the tables behind calculate_micro_xs/_macro_xs: a
 * registered copy of the same host nuclide grid (with the same material
 * arrays when `mats` is given), else one uploaded now and registered. */
static void *lookup_tables(long n_isotopes, long n_gridpoints, int grid_type, int hash_bins,
                           double *egrid, int *index_data, NuclideGridPoint *nuclide_grids,
                           int *num_nucs, double *concs, int *mats, int max_num_nucs)
{
	void *h = NULL;
	pthread_mutex_lock(&g_lock);
	for (int i = 0; i < MAX_STAGED && !h; i++)
		if (g_staged[i].handle && g_staged[i].key == nuclide_grids &&
		    (!mats || (g_staged[i].mats == mats && g_staged[i].num_nucs == num_nucs &&
		               g_staged[i].concs == concs)))
			h = g_staged[i].handle;
	pthread_mutex_unlock(&g_lock);
	if (h)
		return h;
	/* calculate_micro_xs reads no material: one placeholder nuclide per material */
	static int one_nuc[12] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
	static int nuc0[12] = {0};
	static double conc0[12] = {0.0};
	Inputs in;
	memset(&in, 0, sizeof(in));
	in.n_isotopes = n_isotopes;
	in.n_gridpoints = n_gridpoints;
	in.grid_type = grid_type;
	in.hash_bins = hash_bins;
	SimulationData SD;
	memset(&SD, 0, sizeof(SD));
	SD.nuclide_grid = nuclide_grids;
	SD.length_nuclide_grid = (int)(n_isotopes * n_gridpoints);
	SD.num_nucs = mats ? num_nucs : one_nuc;
	SD.mats = mats ? mats : nuc0;
	SD.concs = mats ? concs : conc0;
	SD.max_num_nucs = mats ? max_num_nucs : 1;
	if (grid_type == UNIONIZED) {
		SD.unionized_energy_array = egrid;
		SD.length_unionized_energy_array = (int)(n_isotopes * n_gridpoints);
		SD.index_grid = index_data;
		SD.length_index_grid = (long)SD.length_unionized_energy_array * n_isotopes;
	} else if (grid_type == HASH) {
		SD.index_grid = index_data;
		SD.length_index_grid = (long)hash_bins * n_isotopes;
	}
	xsb_tables t;
	fill_tables(in, SD, &t);
	int rc = xsb_hip_upload(g_opts.device, &t, 0, &h);
	if (rc)
		die(rc, "xsb_hip_upload");
	register_staged(SD, h, 0);
	return h;
}

/* openmp-threading/Simulation.c:241-323, on the GPU: the host reads what the
 * reference reads from index_data at grid index idx (the bracket, or the
 * hash bin's bounds); the search and the interpolation run on the device
 * copy of the tables (xsb_hip_micro_host).  One call is one device round
 * trip: callers with many samples use xsb_hip_micro_host directly. */
void calculate_micro_xs(double p_energy, int nuc, long n_isotopes, long n_gridpoints,
                        double *restrict egrid, int *restrict index_data,
                        NuclideGridPoint *restrict nuclide_grids, long idx,
                        double *restrict xs_vector, int grid_type, int hash_bins)
{
	void *h = lookup_tables(n_isotopes, n_gridpoints, grid_type, hash_bins, egrid, index_data,
	                        nuclide_grids, NULL, NULL, NULL, 0);
	int32_t n32 = nuc, lo = 0, hi = 0;
	if (grid_type == UNIONIZED) {
		lo = index_data[idx * n_isotopes + nuc];
	} else if (grid_type == HASH) {
		lo = index_data[idx * n_isotopes + nuc];
		hi = (idx == hash_bins - 1) ? (int32_t)(n_gridpoints - 1)
		                            : index_data[(idx + 1) * n_isotopes + nuc] + 1;
	}
	int rc = xsb_hip_micro_host(h, 1, &p_energy, &n32, &lo, &hi, xs_vector);
	if (rc)
		die(rc, "xsb_hip_micro_host");
}

/* openmp-threading/Simulation.c:326-375, on the GPU (xsb_hip_macro_host,
 * FP64: the reference's sums bit for bit).  One call is one device round
 * trip: callers with many samples use xsb_hip_macro_host or xsb_hip_lookup. */
void calculate_macro_xs(double p_energy, int mat, long n_isotopes, long n_gridpoints,
                        int *restrict num_nucs, double *restrict concs, double *restrict egrid,
                        int *restrict index_data, NuclideGridPoint *restrict nuclide_grids,
                        int *restrict mats, double *restrict macro_xs_vector, int grid_type,
                        int hash_bins, int max_num_nucs)
{
	void *h = lookup_tables(n_isotopes, n_gridpoints, grid_type, hash_bins, egrid, index_data,
	                        nuclide_grids, num_nucs, concs, mats, max_num_nucs);
	int32_t m32 = mat;
	int rc = xsb_hip_macro_host(h, 1, &p_energy, &m32, XSB_XS_FP64, macro_xs_vector);
	if (rc)
		die(rc, "xsb_hip_macro_host");
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
