/*
 * grid_init.c -- host construction of the lookup tables
 * (openmp-threading/GridInit.c, Materials.c, XSutils.c).
 *
 * Produces arrays identical, byte for byte, to the reference's serial
 * construction, using every host core:
 *  - nuclide grid: the reference's single LCG stream (seed 42, six draws per
 *    point) split into chunks that start from a fast-forwarded state;
 *  - per-nuclide sort: the same libc qsort with the same comparator, one
 *    nuclide per task (equal energies keep qsort's order);
 *  - unionized energies: a multiway merge of the sorted nuclide grids split
 *    into disjoint energy intervals (equal doubles are indistinguishable, so
 *    any correct sort gives the reference's array);
 *  - index grid: when the unionized energies hold no duplicate, the sweep
 *    state after row e equals clamp(#{k : E_i[k] <= U[e]} - 1, 0, n - 2) for
 *    every nuclide i, so row ranges are swept independently from that closed
 *    form; with duplicates the reference's serial sweep runs unchanged.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "xsbench.h"

static int g_last_parallel = 0;
int xsbench_last_init_parallel(void)
{
	return g_last_parallel;
}

/* openmp-threading/XSutils.c:3-14 */
int double_compare(const void *a, const void *b)
{
	double A = *(const double *)a, B = *(const double *)b;
	return (A > B) - (A < B);
}

/* openmp-threading/XSutils.c:16-27 */
int NGP_compare(const void *a, const void *b)
{
	double A = ((const NuclideGridPoint *)a)->energy;
	double B = ((const NuclideGridPoint *)b)->energy;
	return (A > B) - (A < B);
}

/* openmp-threading/XSutils.c:30-47 (estimate in MB) */
size_t estimate_mem_usage(Inputs in)
{
	size_t grids = (size_t)in.n_isotopes * in.n_gridpoints * sizeof(NuclideGridPoint);
	size_t ueg = (size_t)in.n_isotopes * in.n_gridpoints * sizeof(double) +
	             (size_t)in.n_isotopes * in.n_gridpoints * in.n_isotopes * sizeof(int);
	size_t hash = (size_t)in.hash_bins * in.n_isotopes * sizeof(int);
	size_t total = grids;
	if (in.grid_type == UNIONIZED)
		total += ueg;
	else if (in.grid_type == HASH)
		total += hash;
	return (size_t)ceil(total / (1024.0 * 1024.0));
}

/* openmp-threading/XSutils.c:49-63 */
double get_time(void)
{
	return omp_get_wtime();
}

/* ---- materials (openmp-threading/Materials.c) ---- */

/* Nuclide counts per material; the fuel holds 34 nuclides in the small
 * problem (68 isotopes) and 321 otherwise (Materials.c:7-31). */
int *load_num_nucs(long n_isotopes)
{
	static const int rest[11] = {5, 4, 4, 27, 21, 21, 21, 21, 21, 9, 9};
	int *num_nucs = malloc(12 * sizeof(int));
	num_nucs[0] = (n_isotopes == 68) ? 34 : 321;
	memcpy(num_nucs + 1, rest, sizeof(rest));
	return num_nucs;
}

/* Nuclide ids of each material (Materials.c:34-96): fuel, cladding, cold and
 * hot borated water, RPV, five reflector/plate/nozzle regions sharing one
 * composition, and the two assembly ends.  Large fuel appends ids 68..354. */
int *load_mats(int *num_nucs, long n_isotopes, int *max_num_nucs)
{
	static const int fuel[34] = {58, 59, 60, 61, 40, 42, 43, 44, 45, 46, 1,  2,
	                             3,  7,  8,  9,  10, 29, 57, 47, 48, 0,  62, 15,
	                             33, 34, 52, 53, 54, 55, 56, 18, 23, 41};
	static const int clad[5] = {63, 64, 65, 66, 67};
	static const int water[4] = {24, 41, 4, 5};
	static const int rpv[27] = {19, 20, 21, 22, 35, 36, 37, 38, 39, 25, 27, 28, 29, 30,
	                            31, 32, 26, 49, 50, 51, 11, 12, 13, 14, 6,  16, 17};
	static const int refl[21] = {24, 41, 4,  5,  19, 20, 21, 22, 35, 36, 37,
	                             38, 39, 25, 49, 50, 51, 11, 12, 13, 14};
	static const int fa_end[9] = {24, 41, 4, 5, 63, 64, 65, 66, 67};
	const int *lists[12] = {fuel, clad, water, water, rpv, refl, refl, refl, refl, refl, fa_end, fa_end};
	(void)n_isotopes;
	int mx = 0;
	for (int m = 0; m < 12; m++)
		if (num_nucs[m] > mx)
			mx = num_nucs[m];
	*max_num_nucs = mx;
	int *mats = calloc((size_t)12 * mx, sizeof(int));
	for (int m = 0; m < 12; m++)
		for (int j = 0; j < num_nucs[m]; j++)
			mats[m * mx + j] = (m == 0 && j >= 34) ? 68 + (j - 34) : lists[m][j];
	return mats;
}

/* One LCG stream seeded STARTING_SEED^2, material-major (Materials.c:99-116). */
double *load_concs(int *num_nucs, int max_num_nucs)
{
	uint64_t seed = (uint64_t)STARTING_SEED * STARTING_SEED;
	double *concs = calloc((size_t)12 * max_num_nucs, sizeof(double));
	for (int m = 0; m < 12; m++)
		for (int j = 0; j < num_nucs[m]; j++)
			concs[m * max_num_nucs + j] = LCG_random_double(&seed);
	return concs;
}

/* ---- grid construction ---- */

/* clamp(#{k < n : A[k].energy <= q} - 1, 0, n - 2) */
static long bracket_lower(const NuclideGridPoint *A, long n, double q)
{
	long lo = 0, hi = n - 1;
	while (hi - lo > 1) {
		long mid = lo + (hi - lo) / 2;
		if (A[mid].energy > q)
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

/* number of points of A[0..n) with energy < q */
static long count_below(const NuclideGridPoint *A, long n, double q)
{
	long lo = 0, hi = n;
	while (lo < hi) {
		long mid = lo + (hi - lo) / 2;
		if (A[mid].energy < q)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

static void build_nuclide_grid(NuclideGridPoint *g, long n_pts)
{
	const long chunk = 1 << 16;
	const long nchunks = (n_pts + chunk - 1) / chunk;
#pragma omp parallel for schedule(static)
	for (long c = 0; c < nchunks; c++) {
		long begin = c * chunk, end = begin + chunk < n_pts ? begin + chunk : n_pts;
		uint64_t seed = fast_forward_LCG(42, 6ULL * (uint64_t)begin);
		for (long i = begin; i < end; i++) {
			g[i].energy = LCG_random_double(&seed);
			g[i].total_xs = LCG_random_double(&seed);
			g[i].elastic_xs = LCG_random_double(&seed);
			g[i].absorbtion_xs = LCG_random_double(&seed);
			g[i].fission_xs = LCG_random_double(&seed);
			g[i].nu_fission_xs = LCG_random_double(&seed);
		}
	}
}

/* U = all nuclide energies in ascending order.  Interval t of T covers
 * energies [t/T, (t+1)/T) (the last one also everything >= 1); each
 * interval's energies are collected from every nuclide and sorted. */
static void build_unionized(const NuclideGridPoint *g, long n_iso, long n_gp, double *U)
{
	const int T = 256;
	long *cut = malloc((size_t)(T + 1) * n_iso * sizeof(long)); /* cut[t][i] */
#pragma omp parallel for schedule(static)
	for (long i = 0; i < n_iso; i++) {
		const NuclideGridPoint *A = g + i * n_gp;
		cut[i] = 0;
		for (int t = 1; t < T; t++)
			cut[(long)t * n_iso + i] = count_below(A, n_gp, (double)t / T);
		cut[(long)T * n_iso + i] = n_gp;
	}
	long *start = malloc((size_t)(T + 1) * sizeof(long));
	start[0] = 0;
	for (int t = 0; t < T; t++) {
		long c = 0;
		for (long i = 0; i < n_iso; i++)
			c += cut[(long)(t + 1) * n_iso + i] - cut[(long)t * n_iso + i];
		start[t + 1] = start[t] + c;
	}
#pragma omp parallel for schedule(dynamic, 1)
	for (int t = 0; t < T; t++) {
		double *out = U + start[t];
		long k = 0;
		for (long i = 0; i < n_iso; i++) {
			const NuclideGridPoint *A = g + i * n_gp;
			for (long j = cut[(long)t * n_iso + i]; j < cut[(long)(t + 1) * n_iso + i]; j++)
				out[k++] = A[j].energy;
		}
		qsort(out, k, sizeof(double), double_compare);
	}
	free(start);
	free(cut);
}

/* The reference's sweep (GridInit.c:92-122) over rows [e0, e1), starting
 * from the state after row e0 - 1 held in idx_low / e_high. */
static void sweep_rows(const NuclideGridPoint *g, long n_iso, long n_gp, const double *U, long e0,
                       long e1, int *idx_low, double *e_high, int *index_grid)
{
	for (long e = e0; e < e1; e++) {
		const double u = U[e];
		int *out = index_grid + e * n_iso;
		for (long i = 0; i < n_iso; i++) {
			if (!(u < e_high[i]) && idx_low[i] != n_gp - 2) {
				idx_low[i]++;
				e_high[i] = g[i * n_gp + idx_low[i] + 1].energy;
			}
			out[i] = idx_low[i];
		}
	}
}

static void build_index_grid(const NuclideGridPoint *g, long n_iso, long n_gp, const double *U,
                             long n_u, int *index_grid)
{
	long dups = 0;
#pragma omp parallel for reduction(+ : dups)
	for (long e = 1; e < n_u; e++)
		dups += (U[e] == U[e - 1]);
	if (dups) {
		int *idx_low = calloc(n_iso, sizeof(int));
		double *e_high = malloc(n_iso * sizeof(double));
		for (long i = 0; i < n_iso; i++)
			e_high[i] = g[i * n_gp + 1].energy;
		sweep_rows(g, n_iso, n_gp, U, 0, n_u, idx_low, e_high, index_grid);
		free(idx_low);
		free(e_high);
		g_last_parallel = 0;
		return;
	}
	const long rows = 4096;
	const long nblk = (n_u + rows - 1) / rows;
#pragma omp parallel
	{
		int *idx_low = malloc(n_iso * sizeof(int));
		double *e_high = malloc(n_iso * sizeof(double));
#pragma omp for schedule(dynamic, 1)
		for (long b = 0; b < nblk; b++) {
			long e0 = b * rows, e1 = e0 + rows < n_u ? e0 + rows : n_u;
			for (long i = 0; i < n_iso; i++) {
				/* state after row e0 - 1 (initial state for e0 == 0) */
				long l = (e0 == 0) ? 0 : bracket_lower(g + i * n_gp, n_gp, U[e0 - 1]);
				idx_low[i] = (int)l;
				e_high[i] = g[i * n_gp + l + 1].energy;
			}
			sweep_rows(g, n_iso, n_gp, U, e0, e1, idx_low, e_high, index_grid);
		}
		free(idx_low);
		free(e_high);
	}
	g_last_parallel = 1;
}

/* openmp-threading/GridInit.c:3-230 (AML replicas excluded) */
SimulationData grid_init_do_not_profile(Inputs in, int mype)
{
	SimulationData SD;
	memset(&SD, 0, sizeof(SD));
	size_t nbytes = 0;
	const long n_iso = in.n_isotopes, n_gp = in.n_gridpoints, n_pts = n_iso * n_gp;

	if (mype == 0)
		printf("Intializing nuclide grids...\n");
	SD.length_nuclide_grid = (int)n_pts;
	SD.nuclide_grid = malloc((size_t)n_pts * sizeof(NuclideGridPoint));
	if (!SD.nuclide_grid) {
		fprintf(stderr, "grid_init: nuclide grid allocation failed\n");
		exit(1);
	}
	nbytes += (size_t)n_pts * sizeof(NuclideGridPoint);
	build_nuclide_grid(SD.nuclide_grid, n_pts);
#pragma omp parallel for schedule(dynamic, 1)
	for (long i = 0; i < n_iso; i++)
		qsort(&SD.nuclide_grid[i * n_gp], n_gp, sizeof(NuclideGridPoint), NGP_compare);

	if (in.grid_type == UNIONIZED) {
		if (mype == 0)
			printf("Intializing unionized grid...\n");
		SD.length_unionized_energy_array = (int)n_pts;
		SD.unionized_energy_array = malloc((size_t)n_pts * sizeof(double));
		SD.length_index_grid = n_pts * n_iso;
		SD.index_grid = malloc((size_t)SD.length_index_grid * sizeof(int));
		if (!SD.unionized_energy_array || !SD.index_grid) {
			fprintf(stderr, "grid_init: unionized grid allocation failed\n");
			exit(1);
		}
		nbytes += (size_t)n_pts * sizeof(double) + (size_t)SD.length_index_grid * sizeof(int);
		build_unionized(SD.nuclide_grid, n_iso, n_gp, SD.unionized_energy_array);
		build_index_grid(SD.nuclide_grid, n_iso, n_gp, SD.unionized_energy_array, n_pts,
		                 SD.index_grid);
	} else if (in.grid_type == HASH) {
		if (mype == 0)
			printf("Intializing hash grid...\n");
		SD.length_index_grid = (long)in.hash_bins * n_iso;
		SD.index_grid = malloc((size_t)SD.length_index_grid * sizeof(int));
		if (!SD.index_grid) {
			fprintf(stderr, "grid_init: hash grid allocation failed\n");
			exit(1);
		}
		nbytes += (size_t)SD.length_index_grid * sizeof(int);
		const double du = 1.0 / in.hash_bins;
#pragma omp parallel for schedule(static)
		for (long e = 0; e < in.hash_bins; e++) {
			const double energy = e * du;
			for (long i = 0; i < n_iso; i++)
				SD.index_grid[e * n_iso + i] =
				    (int)bracket_lower(SD.nuclide_grid + i * n_gp, n_gp, energy);
		}
	}

	if (mype == 0)
		printf("Intializing material data...\n");
	SD.num_nucs = load_num_nucs(n_iso);
	SD.length_num_nucs = 12;
	SD.mats = load_mats(SD.num_nucs, n_iso, &SD.max_num_nucs);
	SD.length_mats = SD.length_num_nucs * SD.max_num_nucs;
	SD.concs = load_concs(SD.num_nucs, SD.max_num_nucs);
	SD.length_concs = SD.length_mats;

	if (mype == 0)
		printf("Intialization complete. Allocated %.0lf MB of data.\n",
		       nbytes / 1024.0 / 1024.0);
	return SD;
}
