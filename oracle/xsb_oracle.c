/*
 * xsb_oracle.c -- CPU restatement of the XSBench lookup path (test
 * infrastructure only: it checks the HIP product and is never the thing
 * measured or shipped).
 *
 * Restates openmp-threading/{Simulation.c, GridInit.c, Materials.c,
 * XSutils.c} of the reference; every function cites the lines it follows.
 * Compiled with gcc -O2 -fopenmp -ffp-contract=off and no -ffast-math so that
 * each floating-point operation is the IEEE double operation the reference
 * performs, in the reference's order.
 */
#include "xsb_oracle.h"

#include <assert.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define LCG_M 9223372036854775808ULL /* 2^63 */
#define LCG_A 2806196910506780709ULL
#define LCG_C 1ULL

/* openmp-threading/Simulation.c:462-470 */
double or_lcg_random_double(uint64_t *seed)
{
	*seed = (LCG_A * (*seed) + LCG_C) % LCG_M;
	return (double)(*seed) / (double)LCG_M;
}

/* openmp-threading/Simulation.c:472-499 (square-and-multiply skip-ahead) */
uint64_t or_fast_forward_lcg(uint64_t seed, uint64_t n)
{
	uint64_t a = LCG_A, c = LCG_C;
	uint64_t a_acc = 1, c_acc = 0;
	n %= LCG_M;
	for (; n > 0; n >>= 1) {
		if (n & 1) {
			a_acc *= a;
			c_acc = c_acc * a + c;
		}
		c *= (a + 1);
		a *= a;
	}
	return (a_acc * seed + c_acc) % LCG_M;
}

/* Volume fractions of the 12 materials, openmp-threading/Simulation.c:433-445 */
static const double mat_dist[12] = {0.140, 0.052, 0.275, 0.134, 0.154, 0.064,
                                    0.066, 0.055, 0.008, 0.015, 0.025, 0.013};

/* openmp-threading/Simulation.c:426-460: material i is picked when roll is
 * below dist[i] + dist[i-1] + ... + dist[1], summed from j = i downwards;
 * material 0 is the fall-through. */
int or_pick_mat(uint64_t *seed)
{
	double roll = or_lcg_random_double(seed);
	for (int i = 0; i < 12; i++) {
		double running = 0;
		for (int j = i; j > 0; j--)
			running += mat_dist[j];
		if (roll < running)
			return i;
	}
	return 0;
}

void or_pick_mat_thresholds(double thr[12])
{
	for (int i = 0; i < 12; i++) {
		double running = 0;
		for (int j = i; j > 0; j--)
			running += mat_dist[j];
		thr[i] = running;
	}
}

/* openmp-threading/Simulation.c:380-400 */
long or_grid_search(long n, double quarry, const double *A)
{
	long lo = 0, hi = n - 1;
	while (hi - lo > 1) {
		long mid = lo + (hi - lo) / 2;
		if (A[mid] > quarry)
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

/* openmp-threading/Simulation.c:403-423 */
long or_grid_search_nuclide(long n, double quarry, const NuclideGridPoint *A, long low, long high)
{
	(void)n;
	long lo = low, hi = high;
	while (hi - lo > 1) {
		long mid = lo + (hi - lo) / 2;
		if (A[mid].energy > quarry)
			hi = mid;
		else
			lo = mid;
	}
	return lo;
}

/* openmp-threading/Simulation.c:241-323 */
void or_calculate_micro_xs(double p_energy, int nuc, long n_isotopes, long n_gridpoints,
                           const double *egrid, const int *index_data,
                           const NuclideGridPoint *nuclide_grids, long idx,
                           double *xs_vector, int grid_type, int hash_bins)
{
	(void)egrid;
	const NuclideGridPoint *row = &nuclide_grids[nuc * n_gridpoints];
	long lower;
	if (grid_type == OR_NUCLIDE) {
		/* :252-263 */
		lower = or_grid_search_nuclide(n_gridpoints, p_energy, row, 0, n_gridpoints - 1);
	} else if (grid_type == OR_UNIONIZED) {
		/* :264-272 */
		lower = index_data[idx * n_isotopes + nuc];
	} else {
		/* :273-302 hash bins; the p <= e_low branch sets lower to 0 */
		int u_low = index_data[idx * n_isotopes + nuc];
		int u_high;
		if (idx == hash_bins - 1)
			u_high = (int)(n_gridpoints - 1);
		else
			u_high = index_data[(idx + 1) * n_isotopes + nuc] + 1;
		double e_low = row[u_low].energy;
		double e_high = row[u_high].energy;
		if (p_energy <= e_low)
			lower = 0;
		else if (p_energy >= e_high)
			lower = n_gridpoints - 1;
		else
			lower = or_grid_search_nuclide(n_gridpoints, p_energy, row, u_low, u_high);
	}
	if (lower == n_gridpoints - 1)
		lower--;
	const NuclideGridPoint *low = &row[lower];
	const NuclideGridPoint *high = low + 1;

	/* :306-322 */
	double f = (high->energy - p_energy) / (high->energy - low->energy);
	xs_vector[0] = high->total_xs - f * (high->total_xs - low->total_xs);
	xs_vector[1] = high->elastic_xs - f * (high->elastic_xs - low->elastic_xs);
	xs_vector[2] = high->absorbtion_xs - f * (high->absorbtion_xs - low->absorbtion_xs);
	xs_vector[3] = high->fission_xs - f * (high->fission_xs - low->fission_xs);
	xs_vector[4] = high->nu_fission_xs - f * (high->nu_fission_xs - low->nu_fission_xs);
}

/* openmp-threading/Simulation.c:326-375 */
void or_calculate_macro_xs(double p_energy, int mat, long n_isotopes, long n_gridpoints,
                           const int *num_nucs, const double *concs, const double *egrid,
                           const int *index_data, const NuclideGridPoint *nuclide_grids,
                           const int *mats, double *macro_xs_vector, int grid_type,
                           int hash_bins, int max_num_nucs)
{
	long idx = -1;
	for (int k = 0; k < 5; k++)
		macro_xs_vector[k] = 0;
	if (grid_type == OR_UNIONIZED) {
		idx = or_grid_search(n_isotopes * n_gridpoints, p_energy, egrid);
	} else if (grid_type == OR_HASH) {
		double du = 1.0 / hash_bins;
		idx = p_energy / du;
	}
	for (int j = 0; j < num_nucs[mat]; j++) {
		double xs[5];
		int p_nuc = mats[mat * max_num_nucs + j];
		double conc = concs[mat * max_num_nucs + j];
		or_calculate_micro_xs(p_energy, p_nuc, n_isotopes, n_gridpoints, egrid, index_data,
		                      nuclide_grids, idx, xs, grid_type, hash_bins);
		for (int k = 0; k < 5; k++)
			macro_xs_vector[k] += xs[k] * conc;
	}
}

/* openmp-threading/Simulation.c:100-110: max starts at -1.0, strict '>' */
int or_argmax5(const double m[5])
{
	double max = -1.0;
	int max_idx = 0;
	for (int j = 0; j < 5; j++) {
		if (m[j] > max) {
			max = m[j];
			max_idx = j;
		}
	}
	return max_idx;
}

/* ------------------------------------------------------------------------ */
/* Data construction                                                         */
/* ------------------------------------------------------------------------ */

/* openmp-threading/XSutils.c:3-14 */
static int cmp_double(const void *a, const void *b)
{
	double A = *(const double *)a, B = *(const double *)b;
	return (A > B) - (A < B);
}

/* openmp-threading/XSutils.c:16-27 */
static int cmp_point(const void *a, const void *b)
{
	double A = ((const NuclideGridPoint *)a)->energy;
	double B = ((const NuclideGridPoint *)b)->energy;
	return (A > B) - (A < B);
}

/* Material compositions, openmp-threading/Materials.c:7-96.  Row m lists the
 * nuclide ids of material m; the large problem's fuel appends ids 68..354
 * after the 34 small-problem fuel nuclides (Materials.c:54-55). */
static const int fuel_small[34] = {58, 59, 60, 61, 40, 42, 43, 44, 45, 46, 1,  2,
                                   3,  7,  8,  9,  10, 29, 57, 47, 48, 0,  62, 15,
                                   33, 34, 52, 53, 54, 55, 56, 18, 23, 41};
static const int cladding[5] = {63, 64, 65, 66, 67};
static const int water[4] = {24, 41, 4, 5};
static const int rpv[27] = {19, 20, 21, 22, 35, 36, 37, 38, 39, 25, 27, 28, 29, 30,
                            31, 32, 26, 49, 50, 51, 11, 12, 13, 14, 6,  16, 17};
static const int reflector[21] = {24, 41, 4,  5,  19, 20, 21, 22, 35, 36, 37,
                                  38, 39, 25, 49, 50, 51, 11, 12, 13, 14};
static const int assembly_end[9] = {24, 41, 4, 5, 63, 64, 65, 66, 67};

static void fill_materials(long n_isotopes, SimulationData *SD)
{
	static const int counts_rest[11] = {5, 4, 4, 27, 21, 21, 21, 21, 21, 9, 9};
	static const int *lists_rest[11] = {cladding,  water,     water,     rpv,
	                                    reflector, reflector, reflector, reflector,
	                                    reflector, assembly_end, assembly_end};
	int *num_nucs = malloc(12 * sizeof(int));
	num_nucs[0] = (n_isotopes == 68) ? 34 : 321; /* Materials.c:7-31 */
	for (int m = 1; m < 12; m++)
		num_nucs[m] = counts_rest[m - 1];
	int max_n = 0;
	for (int m = 0; m < 12; m++)
		if (num_nucs[m] > max_n)
			max_n = num_nucs[m];
	int *mats = calloc((size_t)12 * max_n, sizeof(int));
	for (int j = 0; j < num_nucs[0]; j++)
		mats[j] = (j < 34) ? fuel_small[j] : 68 + (j - 34);
	for (int m = 1; m < 12; m++)
		memcpy(mats + (size_t)m * max_n, lists_rest[m - 1], num_nucs[m] * sizeof(int));

	/* Materials.c:99-116: one LCG stream seeded 1070^2, material-major */
	double *concs = calloc((size_t)12 * max_n, sizeof(double));
	uint64_t seed = OR_STARTING_SEED * OR_STARTING_SEED;
	for (int m = 0; m < 12; m++)
		for (int j = 0; j < num_nucs[m]; j++)
			concs[m * max_n + j] = or_lcg_random_double(&seed);

	SD->num_nucs = num_nucs;
	SD->length_num_nucs = 12;
	SD->mats = mats;
	SD->max_num_nucs = max_n;
	SD->length_mats = 12 * max_n;
	SD->concs = concs;
	SD->length_concs = 12 * max_n;
}

/* openmp-threading/GridInit.c:3-230 (AML replicas excluded) */
int or_grid_init(const Inputs *in, SimulationData *SD)
{
	memset(SD, 0, sizeof(*SD));
	long n_iso = in->n_isotopes, n_gp = in->n_gridpoints;

	/* GridInit.c:35-47: one LCG stream seeded 42, six draws per point */
	SD->length_nuclide_grid = (int)(n_iso * n_gp);
	SD->nuclide_grid = malloc((size_t)n_iso * n_gp * sizeof(NuclideGridPoint));
	if (!SD->nuclide_grid)
		return 1;
	uint64_t seed = 42;
	for (long i = 0; i < n_iso * n_gp; i++) {
		NuclideGridPoint *p = &SD->nuclide_grid[i];
		p->energy = or_lcg_random_double(&seed);
		p->total_xs = or_lcg_random_double(&seed);
		p->elastic_xs = or_lcg_random_double(&seed);
		p->absorbtion_xs = or_lcg_random_double(&seed);
		p->fission_xs = or_lcg_random_double(&seed);
		p->nu_fission_xs = or_lcg_random_double(&seed);
	}
	/* GridInit.c:50-51 */
	for (long i = 0; i < n_iso; i++)
		qsort(&SD->nuclide_grid[i * n_gp], n_gp, sizeof(NuclideGridPoint), cmp_point);

	if (in->grid_type == OR_UNIONIZED) {
		/* GridInit.c:74-89 */
		long n_u = n_iso * n_gp;
		SD->length_unionized_energy_array = (int)n_u;
		SD->unionized_energy_array = malloc(n_u * sizeof(double));
		if (!SD->unionized_energy_array)
			return 1;
		for (long i = 0; i < n_u; i++)
			SD->unionized_energy_array[i] = SD->nuclide_grid[i].energy;
		qsort(SD->unionized_energy_array, n_u, sizeof(double), cmp_double);

		/* GridInit.c:92-122: the serial sweep; each nuclide's lower index
		 * advances by at most one per unionized energy. */
		SD->length_index_grid = n_u * n_iso;
		SD->index_grid = malloc((size_t)n_u * n_iso * sizeof(int));
		if (!SD->index_grid)
			return 1;
		int *idx_low = calloc(n_iso, sizeof(int));
		double *e_high = malloc(n_iso * sizeof(double));
		for (long i = 0; i < n_iso; i++)
			e_high[i] = SD->nuclide_grid[i * n_gp + 1].energy;
		for (long e = 0; e < n_u; e++) {
			double u = SD->unionized_energy_array[e];
			int *out = &SD->index_grid[e * n_iso];
			for (long i = 0; i < n_iso; i++) {
				if (!(u < e_high[i]) && idx_low[i] != n_gp - 2) {
					idx_low[i]++;
					e_high[i] = SD->nuclide_grid[i * n_gp + idx_low[i] + 1].energy;
				}
				out[i] = idx_low[i];
			}
		}
		free(idx_low);
		free(e_high);
	} else if (in->grid_type == OR_HASH) {
		/* GridInit.c:128-151 */
		SD->length_index_grid = (long)in->hash_bins * n_iso;
		SD->index_grid = malloc(SD->length_index_grid * sizeof(int));
		if (!SD->index_grid)
			return 1;
		double du = 1.0 / in->hash_bins;
		for (long e = 0; e < in->hash_bins; e++) {
			double energy = e * du;
			for (long i = 0; i < n_iso; i++)
				SD->index_grid[e * n_iso + i] = (int)or_grid_search_nuclide(
				    n_gp, energy, SD->nuclide_grid + i * n_gp, 0, n_gp - 1);
		}
	}
	fill_materials(n_iso, SD);
	return 0;
}

void or_free_simulation_data(SimulationData *SD)
{
	free(SD->num_nucs);
	free(SD->concs);
	free(SD->mats);
	free(SD->unionized_energy_array);
	free(SD->index_grid);
	free(SD->nuclide_grid);
	memset(SD, 0, sizeof(*SD));
}

uint64_t or_fnv1a(const void *p, size_t nbytes, uint64_t h)
{
	const unsigned char *c = p;
	if (h == 0)
		h = 1469598103934665603ULL;
	for (size_t i = 0; i < nbytes; i++) {
		h ^= c[i];
		h *= 1099511628211ULL;
	}
	return h;
}

/* ------------------------------------------------------------------------ */
/* Drivers                                                                   */
/* ------------------------------------------------------------------------ */

static inline int lookup_one(const Inputs *in, const SimulationData *SD, double E, int mat)
{
	double macro[5];
	or_calculate_macro_xs(E, mat, in->n_isotopes, in->n_gridpoints, SD->num_nucs, SD->concs,
	                      SD->unionized_energy_array, SD->index_grid, SD->nuclide_grid,
	                      SD->mats, macro, in->grid_type, in->hash_bins, SD->max_num_nucs);
	return or_argmax5(macro);
}

/* openmp-threading/Simulation.c:43-111, over lookups [first, first+count) */
uint64_t or_run_event_range(const Inputs *in, const SimulationData *SD, long first, long count,
                            int nthreads)
{
	unsigned long long verification = 0;
	if (nthreads < 1)
		nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : verification) num_threads(nthreads)
	for (long i = first; i < first + count; i++) {
		uint64_t seed = or_fast_forward_lcg(OR_STARTING_SEED, 2 * (uint64_t)i);
		double E = or_lcg_random_double(&seed);
		int mat = or_pick_mat(&seed);
		verification += lookup_one(in, SD, E, mat) + 1;
	}
	return verification;
}

uint64_t or_run_event(const Inputs *in, const SimulationData *SD, int nthreads)
{
	return or_run_event_range(in, SD, 0, in->lookups, nthreads);
}

/* The particle's starting offset p*lookups*2*5 is evaluated in the
 * reference's 32-bit int arithmetic (Simulation.c:167). */
static uint64_t history_offset(long p, int lookups)
{
	uint32_t prod = (uint32_t)p * (uint32_t)lookups * 10u;
	return (uint64_t)(int64_t)(int32_t)prod;
}

/* openmp-threading/Simulation.c:141-237 for one particle */
static uint64_t history_particle(const Inputs *in, const SimulationData *SD, long p,
                                 uint64_t *final_seed)
{
	uint64_t seed = or_fast_forward_lcg(OR_STARTING_SEED, history_offset(p, in->lookups));
	double E = or_lcg_random_double(&seed);
	int mat = or_pick_mat(&seed);
	uint64_t v = 0;
	for (int i = 0; i < in->lookups; i++) {
		double macro[5];
		or_calculate_macro_xs(E, mat, in->n_isotopes, in->n_gridpoints, SD->num_nucs,
		                      SD->concs, SD->unionized_energy_array, SD->index_grid,
		                      SD->nuclide_grid, SD->mats, macro, in->grid_type,
		                      in->hash_bins, SD->max_num_nucs);
		v += or_argmax5(macro) + 1;
		uint64_t n_forward = 0;
		for (int j = 0; j < 5; j++)
			if (macro[j] > 1.0)
				n_forward++;
		if (n_forward > 0)
			seed = or_fast_forward_lcg(seed, n_forward);
		E = or_lcg_random_double(&seed);
		mat = or_pick_mat(&seed);
	}
	if (final_seed)
		*final_seed = seed;
	return v;
}

uint64_t or_run_history(const Inputs *in, const SimulationData *SD, int nthreads)
{
	unsigned long long verification = 0;
	if (nthreads < 1)
		nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : verification) num_threads(nthreads)
	for (long p = 0; p < in->particles; p++)
		verification += history_particle(in, SD, p, NULL);
	return verification;
}

void or_history_detail(const Inputs *in, const SimulationData *SD, long first, long n,
                       uint64_t *per_particle_sum, uint64_t *final_seed)
{
#pragma omp parallel for schedule(dynamic, 16)
	for (long k = 0; k < n; k++) {
		uint64_t fs;
		uint64_t v = history_particle(in, SD, first + k, &fs);
		if (per_particle_sum)
			per_particle_sum[k] = v;
		if (final_seed)
			final_seed[k] = fs;
	}
}

/* openmp-threading/Simulation.c:698-871: sample every lookup, order the
 * samples by material and then by energy, and run the lookups material by
 * material.  The sort here is a plain qsort of (mat, energy) records; the
 * verification sum does not depend on the order of equal keys. */
typedef struct {
	int mat;
	double energy;
} sample_rec;

static int cmp_sample(const void *a, const void *b)
{
	const sample_rec *A = a, *B = b;
	if (A->mat != B->mat)
		return (A->mat > B->mat) - (A->mat < B->mat);
	return (A->energy > B->energy) - (A->energy < B->energy);
}

uint64_t or_run_event_sorted(const Inputs *in, const SimulationData *SD, int nthreads)
{
	long L = in->lookups;
	if (nthreads < 1)
		nthreads = omp_get_max_threads();
	sample_rec *s = malloc(L * sizeof(sample_rec));
#pragma omp parallel for schedule(dynamic, 100) num_threads(nthreads)
	for (long i = 0; i < L; i++) {
		uint64_t seed = or_fast_forward_lcg(OR_STARTING_SEED, 2 * (uint64_t)i);
		s[i].energy = or_lcg_random_double(&seed);
		s[i].mat = or_pick_mat(&seed);
	}
	qsort(s, L, sizeof(sample_rec), cmp_sample);
	unsigned long long verification = 0;
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : verification) num_threads(nthreads)
	for (long i = 0; i < L; i++)
		verification += lookup_one(in, SD, s[i].energy, s[i].mat) + 1;
	free(s);
	return verification;
}

void or_event_detail(const Inputs *in, const SimulationData *SD, long n, const long *ids,
                     double *energy, int *mat, long *ueg_idx, double *macro, int *argmax)
{
#pragma omp parallel for schedule(dynamic, 16)
	for (long k = 0; k < n; k++) {
		uint64_t seed = or_fast_forward_lcg(OR_STARTING_SEED, 2 * (uint64_t)ids[k]);
		double E = or_lcg_random_double(&seed);
		int m = or_pick_mat(&seed);
		double mx[5];
		or_calculate_macro_xs(E, m, in->n_isotopes, in->n_gridpoints, SD->num_nucs, SD->concs,
		                      SD->unionized_energy_array, SD->index_grid, SD->nuclide_grid,
		                      SD->mats, mx, in->grid_type, in->hash_bins, SD->max_num_nucs);
		if (energy)
			energy[k] = E;
		if (mat)
			mat[k] = m;
		if (ueg_idx)
			ueg_idx[k] = (in->grid_type == OR_UNIONIZED)
			                 ? or_grid_search(in->n_isotopes * in->n_gridpoints, E,
			                                  SD->unionized_energy_array)
			                 : -1;
		if (macro)
			memcpy(macro + 5 * k, mx, sizeof(mx));
		if (argmax)
			argmax[k] = or_argmax5(mx);
	}
}

void or_lookup_detail(const Inputs *in, const SimulationData *SD, long n, const double *energy,
                      const int *mat, double *macro, int *argmax)
{
#pragma omp parallel for schedule(dynamic, 16)
	for (long k = 0; k < n; k++) {
		double mx[5];
		or_calculate_macro_xs(energy[k], mat[k], in->n_isotopes, in->n_gridpoints,
		                      SD->num_nucs, SD->concs, SD->unionized_energy_array,
		                      SD->index_grid, SD->nuclide_grid, SD->mats, mx, in->grid_type,
		                      in->hash_bins, SD->max_num_nucs);
		if (macro)
			memcpy(macro + 5 * k, mx, sizeof(mx));
		if (argmax)
			argmax[k] = or_argmax5(mx);
	}
}

void or_micro_detail(const Inputs *in, const SimulationData *SD, double p_energy, int mat,
                     double *micro)
{
	long idx = -1;
	if (in->grid_type == OR_UNIONIZED)
		idx = or_grid_search(in->n_isotopes * in->n_gridpoints, p_energy,
		                     SD->unionized_energy_array);
	else if (in->grid_type == OR_HASH) {
		double du = 1.0 / in->hash_bins;
		idx = p_energy / du;
	}
	for (int j = 0; j < SD->num_nucs[mat]; j++)
		or_calculate_micro_xs(p_energy, SD->mats[mat * SD->max_num_nucs + j], in->n_isotopes,
		                      in->n_gridpoints, SD->unionized_energy_array, SD->index_grid,
		                      SD->nuclide_grid, idx, micro + 5 * j, in->grid_type,
		                      in->hash_bins);
}

/* ------------------------------------------------------------------------ */
/* Batched workload (no counterpart in the reference: a set of independent   */
/* lookup batches, each with its own material CDF and energy stream).        */
/* ------------------------------------------------------------------------ */

void or_batch_plan(uint64_t seed, int n_batches, long lo, long hi, long *sizes,
                   uint64_t *batch_seeds, double *batch_thresholds)
{
	uint64_t s = seed;
	double llo = log((double)lo), lhi = log((double)hi);
	for (int b = 0; b < n_batches; b++) {
		double u = or_lcg_random_double(&s);
		long size = lround(exp(llo + u * (lhi - llo)));
		if (size < lo)
			size = lo;
		if (size > hi)
			size = hi;
		sizes[b] = size;
		or_lcg_random_double(&s);
		batch_seeds[b] = s;
		double w[12], W = 0;
		for (int m = 0; m < 12; m++) {
			w[m] = 0.05 + or_lcg_random_double(&s);
			W += w[m];
		}
		double run = 0;
		for (int m = 0; m < 12; m++) {
			run += w[m];
			batch_thresholds[b * 12 + m] = run / W;
		}
		batch_thresholds[b * 12 + 11] = 2.0; /* every roll < 1 picks a material */
	}
}

void or_batch_sample(uint64_t batch_seed, const double thr[12], long i, double *energy, int *mat)
{
	uint64_t s = or_fast_forward_lcg(batch_seed, 2 * (uint64_t)i);
	*energy = or_lcg_random_double(&s);
	double roll = or_lcg_random_double(&s);
	int m = 0;
	while (m < 11 && !(roll < thr[m]))
		m++;
	*mat = m;
}

uint64_t or_run_batch(const Inputs *in, const SimulationData *SD, uint64_t batch_seed,
                      const double thr[12], long size, int nthreads)
{
	unsigned long long v = 0;
	if (nthreads < 1)
		nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : v) num_threads(nthreads)
	for (long i = 0; i < size; i++) {
		double E;
		int m;
		or_batch_sample(batch_seed, thr, i, &E, &m);
		v += lookup_one(in, SD, E, m) + 1;
	}
	return v;
}

/* ------------------------------------------------------------------------ */
/* Algorithmic-byte counter                                                  */
/* ------------------------------------------------------------------------ */

/* number of probes (distinct words read) of the reference's bisection over
 * [low, high] (Simulation.c:380-423) */
static int bisect_probes(const double *E, long stride_words, long low, long high, double q)
{
	int probes = 0;
	long lo = low, hi = high;
	while (hi - lo > 1) {
		long mid = lo + (hi - lo) / 2;
		probes++;
		if (E[mid * stride_words] > q)
			hi = mid;
		else
			lo = mid;
	}
	return probes;
}

/* Mean stored size of the distinct words one lookup reads.
 * fmt[0] energy word, fmt[1] xs block of one point, fmt[2] index word,
 * fmt[3] concentration word, fmt[4] material-list word. */
double or_balg_event(const Inputs *in, const SimulationData *SD, long first, long count,
                     const int fmt[5], int nthreads)
{
	const long n_gp = in->n_gridpoints, n_iso = in->n_isotopes;
	const double *E_aos = &SD->nuclide_grid[0].energy;
	const long stride = sizeof(NuclideGridPoint) / sizeof(double);
	double total = 0;
	if (nthreads < 1)
		nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total) num_threads(nthreads)
	for (long i = first; i < first + count; i++) {
		uint64_t seed = or_fast_forward_lcg(OR_STARTING_SEED, 2 * (uint64_t)i);
		double p = or_lcg_random_double(&seed);
		int mat = or_pick_mat(&seed);
		int nn = SD->num_nucs[mat];
		double b = 4; /* num_nucs[mat] */
		b += (double)nn * (fmt[3] + fmt[4]);
		long idx = -1;
		if (in->grid_type == OR_UNIONIZED) {
			b += 8.0 * bisect_probes(SD->unionized_energy_array, 1, 0, n_iso * n_gp - 1, p);
			idx = or_grid_search(n_iso * n_gp, p, SD->unionized_energy_array);
		} else if (in->grid_type == OR_HASH) {
			double du = 1.0 / in->hash_bins;
			idx = p / du;
		}
		for (int j = 0; j < nn; j++) {
			int nuc = SD->mats[mat * SD->max_num_nucs + j];
			const double *Er = E_aos + nuc * n_gp * stride;
			long lower;
			if (in->grid_type == OR_UNIONIZED) {
				b += fmt[2];
				lower = SD->index_grid[idx * n_iso + nuc];
				if (lower == n_gp - 1)
					lower--;
				b += 2.0 * (fmt[0] + fmt[1]);
			} else if (in->grid_type == OR_NUCLIDE) {
				/* probes lie strictly inside (0, n-1); the bracket points
				 * may coincide with probed energies */
				lower = or_grid_search_nuclide(n_gp, p, SD->nuclide_grid + nuc * n_gp, 0,
				                               n_gp - 1);
				int probes = bisect_probes(Er, stride, 0, n_gp - 1, p);
				/* the final bracket [lower, lower+1]: energies of the
				 * interior end points were probed unless they are 0 / n-1 */
				int seen = (lower > 0) + (lower + 1 < n_gp - 1);
				b += (double)probes * fmt[0] + 2.0 * (fmt[0] + fmt[1]) - seen * fmt[0];
			} else {
				int u_low = SD->index_grid[idx * n_iso + nuc];
				int u_high;
				b += fmt[2];
				if (idx == in->hash_bins - 1) {
					u_high = (int)(n_gp - 1);
				} else {
					u_high = SD->index_grid[(idx + 1) * n_iso + nuc] + 1;
					b += fmt[2];
				}
				/* words touched: E[u_low], E[u_high], bisection probes in
				 * (u_low, u_high), then the full bracket points */
				long words_lo = u_low, words_hi = u_high;
				double e_low = Er[u_low * stride], e_high = Er[u_high * stride];
				int probes = 0;
				if (p <= e_low)
					lower = 0;
				else if (p >= e_high)
					lower = n_gp - 1;
				else {
					lower = or_grid_search_nuclide(n_gp, p, SD->nuclide_grid + nuc * n_gp,
					                               u_low, u_high);
					probes = bisect_probes(Er, stride, u_low, u_high, p);
				}
				if (lower == n_gp - 1)
					lower--;
				/* energies already counted: u_low, u_high and the probes;
				 * the bracket energies are free when they were probed */
				double eb = 2.0 * fmt[0] + probes * fmt[0];
				int lo_seen = (lower == words_lo) || (lower == words_hi) ||
				              (lower > words_lo && lower < words_hi && probes > 0);
				int hi_seen = (lower + 1 == words_lo) || (lower + 1 == words_hi) ||
				              (lower + 1 > words_lo && lower + 1 < words_hi && probes > 0);
				eb += (!lo_seen) * fmt[0] + (!hi_seen) * fmt[0];
				b += eb + 2.0 * fmt[1];
			}
			(void)lower;
		}
		total += b;
	}
	return count > 0 ? total / count : 0.0;
}
