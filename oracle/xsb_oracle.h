/*
 * xsb_oracle.h -- CPU restatement of XSBench's macroscopic cross-section
 * lookup, used ONLY to check the HIP product (tests, smoke()) and as the
 * cpu_baseline leg of bench.py.  Nothing in the product links it.
 *
 * Every function follows openmp-threading in the reference; each definition
 * in xsb_oracle.c cites the file:line it restates.  The structs keep the
 * reference's exact memory layout (openmp-threading/XSbench_header.h:54-102,
 * non-AML build) so that one ctypes description serves this library, the
 * reference build in oracle/_ref/ and the host surface of the product.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   LCG / fast-forward ........ brute-force stepping + Python big-int closed form
 *   pick_mat .................. reference .so on 10^5 seeds + frequency check
 *   grid_search(_nuclide) ..... brute-force clamped upper_bound incl. ties
 *   grid init ................. FNV-1a digests of every array vs reference .so
 *   micro / macro XS .......... hand-computed toy tables + golden vectors
 *   event / history / sorted .. published hashes (io.c:89-95) + golden hashes
 *   batched sampling .......... no reference counterpart: pinned by brute
 *                               force (per-batch hash == event loop over the
 *                               same explicit (E, mat) list)
 *   B_alg counter ............. closed form for the unionized grid
 */
#ifndef XSB_ORACLE_H
#define XSB_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#define OR_UNIONIZED 0
#define OR_NUCLIDE 1
#define OR_HASH 2
#define OR_HISTORY_BASED 1
#define OR_EVENT_BASED 2
#define OR_STARTING_SEED 1070ULL

typedef struct {
	double energy;
	double total_xs;
	double elastic_xs;
	double absorbtion_xs;
	double fission_xs;
	double nu_fission_xs;
} NuclideGridPoint;

typedef struct {
	int nthreads;
	long n_isotopes;
	long n_gridpoints;
	int lookups;
	char *HM;
	int grid_type;
	int hash_bins;
	int particles;
	int simulation_method;
	int binary_mode;
	int kernel_id;
} Inputs;

typedef struct {
	int *num_nucs;
	double *concs;
	int *mats;
	double *unionized_energy_array;
	int *index_grid;
	NuclideGridPoint *nuclide_grid;
	int length_num_nucs;
	int length_concs;
	int length_mats;
	int length_unionized_energy_array;
	long length_index_grid;
	int length_nuclide_grid;
	int max_num_nucs;
	double *p_energy_samples;
	int length_p_energy_samples;
	int *mat_samples;
	int length_mat_samples;
} SimulationData;

/* sampling */
double or_lcg_random_double(uint64_t *seed);
uint64_t or_fast_forward_lcg(uint64_t seed, uint64_t n);
int or_pick_mat(uint64_t *seed);
void or_pick_mat_thresholds(double thr[12]);

/* searches */
long or_grid_search(long n, double quarry, const double *A);
long or_grid_search_nuclide(long n, double quarry, const NuclideGridPoint *A, long low, long high);

/* lookups */
void or_calculate_micro_xs(double p_energy, int nuc, long n_isotopes, long n_gridpoints,
                           const double *egrid, const int *index_data,
                           const NuclideGridPoint *nuclide_grids, long idx,
                           double *xs_vector, int grid_type, int hash_bins);
void or_calculate_macro_xs(double p_energy, int mat, long n_isotopes, long n_gridpoints,
                           const int *num_nucs, const double *concs, const double *egrid,
                           const int *index_data, const NuclideGridPoint *nuclide_grids,
                           const int *mats, double *macro_xs_vector, int grid_type,
                           int hash_bins, int max_num_nucs);
int or_argmax5(const double m[5]);

/* data construction (returns 0 on success) */
int or_grid_init(const Inputs *in, SimulationData *SD);
void or_free_simulation_data(SimulationData *SD);
uint64_t or_fnv1a(const void *p, size_t nbytes, uint64_t h);

/* drivers: return the un-modded verification sum */
uint64_t or_run_event(const Inputs *in, const SimulationData *SD, int nthreads);
uint64_t or_run_event_range(const Inputs *in, const SimulationData *SD, long first, long count,
                            int nthreads);
uint64_t or_run_history(const Inputs *in, const SimulationData *SD, int nthreads);
uint64_t or_run_event_sorted(const Inputs *in, const SimulationData *SD, int nthreads);

/* per-lookup / per-particle detail for parity tests (arrays of length n,
 * macro is n x 5; any output pointer may be NULL). ids[] selects lookups. */
void or_event_detail(const Inputs *in, const SimulationData *SD, long n, const long *ids,
                     double *energy, int *mat, long *ueg_idx, double *macro, int *argmax);
void or_lookup_detail(const Inputs *in, const SimulationData *SD, long n, const double *energy,
                      const int *mat, double *macro, int *argmax);
void or_micro_detail(const Inputs *in, const SimulationData *SD, double p_energy, int mat,
                     double *micro /* num_nucs[mat] x 5 */);
void or_history_detail(const Inputs *in, const SimulationData *SD, long first, long n,
                       uint64_t *per_particle_sum, uint64_t *final_seed);

/* batched workload: batch b of a seeded set of independent lookup batches.
 * sizes[b] = round(exp(U(ln lo, ln hi))); each batch has its own material
 * CDF (12 weights) and its own energy stream.  See xsb_oracle.c. */
void or_batch_plan(uint64_t seed, int n_batches, long lo, long hi, long *sizes,
                   uint64_t *batch_seeds, double *batch_thresholds /* n x 12 */);
void or_batch_sample(uint64_t batch_seed, const double thr[12], long i, double *energy, int *mat);
uint64_t or_run_batch(const Inputs *in, const SimulationData *SD, uint64_t batch_seed,
                      const double thr[12], long size, int nthreads);

/* algorithmic bytes: mean over lookups [first, first+count) of the stored
 * size of the distinct words the reference algorithm reads per lookup, for
 * a table format given by word sizes (bytes).  fmt = {energy word, xs block
 * per point, index word, conc word, mats word}. */
double or_balg_event(const Inputs *in, const SimulationData *SD, long first, long count,
                     const int fmt[5], int nthreads);

#endif
