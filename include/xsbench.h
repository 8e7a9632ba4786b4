/*
 * xsbench.h -- host surface of the MI355X XSBench lookup library
 * (libxsbench.so).  Keeps the reference's C interface: the same constants,
 * the same struct layouts and the same function signatures as
 * openmp-threading/XSbench_header.h:104-150, every one of which this library
 * defines, so that code written against the reference (its Main.c included)
 * links against this library unchanged.  The lookup drivers (run_*) and
 * calculate_micro_xs / calculate_macro_xs execute on the GPU through
 * libxsb_hip.so (include/xsb_hip.h); there is no CPU implementation of the
 * lookup in this library.  grid_search, grid_search_nuclide, pick_mat and
 * the LCG functions are the reference's host helpers, defined here with its
 * semantics; the GPU path does not call them.
 */
#ifndef XSBENCH_H
#define XSBENCH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid types (openmp-threading/XSbench_header.h:37-39) */
#define UNIONIZED 0
#define NUCLIDE 1
#define HASH 2

/* Simulation types (XSbench_header.h:42-43) */
#define HISTORY_BASED 1
#define EVENT_BASED 2

/* Binary mode (XSbench_header.h:46-48) */
#define NONE 0
#define READ 1
#define WRITE 2

/* Starting seed (XSbench_header.h:51) */
#define STARTING_SEED 1070

/* XSbench_header.h:54-61 -- one point of a nuclide's energy grid */
typedef struct {
	double energy;
	double total_xs;
	double elastic_xs;
	double absorbtion_xs;
	double fission_xs;
	double nu_fission_xs;
} NuclideGridPoint;

/* XSbench_header.h:63-75 */
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

/* XSbench_header.h:77-102 (non-AML layout) */
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

/* ---- io.c (openmp-threading/io.c) ---- */
void logo(int version);                                           /* io.c:8 */
void center_print(const char *s, int width);                      /* io.c:28 */
void border_print(void);                                          /* io.c:179 */
void fancy_int(long a);                                           /* io.c:187 */
Inputs read_CLI(int argc, char *argv[]);                          /* io.c:227 */
void print_CLI_error(void);                                       /* io.c:208 */
void print_inputs(Inputs in, int nprocs, int version);            /* io.c:120 */
int print_results(Inputs in, int mype, double runtime, int nprocs,
                  unsigned long long vhash);                      /* io.c:39 */
void binary_write(Inputs in, SimulationData SD);                  /* io.c:456 */
SimulationData binary_read(Inputs in);                            /* io.c:476 */

/* ---- Simulation.c drivers (GPU-backed) ---- */
unsigned long long run_event_based_simulation(Inputs in, SimulationData SD, int mype);    /* Simulation.c:15 */
unsigned long long run_history_based_simulation(Inputs in, SimulationData SD, int mype);  /* Simulation.c:116 */
unsigned long long run_event_based_simulation_optimization_1(Inputs in, SimulationData SD,
                                                             int mype);                   /* Simulation.c:698 */
/* Per-lookup functions (Simulation.c:241-499).  calculate_micro_xs and
 * calculate_macro_xs run on the GPU: on a copy of the tables staged by
 * move_simulation_data_to_device for the same host nuclide grid (and, for
 * calculate_macro_xs, the same material arrays), else on one uploaded at the
 * first call and kept until release_device_memory; one call is one device
 * round trip (xsb_hip_macro_host / xsb_hip_micro_host take many samples). */
void calculate_micro_xs(double p_energy, int nuc, long n_isotopes, long n_gridpoints,
                        double *egrid, int *index_data, NuclideGridPoint *nuclide_grids,
                        long idx, double *xs_vector, int grid_type, int hash_bins); /* Simulation.c:241 */
void calculate_macro_xs(double p_energy, int mat, long n_isotopes, long n_gridpoints,
                        int *num_nucs, double *concs, double *egrid, int *index_data,
                        NuclideGridPoint *nuclide_grids, int *mats, double *macro_xs_vector,
                        int grid_type, int hash_bins, int max_num_nucs);        /* Simulation.c:326 */
long grid_search(long n, double quarry, double *A);                             /* Simulation.c:380 */
long grid_search_nuclide(long n, double quarry, NuclideGridPoint *A, long low, long high); /* :403 */
int pick_mat(uint64_t *seed);                                                   /* Simulation.c:426 */
double LCG_random_double(uint64_t *seed);                                       /* Simulation.c:462 */
uint64_t fast_forward_LCG(uint64_t seed, uint64_t n);                           /* Simulation.c:472 */

/* ---- GridInit.c / Materials.c / XSutils.c ---- */
SimulationData grid_init_do_not_profile(Inputs in, int mype);     /* GridInit.c:3 */
int *load_num_nucs(long n_isotopes);                              /* Materials.c:7 */
int *load_mats(int *num_nucs, long n_isotopes, int *max_num_nucs);/* Materials.c:34 */
double *load_concs(int *num_nucs, int max_num_nucs);              /* Materials.c:99 */
int NGP_compare(const void *a, const void *b);                    /* XSutils.c:16 */
int double_compare(const void *a, const void *b);                 /* XSutils.c:3 */
size_t estimate_mem_usage(Inputs in);                             /* XSutils.c:30 */
double get_time(void);                                            /* XSutils.c:49 */

/* ---- device staging, as in the reference's GPU port ---- */
/* cuda/GridInit.cu:4 -- uploads SD to the current GPU and registers the
 * device copy so that later run_* calls on SD use it.  Returns SD. */
SimulationData move_simulation_data_to_device(Inputs in, int mype, SimulationData SD);
/* cuda/GridInit.cu:81 -- frees the device copy registered for SD. */
void release_device_memory(SimulationData SD);

/* The program flow of openmp-threading/Main.c:7-123 (bin/XSBench calls it);
 * returns the process exit code: 1 when the hash is not the published one. */
int xsbench_main(int argc, char *argv[]);

/* ---- batched lookups (no reference counterpart): many independent
 * batches, each with its own sample stream and material CDF ---- */
/* The batched workload's generator: sizes log-uniform in [lo, hi], one
 * stream seed and 12 cumulative material thresholds per batch. */
void xsbench_make_batches(uint64_t seed, int n_batches, long lo, long hi, long *sizes,
                          uint64_t *batch_seeds, double *thresholds);
/* Lookup i of a batch (energy, material) from its stream. */
void xsbench_batch_sample(uint64_t batch_seed, const double *thresholds, long i, double *energy,
                          int *mat);
/* Runs every batch on the GPU; returns the un-modded sum over all batches. */
unsigned long long run_batched_simulation(Inputs in, SimulationData SD, int n_batches,
                                          const long *sizes, const uint64_t *batch_seeds,
                                          const double *thresholds, unsigned long long *out_vsum,
                                          double *out_done_ms);

/* ---- options of this implementation (no reference counterpart) ---- */
#define XSBENCH_SCHEDULE_ASKED 0 /* run_* follow in.kernel_id (-k) */
#define XSBENCH_SCHEDULE_AUTO 1  /* run_* take the fastest exact schedule whatever -k says */
typedef struct {
	int xs_precision;  /* 0 FP64 (bit-exact), 1 FP32 tables, 2 BF16 tables */
	int device;        /* HIP device ordinal */
	int verbose;       /* print per-phase GPU timings */
	double last_kernel_ms;  /* out: GPU time of the last run_* call */
	double last_lookup_ms;  /* out: GPU time of its lookup kernels */
	long last_fallbacks;    /* out: lookups re-evaluated in FP64 */
	int last_error;         /* out: 0 or an XSB_ERR_* code */
	int schedule;           /* XSBENCH_SCHEDULE_*; initially AUTO when the environment
	                           holds XSBENCH_SCHEDULE=auto, else ASKED */
	int last_kernel_id;     /* out: the schedule (-k id) the last run_* call ran */
} xsbench_options;
xsbench_options *xsbench_get_options(void);
/* parallel exact construction of the unionized index grid (returns 1 when
 * the parallel path was used, 0 when duplicate energies forced the serial
 * sweep) -- reported for tests */
int xsbench_last_init_parallel(void);

#ifdef __cplusplus
}
#endif
#endif
