/*
 * xsb_hip.h -- C ABI of libxsb_hip.so, the MI355X (gfx950) implementation of
 * XSBench's macroscopic cross-section lookup.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, opaque handles,
 * integer return codes (never exit()).  No HIP or torch types appear here;
 * a stream argument is a hipStream_t passed as void* (NULL = the handle's
 * own stream).  One handle owns one device copy of the tables and one
 * stream; handles on different devices are independent and may be used
 * from different host threads.  Calls on one handle must not overlap.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef XSB_HIP_H
#define XSB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define XSB_OK 0
#define XSB_ERR_INVALID 1     /* bad argument / inconsistent sizes */
#define XSB_ERR_NO_DEVICE 2   /* no usable HIP device */
#define XSB_ERR_OOM 3         /* device allocation failed */
#define XSB_ERR_HIP 4         /* a HIP runtime call failed */
#define XSB_ERR_UNSUPPORTED 5 /* option not available for this table set */

/* ---- grid types: same values as openmp-threading/XSbench_header.h:37-39 ---- */
#define XSB_UNIONIZED 0
#define XSB_NUCLIDE 1
#define XSB_HASH 2

/* ---- XS table precision ---- */
#define XSB_XS_FP64 0 /* reference tables, bit-exact */
#define XSB_XS_FP32 1 /* FP64 energies + FP32 XS, FP32 Kahan sums, FP64 fallback */
#define XSB_XS_BF16 2 /* FP64 energies + BF16 XS scaled per nuclide, same fallback */

/* ---- upload flags ---- */
#define XSB_BUILD_FP32 0x1u /* also build the FP32 XS table on the device */
#define XSB_BUILD_BF16 0x2u /* also build the BF16 XS table on the device */
#define XSB_BUILD_GRIDS 0x4u /* generate the nuclide grid, unionized energies and index grid
                                on the device (GridInit.c:3-151) instead of copying them:
                                only n_isotopes, n_gridpoints, grid_type, hash_bins and the
                                material tables of xsb_tables are read */
#define XSB_BUILD_IDX32 0x8u /* keep the per-material unionized index tables 32-bit; by default
                                they are 16-bit whenever n_gridpoints <= 65536 (same results,
                                half the index traffic) */

/* ---- lookup schedules: the reference's -k kernel ids
 *      (openmp-threading/Main.c:80-93, cuda/Main.cu:62-83) ---- */
#define XSB_K_BASELINE 0     /* sample + lookup fused, one lookup per lane, lookup order */
#define XSB_K_SPLIT 1        /* sampling kernel, then lookup kernel (cuda opt 1) */
#define XSB_K_MAT_LAUNCH 2   /* samples partitioned by material, one launch per material (opt 2) */
#define XSB_K_FUEL_LAUNCH 3  /* fuel / non-fuel partition, two launches (opt 3) */
#define XSB_K_MAT_SORT 4     /* material sort, per-material segments (opt 4) */
#define XSB_K_FUEL_SORT 5    /* fuel-first partition, two segments (opt 5) */
#define XSB_K_MAT_E_SORT 6   /* material then energy sort, per-material segments (opt 6) */
#define XSB_K_AUTO -1        /* fastest exact schedule for the table set */

/* Host view of the reference's SimulationData (openmp-threading/
 * XSbench_header.h:77-102) plus the Inputs fields the lookup needs. */
typedef struct {
	int64_t n_isotopes;
	int64_t n_gridpoints;
	int32_t grid_type;
	int32_t hash_bins;
	const int32_t *num_nucs; /* 12 */
	const int32_t *mats;     /* 12 x max_num_nucs */
	const double *concs;     /* 12 x max_num_nucs */
	int32_t max_num_nucs;
	int32_t reserved0;
	const double *unionized_energy_array; /* length_unionized_energy_array */
	int64_t length_unionized_energy_array;
	const int32_t *index_grid; /* length_index_grid */
	int64_t length_index_grid;
	const double *nuclide_grid; /* n_isotopes x n_gridpoints points of 6 doubles */
	int64_t length_nuclide_grid;
} xsb_tables;

/* Event-based run (openmp-threading/Simulation.c:15-114) */
typedef struct {
	int64_t lookups;      /* number of lookups */
	int64_t first_lookup; /* id of the first lookup in the sample stream */
	uint64_t seed;        /* STARTING_SEED (1070) in the reference */
	int32_t kernel_id;    /* XSB_K_* */
	int32_t xs_precision; /* XSB_XS_* */
	/* optional per-lookup results, host pointers of length `lookups`
	 * (macro: lookups x 5); NULL skips the output */
	double *out_energy;
	int32_t *out_mat;
	int64_t *out_index; /* unionized-grid index, or hash bin */
	double *out_macro;
	int32_t *out_argmax;
	/* mixed precision only: the per-lookup error bound B with
	 * |macro_c - FP64 macro_c| <= B for every channel, and 1 where the
	 * lookup was re-evaluated in FP64 */
	double *out_bound;
	int32_t *out_fallback;
} xsb_event_args;

/* History-based run (openmp-threading/Simulation.c:116-238) */
typedef struct {
	int64_t particles;
	int32_t lookups_per_particle;
	int32_t kernel_id;    /* 0: one particle per lane, its lookups in sequence;
	                         6: rounds -- lookup r of every particle, ordered by material
	                         and energy, one round after another (a replayed hipGraph);
	                         -1: rounds from 65536 particles on, else 0 */
	int32_t xs_precision; /* XSB_XS_* */
	int32_t reserved0;
	uint64_t seed;                 /* STARTING_SEED */
	uint64_t *out_particle_sum;    /* optional host array [particles] */
	uint64_t *out_final_seed;      /* optional host array [particles] */
} xsb_history_args;

/* A set of independent lookup batches, each with its own sample stream and
 * material CDF (no reference counterpart; the producer of many small event
 * batches).  Batch b holds sizes[b] lookups; lookup i of batch b samples
 * s = FF(seeds[b], 2i), E = LCG(s), roll = LCG(s), mat = first m with
 * roll < thresholds[12 b + m].
 *
 * mode 0 runs one pipeline captured once in a hipGraph and replayed while
 * the call shape (batch count, total lookups, plan, precision) is unchanged:
 * a key kernel, an LSD radix sort of all lookups by (sort window, material,
 * energy), then one persistent lookup kernel whose workgroups take tiles of
 * the sorted order from per-XCD work queues (stealing across XCDs when
 * their own queue runs dry).  Batches are grouped into sort windows that
 * run one after the other: by default (window 0) the smallest batches -- at
 * most 55% of them, holding at most 20% of the lookups -- form a first
 * window, so that most batches complete after a small part of the work;
 * window > 0 groups the batches in order into windows of about `window`
 * lookups; window -1 runs all batches in one window (the highest throughput,
 * every batch completing at the end).  Zero fields of
 * the plan (tile, stage, sort, mapping) are chosen by an autotuner keyed on the
 * nuclide count, grid type, precision and batch-size histogram, whose choice
 * is cached on the handle. */
typedef struct {
	int32_t tile;      /* lookups per work-queue item */
	int32_t stage;     /* index-grid words per lane staged in LDS per chunk (8 or 16) */
	int32_t sort;      /* 1: material, 2: material then energy */
	int32_t windows;   /* sort windows the batches were grouped into */
	int32_t tuned;     /* 1 if the autotuner ran during this call */
	int32_t graph;     /* 1 if the pipeline ran as a hipGraph */
	int32_t mapping;   /* 1: one lane per lookup; 2: wave-uniform path where a wave's
	                      lanes share the material (and on the hash grid the bin);
	                      3: as 2 without the two-interval shared step */
	int32_t reserved0;
	double ms_tuning;  /* host wall time spent tuning during this call */
} xsb_batch_plan;

typedef struct {
	int32_t n_batches;
	int32_t xs_precision;
	const int64_t *sizes;      /* host [n_batches] */
	const uint64_t *seeds;     /* host [n_batches] */
	const double *thresholds;  /* host [n_batches x 12], non-decreasing, thresholds[12b+11] >= 1 */
	int32_t mode;              /* 0 sorted pipeline (above), 1 one launch per batch in lookup
	                              order, 2 persistent queue in lookup order (each tile ordered
	                              by material in LDS) */
	int32_t tile;              /* lookups per work item (0 = autotuned in mode 0, 4096 in mode 2) */
	int32_t stage;             /* mode 0: 8 or 16 (0 = autotuned) */
	int32_t sort;              /* mode 0: 1 or 2 (0 = autotuned) */
	int64_t window;            /* mode 0: 0 small batches first (two windows), > 0 about this
	                              many lookups per window in batch order, -1 one window */
	int32_t eager;             /* mode 0: 1 = launch the kernels directly instead of the graph */
	int32_t mapping;           /* mode 0: 1 lane per lookup, 2 wave-uniform, 3 wave-uniform
	                              without the two-interval step (0 = autotuned) */
	uint64_t *out_vsum;        /* optional host [n_batches] */
	double *out_done_ms;       /* optional host [n_batches]: completion time of each batch,
	                              from the start of the call's first kernel */
	xsb_batch_plan *out_plan;  /* optional: the plan that ran */
} xsb_batched_args;

typedef struct {
	uint64_t vsum;        /* un-modded verification sum (main applies % 999983) */
	double ms_total;      /* GPU time of every kernel of the call (hipEvents) */
	double ms_lookup;     /* GPU time of the lookup kernel(s) */
	double ms_sample;     /* sampling / sort kernels */
	int64_t n_fallback;   /* lookups re-evaluated in FP64 (mixed precision) */
	int64_t material_counts[12];
	int32_t kernel_id;    /* schedule that ran */
	int32_t reserved0;
} xsb_result;

typedef struct {
	int32_t device;
	int32_t grid_type;
	int64_t n_isotopes;
	int64_t n_gridpoints;
	int64_t n_union;
	int64_t bytes_fp64;       /* FP64 AoS grid */
	int64_t bytes_fp32;       /* FP32 XS table (0 if not built) */
	int64_t bytes_bf16;       /* BF16 XS table (0 if not built) */
	int64_t bytes_index;      /* hash grid search tables: the index grid (int32 rows padded to
	                             16 bytes), 96-byte window records per bin and nuclide (same
	                             rows: u_low, u_high and ten energies) and the nuclides'
	                             energies by row (n_gridpoints + 16 rounded up to 8 doubles
	                             each); the unionized index grid is not kept whole on the
	                             device (0) */
	int64_t bytes_midx;       /* per-material copies of the unionized index grid */
	int64_t bytes_union;      /* unionized energies */
	int64_t bytes_ivl;        /* interval records of the unionized grid: FP64, plus FP32 with
	                             XSB_BUILD_FP32 */
	int32_t cu_count;
	int32_t xcd_count;
	char arch[32];
} xsb_info;

/* cuda/GridInit.cu:4 move_simulation_data_to_device: copy the tables to
 * `device` (hipMalloc + hipMemcpy, FP64 AoS + index grid + materials), and
 * build the compact XS tables named by `flags` on the device. */
int xsb_hip_upload(int device, const xsb_tables *t, uint32_t flags, void **handle);

/* Copies the handle's tables back to host arrays (any may be NULL):
 * nuclide_grid [n_isotopes x n_gridpoints x 6], unionized_energy_array
 * [n_union], index_grid [rows x n_isotopes] (unpadded, as the reference lays
 * it out).  For checking tables built with XSB_BUILD_GRIDS against
 * grid_init_do_not_profile (openmp-threading/GridInit.c:3). */
int xsb_hip_download(void *handle, double *nuclide_grid, double *unionized_energy_array,
                     int32_t *index_grid);

/* cuda/GridInit.cu:81 release_device_memory: free everything the handle owns. */
int xsb_hip_release(void *handle);

/* openmp-threading/Simulation.c:15 run_event_based_simulation and
 * Simulation.c:698 / cuda/Simulation.cu:388-1099 run_event_based_simulation_
 * optimization_k, selected by args->kernel_id. */
int xsb_hip_event(void *handle, const xsb_event_args *args, xsb_result *res);

/* No counterpart in the reference: xsb_hip_event for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, schedules and results,
 * but it allocates nothing: the sample and sort buffers come from one fixed
 * array in the library per GPU (34.6 MB, held until the process ends,
 * its address looked up by xsb_hip_upload), so at most
 * XSB_NOALLOC_MAX_LOOKUPS lookups per call (larger counts return
 * XSB_ERR_INVALID), and no per-lookup outputs (the out_* pointers must be
 * NULL).  Every call on one GPU shares that array, from any handle on that
 * GPU: calls that overlap on one GPU race. */
#define XSB_NOALLOC_MAX_LOOKUPS (1 << 21)
int xsb_hip_event_noalloc(void *handle, const xsb_event_args *args, xsb_result *res);

/* openmp-threading/Simulation.c:116 run_history_based_simulation */
int xsb_hip_history(void *handle, const xsb_history_args *args, xsb_result *res);

/* openmp-threading/Simulation.c:326 calculate_macro_xs over n explicit
 * samples held in device memory: d_energy[n], d_mat[n] -> d_macro[n x 5]
 * (nullable), d_argmax[n] (nullable), *d_vsum += sum(argmax + 1) (nullable).
 * A sample whose material is outside [0, 12) is skipped (argmax -1, adds
 * nothing); on the hash grid energies outside [0, 1) use the nearest bin.
 * Enqueued on `stream` (NULL = the handle's stream); does not synchronize. */
int xsb_hip_lookup(void *handle, int64_t n, const double *d_energy, const int32_t *d_mat,
                   int32_t xs_precision, double *d_macro, int32_t *d_argmax,
                   unsigned long long *d_vsum, void *stream);

/* openmp-threading/Simulation.c:326 calculate_macro_xs over n samples held in
 * host memory: energy[n], mat[n] (each in [0, 12)) -> macro[n x 5];
 * synchronous, on the handle's stream.  Backs the host export
 * calculate_macro_xs (include/xsbench.h); xsb_hip_lookup is the
 * asynchronous form over device memory. */
int xsb_hip_macro_host(void *handle, int64_t n, const double *energy, const int32_t *mat,
                       int32_t xs_precision, double *macro);

/* openmp-threading/Simulation.c:241 calculate_micro_xs over n samples held in
 * host memory, FP64: energy[n], nuc[n] -> xs[n x 5].  lo/hi carry what the
 * reference reads from index_data at the sample's grid index idx: on the
 * unionized grid lo[i] = index_data[idx * n_isotopes + nuc] (hi unused), on
 * the hash grid lo[i] = u_low and hi[i] = u_high (Simulation.c:273-281), on
 * the nuclide grid neither (both may be NULL).  Every lo/hi must lie in
 * [0, n_gridpoints - 1].  Synchronous.  Backs the host export
 * calculate_micro_xs. */
int xsb_hip_micro_host(void *handle, int64_t n, const double *energy, const int32_t *nuc,
                       const int32_t *lo, const int32_t *hi, double *xs);

/* Many independent batches (see xsb_batched_args); no reference counterpart. */
int xsb_hip_batched(void *handle, const xsb_batched_args *args, xsb_result *res);

/* Mean algorithmic bytes per lookup of an event run (the stored size of the
 * distinct words the reference algorithm reads), counted on the device for
 * the table format of `xs_precision`. */
int xsb_hip_balg(void *handle, const xsb_event_args *args, double *bytes_per_lookup);

/* Measurement helpers, no reference counterpart: the compulsory HBM bytes of
 * a workload -- the distinct 128-byte lines of device memory its lookups read
 * at least once (a line shared by many lookups counts once), in the layout
 * of the handle's tables and the format of xs_precision, plus the buffers of
 * the schedule that the lookup kernel reads and writes and the small tables.
 * No kernel can move fewer bytes from HBM for the same launch on this layout,
 * so bytes / lookup-kernel time / HBM peak is a roofline fraction of at most
 * 1 (up to lines still resident in the last-level cache from earlier work).
 * The FP64 re-evaluations of mixed-precision lookups are not counted. */
#define XSB_LINE_REGIONS 11
typedef struct {
	int64_t bytes;           /* bytes_tables + bytes_schedule + bytes_small */
	int64_t bytes_tables;    /* lines of the lookup tables (sum of region_bytes) */
	int64_t bytes_schedule;  /* sort pairs / sample arrays / particle states the kernel touches */
	int64_t bytes_small;     /* skip table, material tables and constants */
	int64_t launches;        /* lookup-kernel launches the bytes spread over (history: rounds) */
	/* per table: unionized energies, unionized search brackets, per-material
	 * index tables, FP64 interval records, FP32 interval records, hash index
	 * grid, FP64 points, FP32 records, BF16 records, hash window records,
	 * hash energy rows */
	int64_t region_bytes[XSB_LINE_REGIONS];
} xsb_lines;

/* The event run of `args` (per-lookup outputs ignored): lookups of the
 * sorted schedules' kernels for kernel_id >= 2, one lane per lookup below. */
int xsb_hip_lines_event(void *handle, const xsb_event_args *args, xsb_lines *out);
/* The history run of `args`, as its sorted rounds read it; one launch per round. */
int xsb_hip_lines_history(void *handle, const xsb_history_args *args, xsb_lines *out);
/* The batched run of `args` (sizes, seeds, thresholds, xs_precision). */
int xsb_hip_lines_batched(void *handle, const xsb_batched_args *args, xsb_lines *out);

/* Device copy bandwidth probe: copies `bytes` once per rep with a plain
 * vector copy kernel and returns the mean kernel time (for scaling HBM
 * counters). */
int xsb_hip_copy_probe(void *handle, int64_t bytes, int32_t reps, double *ms_per_copy);

int xsb_hip_info(void *handle, xsb_info *info);
/* No counterpart: the table sizes (xsb_info's bytes_* fields) that
 * xsb_hip_upload of t with `flags` would keep resident, computed on the host
 * from the sizes and the material tables alone (no device needed); for
 * checking that a problem fits one GPU before building it. */
int xsb_hip_table_bytes(const xsb_tables *t, uint32_t flags, xsb_info *info);
const char *xsb_hip_strerror(int code);
/* message of the last failure on the calling thread ("" if none) */
const char *xsb_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
