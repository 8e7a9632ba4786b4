/*
 * main.c -- the XSBench program flow on the GPU path
 * (openmp-threading/Main.c, cuda/Main.cu), exported as xsbench_main() so a
 * host process can run it in-process; bin/XSBench (exe.c) only calls it.
 *
 * Same flow as the reference: read the CLI, build (or read) the tables, stage
 * them on the GPU outside the timer, time the run_* call, reduce the hash
 * % 999983 and return "hash invalid" as the exit code.
 */
#include <stdio.h>
#include <stdlib.h>

#include "xsbench.h"

int xsbench_main(int argc, char *argv[])
{
	const int version = 20;
	const int mype = 0, nprocs = 1;
	Inputs in = read_CLI(argc, argv);
#ifdef _OPENMP
	extern void omp_set_num_threads(int);
	omp_set_num_threads(in.nthreads);
#endif
	if (in.kernel_id < 0 || in.kernel_id > 6) {
		printf("Error: No kernel ID %d found!\n", in.kernel_id);
		exit(1);
	}
	print_inputs(in, nprocs, version);

	SimulationData SD;
	if (in.binary_mode == READ)
		SD = binary_read(in);
	else
		SD = grid_init_do_not_profile(in, mype);
	if (in.binary_mode == WRITE)
		binary_write(in, SD);

	/* staging is outside the timed region, as in cuda/Main.cu:42 */
	SD = move_simulation_data_to_device(in, mype, SD);

	printf("\n");
	border_print();
	center_print("SIMULATION", 79);
	border_print();

	double t0 = get_time();
	unsigned long long verification;
	if (in.simulation_method == EVENT_BASED)
		verification = run_event_based_simulation(in, SD, mype);
	else
		verification = run_history_based_simulation(in, SD, mype);
	printf("\nSimulation complete.\n");
	double t1 = get_time();

	release_device_memory(SD);
	verification = verification % 999983;
	int invalid = print_results(in, mype, t1 - t0, nprocs, verification);
	fflush(stdout);
	return invalid;
}
