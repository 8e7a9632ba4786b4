/* exe.c -- bin/XSBench: the reference's command line (openmp-threading/Main.c) */
#include "xsbench.h"

int main(int argc, char *argv[])
{
	return xsbench_main(argc, argv);
}
