# Recipe that compiles the reference's openmp-threading sources, where they
# lie, into oracle/_ref/ (never into the tree's tracked files).  Flags follow
# openmp-threading/Makefile:20-63 (gcc, -std=gnu99 -flto -fopenmp -DOPENMP -O3).
#   _ref/XSBench            the reference executable
#   _ref/libxsbench_ref.so  every reference source except Main.c, for the
#                           golden-vector dumper and the CPU baseline
REF ?= /root/reference/openmp-threading
OUT := _ref
CC := gcc
REF_CFLAGS := -std=gnu99 -Wall -flto -fopenmp -DOPENMP -O3
LIB_SRC := $(REF)/io.c $(REF)/Simulation.c $(REF)/GridInit.c $(REF)/XSutils.c $(REF)/Materials.c

all: $(OUT)/XSBench $(OUT)/libxsbench_ref.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/XSBench: $(REF)/Main.c $(LIB_SRC) $(REF)/XSbench_header.h | $(OUT)
	$(CC) $(REF_CFLAGS) $(REF)/Main.c $(LIB_SRC) -o $@ -lm

$(OUT)/libxsbench_ref.so: $(LIB_SRC) $(REF)/XSbench_header.h | $(OUT)
	$(CC) $(REF_CFLAGS) -fPIC -shared $(LIB_SRC) -o $@ -lm

.PHONY: all
