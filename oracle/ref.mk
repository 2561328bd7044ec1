# Builds the reference's CPU paths from their sources under /root/reference
# (read in place, never copied) into oracle/_ref/.  Used only to pin the
# oracle (golden vectors) and as a timed CPU baseline in this container.
#   make -C oracle -f ref.mk
# Same flags as the reference's own build (tools/makefile/build.mk): gcc -O3
# -std=c11, OpenMP for the parallel path, no MPI (mpicc is absent here).

REF    ?= /root/reference
OUT    ?= _ref
CC     ?= gcc
CFLAGS := -O3 -std=c11 -D_GNU_SOURCE -I$(REF)/include
COMMON := $(REF)/src/common/aco_config.c $(REF)/src/common/aco_shared.c \
          $(REF)/src/common/cli_common.c $(REF)/src/common/instance_parser.c \
          $(REF)/src/common/matrix.c $(REF)/src/common/solution.c

all: $(OUT)/aco_seq $(OUT)/aco_omp $(OUT)/aco_omp_l3 $(OUT)/ref_knn $(OUT)/aco_vrp_cuda_linked

$(OUT):
	mkdir -p $@

$(OUT)/aco_seq: | $(OUT)
	$(CC) $(CFLAGS) -o $@ $(REF)/src/main.c $(REF)/src/seq/aco_sequential.c $(COMMON) -lm

$(OUT)/aco_omp: | $(OUT)
	$(CC) $(CFLAGS) -fopenmp -o $@ $(REF)/src/main.c $(REF)/src/openmp-mpi/aco_parallel.c $(COMMON) -lm

# the OpenMP CLI with its L3-size probe settable (REF_L3_SIZE), to time it
# at a chosen candidate width
$(OUT)/aco_omp_l3: ref_omp_l3_driver.c | $(OUT)
	$(CC) $(CFLAGS) -fopenmp -I$(REF)/src/openmp-mpi -o $@ $(REF)/src/main.c ref_omp_l3_driver.c $(COMMON) -lm

$(OUT)/ref_knn: ref_knn_driver.c | $(OUT)
	$(CC) $(CFLAGS) -fopenmp -I$(REF)/src/openmp-mpi -o $@ ref_knn_driver.c $(COMMON) -lm

# the reference CUDA CLI (src/cuda/main_vrp.cu: host code only, compiled as
# C++) with its own common sources, linked against this build's
# libaco_hip.so instead of src/cuda/aco_cuda.cu + aco_cuda_kernels.cu: the
# check that the drop-in entry point links unchanged (no GPU needed to link)
LIBDIR ?= ../vehicleroutingproblem_amd/build
$(OUT)/aco_vrp_cuda_linked: $(LIBDIR)/libaco_hip.so | $(OUT)
	g++ -O2 -x c++ -I$(REF)/include -c $(REF)/src/cuda/main_vrp.cu -o $(OUT)/main_vrp.o
	for f in aco_config aco_shared cli_common instance_parser matrix solution; do \
	  $(CC) $(CFLAGS) -c $(REF)/src/common/$$f.c -o $(OUT)/$$f.o || exit 1; done
	g++ -o $@ $(OUT)/main_vrp.o $(addprefix $(OUT)/,aco_config.o aco_shared.o cli_common.o \
	  instance_parser.o matrix.o solution.o) -L$(LIBDIR) -laco_hip \
	  -Wl,-rpath,$(abspath $(LIBDIR)) -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
