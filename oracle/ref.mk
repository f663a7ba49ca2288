# Builds the reference Fortran hot-path modules, unmodified and straight from
# /root/reference/playground/src, together with this repository's own driver
# (oracle/ref_harness.f90), into oracle/_ref/<kernel>/ref_harness.
#
# Test infrastructure only: the binaries generate golden vectors and time the
# reference OpenMP path.  Nothing in the product links or loads them.
#
# Only the modules on the two-loop path are compiled (the reference Makefile's
# SRC list minus dumper/printer/IC/args/err*/main, which need HDF5 or are not on
# the path); args.f90 (the ./execute flag parser, which needs only const and
# state) is built for oracle/args_harness.f90.  The reference Makefile itself
# is not run: it hard-codes gfortran and a Homebrew HDF5 path.
#
#   make -f oracle/ref.mk KERNEL=m6 args   (the flag-parser driver)
#
#   make -f oracle/ref.mk KERNEL=m6      (also m4, gauss)
#
# Flags: FP64 default reals exactly as the reference Makefile (-fdefault-real-8
# -fdefault-double-8), OpenMP on, -O2 without fast-math, and FP contraction off
# so that `sd + d**2` in the neighbour test is rounded as written (the
# reference's generic x86-64 gfortran build emits no FMA either).

REFSRC  ?= /root/reference/playground/src
KERNEL  ?= m6
REF_FC  ?= /opt/rocm/llvm/bin/amdflang
OUT     := oracle/_ref/$(KERNEL)
FFLAGS  := -fdefault-real-8 -fdefault-double-8 -O2 -cpp -fopenmp -ffp-contract=off

MODS := utils/errprinter utils/arrayresize utils/timing utils/list utils/kdtree2 \
        const kernel/base_$(KERNEL) state kernel/kernel BC utils/neighboursearch \
        eos circuit1 circuit2 iterator timestep sts_integrator rad_exchange

OBJS := $(addprefix $(OUT)/,$(addsuffix .o,$(notdir $(MODS))))

all: $(OUT)/ref_harness

$(OUT):
	mkdir -p $(OUT)

# Modules must compile in dependency order (each writes its .mod into $(OUT)).
$(OUT)/stamp: oracle/ref.mk | $(OUT)
	set -e; for m in $(MODS); do \
	  $(REF_FC) $(FFLAGS) -module-dir $(OUT) -I $(OUT) -c $(REFSRC)/$$m.f90 -o $(OUT)/$$(basename $$m).o; \
	done
	touch $@

$(OUT)/ref_harness: $(OUT)/stamp oracle/ref_harness.f90
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -I $(OUT) oracle/ref_harness.f90 $(OBJS) -o $@

# ./execute flag parser (args.f90) with this repository's driver
$(OUT)/args_harness: $(OUT)/stamp oracle/args_harness.f90
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -I $(OUT) -c $(REFSRC)/args.f90 -o $(OUT)/args.o
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -I $(OUT) oracle/args_harness.f90 $(OUT)/args.o $(OBJS) -o $@

args: $(OUT)/args_harness

clean:
	rm -rf $(OUT)

.PHONY: all args clean

# Side-by-side program: reference iterate() and the HIP drop-in (needs libsph_hip.so)
HIPLIB := sph_playground_amd/libsph_hip.so
$(OUT)/ref_vs_hip: $(OUT)/stamp oracle/ref_vs_hip.f90 sph_playground_amd/fortran/sph_hip_capi.f90 \
                   sph_playground_amd/fortran/iterator_hip.f90 $(HIPLIB)
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -I $(OUT) -c sph_playground_amd/fortran/sph_hip_capi.f90 -o $(OUT)/sph_hip_capi.o
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -I $(OUT) -c sph_playground_amd/fortran/iterator_hip.f90 -o $(OUT)/iterator_hip.o
	$(REF_FC) $(FFLAGS) -module-dir $(OUT) -I $(OUT) oracle/ref_vs_hip.f90 $(OBJS) $(OUT)/sph_hip_capi.o \
	  $(OUT)/iterator_hip.o -L sph_playground_amd -lsph_hip -Wl,-rpath,'$$ORIGIN/../../../sph_playground_amd' -o $@

dropin: $(OUT)/ref_vs_hip
.PHONY: dropin
