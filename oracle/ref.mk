# Builds oracle/_ref/libqxref.so: the reference QXMD sources, compiled where
# they lie under /root/reference with ROCm flang, plus oracle/ref_harness.f90
# (C-callable wrappers).  Test infrastructure only: the library is a checker for
# the FFT-free routines of the path.
#
# Sources/fft3.f90 is left out: it needs FFTW3, which the image lacks, so the
# two symbols it defines (rfft3_, fftw3_alloc_) stay unresolved and every
# routine reached through them (hcnonloc, wvg2r, softchg, eigen_BKL) is pinned
# by the oracle's own tests instead.  Sources/mpi_comm.f90 is replaced by the
# reference's own serial stubs Sources/no_comm.f90, as its serial target does.
#
#   make -f oracle/ref.mk          (from the repository root)
REFSRC  ?= /root/reference/Sources
REFINC  ?= /root/reference/Include
ifeq ($(origin FC),default)
FC      := flang
endif
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
BLD     := $(OUT)/build
FFLAGS  := -cpp -DLIBFFTW3 -DPCHOLESKY=0 -DCCREAL4=0 -DIFC -DVECTOR -I$(REFINC) -O1 -fPIC \
           -module-dir $(BLD) -I$(BLD)

# dependency order of the reference's modules
UNITS := allinc constant modules remd_modules ftmain input pwlda pwreset kpoints bulk stress \
         openfile reclat ba eigen schmidt fermi vxc chgdns funcs commwv ppkb ppkbr ppvd nlkbpp \
         nlkbppr pcc force tddft_fssh sbalin savedd engrad datadump remd remd_in no_comm
OBJS  := $(addprefix $(BLD)/,$(addsuffix .o,$(UNITS)))

all: $(OUT)/libqxref.so

$(BLD):
	mkdir -p $(BLD)

# serialised: each unit may need the .mod files of the ones before it
$(OBJS): | $(BLD)
.NOTPARALLEL:
$(BLD)/%.o: $(REFSRC)/%.f90
	$(FC) $(FFLAGS) -c $< -o $@

$(BLD)/ref_harness.o: $(HERE)ref_harness.f90 $(OBJS)
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/libqxref.so: $(OBJS) $(BLD)/ref_harness.o
	$(FC) -shared -fPIC -o $@ $(OBJS) $(BLD)/ref_harness.o

clean:
	rm -rf $(OUT)
.PHONY: all clean
