# Builds the reference MOCCa Fortran modules (read in place, never copied)
# plus this repository's own golden-vector driver (oracle/ref_driver.f90)
# into oracle/_ref/.  Used only to produce and check golden vectors; the
# product never links anything from here.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Requirements: amdflang (ROCm), LAPACK symbols from the image's MKL
# (libmkl_rt) -- the modules DensityMixing.f90 and SpwfStorage.f90 reference
# DSYSV/ZHEEV off the iteration path.
REFSRC  ?= /root/reference/src
OUT     := oracle/_ref
OBJDIR  := $(OUT)/obj
MODDIR  := $(OUT)/mod
FC      := amdflang
FFLAGS  ?= -O2
MKLDIR  ?= /opt/conda/lib

# Module order follows the reference Makefile (Makefile:9-17); Main.f90 and
# its generated version include are not built (the driver replaces Main).
MODS := CompilationInfo GenInfo Force OptimizedDerivatives Derivatives \
        CoulombDerivatives Mesh Spinor pfaff Spwf SpwfStorage Damping \
        Densities TrapPotential Moments SpecialMoments MultiGrid Coulomb \
        PairingInteraction LipkinNogami HartreeFock HFB GradientHFB BCS \
        Pairing Cranking MeanFields Energy PotentialMixing ImaginaryTime \
        DensityMixing Transform nil8 SpwfFactory Interfaces InOut
OBJS := $(patsubst %,$(OBJDIR)/%.o,$(MODS))

all: $(OUT)/mocca_ref_driver

$(OBJDIR) $(MODDIR):
	mkdir -p $@

# Serial chain: each module depends on the previous one (Fortran .mod order).
$(OBJDIR)/%.o: $(REFSRC)/%.f90 | $(OBJDIR) $(MODDIR)
	$(FC) $(FFLAGS) -c $< -o $@ -module-dir $(MODDIR) -I$(MODDIR)

define chain
$(OBJDIR)/$(2).o: $(OBJDIR)/$(1).o
endef
$(foreach i,$(shell seq 2 $(words $(MODS))),$(eval $(call chain,$(word $(shell expr $(i) - 1),$(MODS)),$(word $(i),$(MODS)))))

$(OBJDIR)/ref_driver.o: oracle/ref_driver.f90 $(OBJS)
	$(FC) $(FFLAGS) -c $< -o $@ -module-dir $(MODDIR) -I$(MODDIR)

$(OUT)/mocca_ref_driver: $(OBJDIR)/ref_driver.o $(OBJS)
	$(FC) $(FFLAGS) -o $@ $^ -L$(MKLDIR) -lmkl_rt -Wl,-rpath,$(MKLDIR)

clean:
	rm -rf $(OUT)

.PHONY: all clean
