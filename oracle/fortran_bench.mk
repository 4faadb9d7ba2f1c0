# The reference's Fortran benchmark driver, test/stellarator_bench.f90, and its
# tensor-patch smoke test, test/test_lap3d_close.f90, compiled unchanged against
# this library's Fortran drop-in modules
# (qotential_amd/fortran/qotential_modules.f90) and linked to
# libqotential_hip.so: the drop-in boundary exercised by the reference's own
# caller.  Output only under _ref/ (git-ignored; it travels with the built
# library).  Needs the reference checkout (REF) and amdflang.
REF ?= /root/reference
FLANG ?= /opt/rocm/bin/amdflang
OUT = _ref
MODSRC = ../qotential_amd/fortran/qotential_modules.f90
LIB = ../qotential_amd/libqotential_hip.so

all: $(OUT)/stellarator_bench $(OUT)/test_lap3d_close

$(OUT)/stellarator_bench: $(REF)/test/stellarator_bench.f90 $(MODSRC) $(LIB)
	@mkdir -p $(OUT)/mod
	$(FLANG) -O2 -module-dir $(OUT)/mod -c $(MODSRC) -o $(OUT)/qotential_modules.o
	$(FLANG) -O2 -I$(OUT)/mod $(REF)/test/stellarator_bench.f90 $(OUT)/qotential_modules.o \
	  -L../qotential_amd -lqotential_hip -Wl,-rpath,'$$ORIGIN/../../qotential_amd' -o $@

$(OUT)/test_lap3d_close: $(REF)/test/test_lap3d_close.f90 $(OUT)/stellarator_bench
	$(FLANG) -O2 -I$(OUT)/mod $(REF)/test/test_lap3d_close.f90 $(OUT)/qotential_modules.o \
	  -L../qotential_amd -lqotential_hip -Wl,-rpath,'$$ORIGIN/../../qotential_amd' -o $@

clean:
	rm -rf $(OUT)/mod $(OUT)/qotential_modules.o $(OUT)/stellarator_bench $(OUT)/test_lap3d_close
