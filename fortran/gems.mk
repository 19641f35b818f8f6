# Builds GEMS itself, from its own sources where they lie, with the SMA-TB and
# Lennard-Jones force calls routed to the MI355X library when GEMS_BACKEND=hip:
# the drop-in at the reference's own plugin boundary (smatb_interact,
# src/TB.F90:457-586; lj_interact, src/Pairs.F90:207-279).
#
#   make -f fortran/gems.mk REF=/root/reference
#
# Output: fortran/_build/gems (links gems_amd/lib/libgems_hip.so by rpath).
# The reference sources are read in place; the hooks are applied to temporary
# copies outside the repository, compiled, and removed:
#   * `use gems_plugin_hip` beside `use gems_neighbour` (TB.F90:47, Pairs.F90:24),
#   * at the top of smatb_interact (before TB.F90:465) and of lj_interact
#     (before Pairs.F90:215):
#       if (gems_hip_on()) then; call tb_hip_interact(ig, ...); return; endif
# Module order and flags follow oracle/ref.mk (TB.F90 without -fopenmp: Flang
# 22 cannot build its polymorphic FIRSTPRIVATE).
REF     ?= /root/reference
FC      := /opt/rocm/lib/llvm/bin/flang
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_build
OBJ     := $(OUT)/gems_obj
LIB     := $(HERE)../gems_amd/lib
FFLAGS  := -O2 -cpp -DDIM3 -DPACKAGE_VERSION='"gems-hip"' \
           -I$(REF)/lib/fpt/include -I$(REF)/lib/fparser/src -module-dir $(OBJ) -I$(OBJ)

MODS_A := Constants.F90 Strings.f90 Errors.f90 Random.F90 Algebra.F90 Elements.F90 \
          Input_Parsing.F90 Atoms.F90 Groups.F90 Program_Types.F90 Variables.F90 \
          DDDA.F90 Tables.f90 Inq_Properties.F90 Set_Properties.f90 Select_Create.F90 \
          Output.F90 Neighbour.F90 ForceField.F90 Bias.f90
MODS_B := Tersoff.f90 Bond_Boost.f90 Fields.f90 Graphs.F90 CVS.F90 Integration.F90 \
          Interaction.f90 Quasi_Newton.f90 Checkpoint.F90 Programs.f90 NEB.f90 \
          Hyperdinamics.f90 Metadynamics.f90 MPI.F90 CLInterpreter.F90 Main.F90

all: $(OUT)/gems

$(OUT)/gems: $(HERE)gems_hip.f90 $(HERE)gems_plugin_hip.f90 $(LIB)/libgems_hip.so
	@test -d $(REF)/src || { echo "reference sources not found at $(REF)"; exit 1; }
	rm -rf $(OBJ) && mkdir -p $(OBJ)
	cd $(OBJ) && $(FC) -O2 -c $(REF)/lib/fparser/src/parameters.f90 -o parameters.o
	cd $(OBJ) && $(FC) -O2 -c $(REF)/lib/fparser/src/fparser.f90 -o fparser.o
	cd $(OBJ) && $(FC) -O2 -module-dir $(OBJ) -c $(HERE)gems_hip.f90 -o gems_hip.o
	cd $(OBJ) && for f in $(MODS_A); do \
	  $(FC) $(FFLAGS) -fopenmp -c $(REF)/src/$$f -o $${f%.*}.o || exit 1; \
	done
	cd $(OBJ) && $(FC) $(FFLAGS) -c $(HERE)gems_plugin_hip.f90 -o gems_plugin_hip.o
	tmp=$$(mktemp -d) && \
	sed -e '0,/^use gems_neighbour *$$/s//use gems_neighbour\nuse gems_plugin_hip, only: gems_hip_on, lj_hip_interact/' \
	    -e '0,/^ig%epot = 0._dp$$/s//if (gems_hip_on()) then\n  call lj_hip_interact(ig, ig%e, ig%s, ig%rcut)\n  return\nendif\nig%epot = 0._dp/' \
	    $(REF)/src/Pairs.F90 > $$tmp/Pairs.F90 && \
	sed -e '0,/^use gems_neighbour *$$/s//use gems_neighbour\nuse gems_plugin_hip, only: gems_hip_on, tb_hip_interact/' \
	    -e '0,/^call smatb_preinteraction(ig)$$/s//if (gems_hip_on()) then\n  call tb_hip_interact(ig, ig%nz, ig%z, ig%ca, ig%eps, ig%p, ig%q, ig%r0, ig%rci, ig%rce)\n  return\nendif\ncall smatb_preinteraction(ig)/' \
	    $(REF)/src/TB.F90 > $$tmp/TB.F90 && \
	grep -q 'lj_hip_interact(ig' $$tmp/Pairs.F90 && grep -q 'tb_hip_interact(ig' $$tmp/TB.F90 && \
	(cd $(OBJ) && $(FC) $(FFLAGS) -fopenmp -c $$tmp/Pairs.F90 -o Pairs.o && \
	 $(FC) $(FFLAGS) -c $$tmp/TB.F90 -o TB.o); rc=$$?; rm -rf $$tmp; exit $$rc
	cd $(OBJ) && for f in $(MODS_B); do \
	  $(FC) $(FFLAGS) -fopenmp -c $(REF)/src/$$f -o $${f%.*}.o || exit 1; \
	done
	cd $(OBJ) && $(FC) -fopenmp -O2 parameters.o fparser.o gems_hip.o \
	  $(patsubst %,%.o,$(basename $(MODS_A))) gems_plugin_hip.o Pairs.o TB.o \
	  $(patsubst %,%.o,$(basename $(MODS_B))) \
	  -L$(LIB) -lgems_hip -Wl,-rpath,'$$ORIGIN/../../gems_amd/lib' -o $(OUT)/gems
	rm -rf $(OBJ)

clean:
	rm -f $(OUT)/gems
	rm -rf $(OBJ)

.PHONY: all clean
