# Builds GEMS itself, from its own sources where they lie, with its force
# calls and its `dinamica` loop routed to the MI355X library when
# GEMS_BACKEND=hip: the drop-in at the reference's own plugin boundary
# (smatb_interact, src/TB.F90:457-586; lj/plj/slj/wca_interact,
# src/Pairs.F90:207-560; dinamic, src/Programs.f90:114-167).
#
#   make -f fortran/gems.mk REF=/root/reference
#
# Output: fortran/_build/gems (links gems_amd/lib/libgems_hip.so by rpath).
# The reference sources are read in place; fortran/patch_gems.py writes hooked
# copies of Pairs.F90, TB.F90, Integration.F90, Random.F90 and Programs.f90
# to a temporary directory outside the repository (the hooks are listed
# there), which is compiled in their place and removed.  The plugin modules
# come in where their dependencies allow: gems_plugin_hip after Bias.f90
# (before Pairs/TB, which call it), gems_plugin_hip_md after Checkpoint.F90
# (before Programs.f90).  Module order and flags follow oracle/ref.mk (TB.F90
# without -fopenmp: Flang 22 cannot build its polymorphic FIRSTPRIVATE).
REF     ?= /root/reference
FC      := /opt/rocm/lib/llvm/bin/flang
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_build
OBJ     := $(OUT)/gems_obj
LIB     := $(HERE)../gems_amd/lib
FFLAGS  := -O2 -cpp -DDIM3 -DPACKAGE_VERSION='"gems-hip"' \
           -I$(REF)/lib/fpt/include -I$(REF)/lib/fparser/src -module-dir $(OBJ) -I$(OBJ)

# build order; +name is a module of this directory
SRCS := Constants.F90 Strings.f90 Errors.f90 Random.F90 Algebra.F90 Elements.F90 \
        Input_Parsing.F90 Atoms.F90 Groups.F90 Program_Types.F90 Variables.F90 \
        DDDA.F90 Tables.f90 Inq_Properties.F90 Set_Properties.f90 Select_Create.F90 \
        Output.F90 Neighbour.F90 ForceField.F90 Bias.f90 +gems_plugin_hip.f90 \
        Pairs.F90 TB.F90 Tersoff.f90 Bond_Boost.f90 Fields.f90 Graphs.F90 CVS.F90 \
        Integration.F90 Interaction.f90 Quasi_Newton.f90 Checkpoint.F90 \
        +gems_plugin_hip_md.f90 Programs.f90 NEB.f90 Hyperdinamics.f90 \
        Metadynamics.f90 MPI.F90 CLInterpreter.F90 Main.F90
OBJS := $(patsubst %,%.o,$(basename $(subst +,,$(SRCS))))

all: $(OUT)/gems

$(OUT)/gems: $(HERE)gems_hip.f90 $(HERE)gems_plugin_hip.f90 $(HERE)gems_plugin_hip_md.f90 \
             $(HERE)patch_gems.py $(LIB)/libgems_hip.so
	@test -d $(REF)/src || { echo "reference sources not found at $(REF)"; exit 1; }
	rm -rf $(OBJ) && mkdir -p $(OBJ)
	cd $(OBJ) && $(FC) -O2 -c $(REF)/lib/fparser/src/parameters.f90 -o parameters.o
	cd $(OBJ) && $(FC) -O2 -c $(REF)/lib/fparser/src/fparser.f90 -o fparser.o
	cd $(OBJ) && $(FC) -O2 -module-dir $(OBJ) -c $(HERE)gems_hip.f90 -o gems_hip.o
	tmp=$$(mktemp -d) && python3 $(HERE)patch_gems.py $(REF)/src $$tmp && \
	(cd $(OBJ) && for f in $(SRCS); do \
	  case $$f in +*) src=$(HERE)$${f#+}; f=$${f#+};; *) src=$(REF)/src/$$f;; esac; \
	  [ -f $$tmp/$$f ] && src=$$tmp/$$f; \
	  omp=-fopenmp; [ $$f = TB.F90 ] && omp=; \
	  $(FC) $(FFLAGS) $$omp -c $$src -o $${f%.*}.o || exit 1; \
	done); rc=$$?; rm -rf $$tmp; exit $$rc
	cd $(OBJ) && $(FC) -fopenmp -O2 parameters.o fparser.o gems_hip.o $(OBJS) \
	  -L$(LIB) -lgems_hip -Wl,-rpath,'$$ORIGIN/../../gems_amd/lib' -o $(OUT)/gems
	rm -rf $(OBJ)

clean:
	rm -f $(OUT)/gems
	rm -rf $(OBJ)

.PHONY: all clean
