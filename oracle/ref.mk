# Compiles the reference GEMS binary from its own sources, where they lie,
# with LLVM Flang (the only Fortran compiler in this image).  Outputs go to
# oracle/_ref/ only.  This is the checker's strongest leg; it never ships as
# product code.
#
#   make -f oracle/ref.mk REF=/root/reference
#
# Recipe notes:
#  * Module order follows the dependency lines of src/Makefile.am.
#  * -DDIM3 as src/Makefile.am:76; OpenMP as the reference's default build.
#  * PACKAGE_VERSION is the autoconf DEFS string, passed on the command line.
#  * TB.F90 is compiled without -fopenmp: Flang 22 aborts on the polymorphic
#    FIRSTPRIVATE(ig) of TB.F90:384-387 ("not yet implemented: create
#    polymorphic host associated copy").  Its loops then run serially with the
#    same full-list / newton-off semantics the OpenMP build selects in
#    Neighbour.F90:160-171 (those flags come from Neighbour.F90, built with
#    OpenMP).
REF     ?= /root/reference
FC      := /opt/rocm/lib/llvm/bin/flang
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
OBJ     := $(OUT)/obj
FFLAGS  := -O2 -cpp -DDIM3 -DPACKAGE_VERSION='"flang-oracle"' \
           -I$(REF)/lib/fpt/include -I$(REF)/lib/fparser/src -module-dir $(OBJ) -I$(OBJ)

MODS := Constants.F90 Strings.f90 Errors.f90 Random.F90 Algebra.F90 Elements.F90 \
        Input_Parsing.F90 Atoms.F90 Groups.F90 Program_Types.F90 Variables.F90 \
        DDDA.F90 Tables.f90 Inq_Properties.F90 Set_Properties.f90 Select_Create.F90 \
        Output.F90 Neighbour.F90 ForceField.F90 Bias.f90 Pairs.F90 TB.F90 Tersoff.f90 \
        Bond_Boost.f90 Fields.f90 Graphs.F90 CVS.F90 Integration.F90 Interaction.f90 \
        Quasi_Newton.f90 Checkpoint.F90 Programs.f90 NEB.f90 Hyperdinamics.f90 \
        Metadynamics.f90 MPI.F90 CLInterpreter.F90 Main.F90

all: $(OUT)/gems

$(OUT)/gems:
	@test -d $(REF)/src || { echo "reference sources not found at $(REF)"; exit 1; }
	mkdir -p $(OBJ)
	cd $(OBJ) && $(FC) -O2 -c $(REF)/lib/fparser/src/parameters.f90 -o parameters.o
	cd $(OBJ) && $(FC) -O2 -c $(REF)/lib/fparser/src/fparser.f90 -o fparser.o
	cd $(OBJ) && for f in $(MODS); do \
	  omp=-fopenmp; [ $$f = TB.F90 ] && omp=; \
	  $(FC) $(FFLAGS) $$omp -c $(REF)/src/$$f -o $${f%.*}.o || exit 1; \
	done
	cd $(OBJ) && $(FC) -fopenmp -O2 parameters.o fparser.o $(patsubst %,%.o,$(basename $(MODS))) -o $(OUT)/gems
	rm -f $(OBJ)/*.o

clean:
	rm -rf $(OUT)

.PHONY: all clean
