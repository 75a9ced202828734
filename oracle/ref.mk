# Builds the reference FastGrid CPU path (OpenMP, no CUDA) from its own
# sources where they lie under /root/reference, into oracle/_ref/.
# Test infrastructure only: the binary is the checker for the oracle and an
# optional CPU baseline, never part of the product.
#
#   make -f oracle/ref.mk -j8          (from the repository root)
#
# Sources: fastgrid_master/fastgrid/CMakeLists.txt:7-46 (the non-CUDA list).
# Also AutoDock 4.2.1's trilinear interpolation (AD/trilinterp.cc, AD =
# autogrid_original_4_2_1_ref/autodock) with the driver
# oracle/ref_trilinterp_main.cpp, compiled as AutoDock's default build does
# (Real = float, no FMA contraction on x86-64).
# AG_CUDA is left undefined so electrostatics/Cpu.cpp:114-127 provides the
# e-map backend.  The only portability fix is a forced <stdint.h>/<cassert>
# include (math/Int.h uses int32_t without including it); no source is edited
# or copied.

REF      ?= /root/reference
FG       := $(REF)/fastgrid_master/fastgrid
OUT      := oracle/_ref
OBJ      := $(OUT)/obj

CXX      ?= g++
CXXFLAGS := -std=gnu++98 -O2 -DLinux -DFASTGRID -DAG_OPENMP -fopenmp \
            -include stdint.h -include cassert -include unistd.h -w -I$(FG)

SRCS := ann/src/perf.cpp ann/src/kd_dump.cpp ann/src/bd_tree.cpp ann/src/ANN.cpp \
        ann/src/kd_search.cpp ann/src/brute.cpp ann/src/kd_tree.cpp ann/src/bd_search.cpp \
        ann/src/bd_fix_rad_search.cpp ann/src/kd_util.cpp ann/src/kd_pr_search.cpp \
        ann/src/kd_fix_rad_search.cpp ann/src/bd_pr_search.cpp ann/src/kd_split.cpp \
        math/Geometric/Sphere3.cpp math/Geometric/AxisAlignedBox3.cpp \
        math/Geometric/Vec4.cpp math/Geometric/Intersection3.cpp \
        times.cpp DesolvExpFunc.cpp InputDataLoader.cpp ParameterLibrary.cpp \
        PairwiseInteractionEnergies.cpp main.cpp BondVectors.cpp LogFile.cpp \
        CalculateGridmaps.cpp ProgramParameters.cpp Utils.cpp GridMap.cpp \
        electrostatics/Cpu.cpp

OBJS := $(addprefix $(OBJ)/,$(SRCS:.cpp=.o))

AD       := $(REF)/autogrid_original_4_2_1_ref/autodock

all: $(OUT)/fastgrid4_ref $(OUT)/testdiff $(OUT)/ref_trilinterp $(OUT)/fastgrid4_amd

$(OBJ)/%.o: $(FG)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/fastgrid4_ref: $(OBJS)
	$(CXX) -fopenmp -o $@ $(OBJS) -lpthread -lm

$(OUT)/testdiff: $(REF)/testdiff_master/main.cpp
	@mkdir -p $(OUT)
	$(CXX) -O2 -w -include cstdlib -include cstring -include climits -o $@ $<

$(OUT)/ref_trilinterp: $(AD)/trilinterp.cc oracle/ref_trilinterp_main.cpp
	@mkdir -p $(OBJ)
	$(CXX) -std=gnu++98 -O2 -w -I$(AD) -c $(AD)/trilinterp.cc -o $(OBJ)/ad_trilinterp.o
	$(CXX) -std=gnu++98 -O2 -w -I$(AD) -c oracle/ref_trilinterp_main.cpp -o $(OBJ)/ref_trilinterp_main.o
	$(CXX) -o $@ $(OBJ)/ad_trilinterp.o $(OBJ)/ref_trilinterp_main.o -lm

.PHONY: all

# The reference's own front end with libfgmaps in place of its two map
# backends (electrostatics/Cpu.cpp and CalculateGridmaps.cpp): the binding in
# fastgrid_amd/integration/amd_backend.cpp compiled against the reference's
# headers and linked with its other objects.  Checks that the drop-in
# boundary compiles and links as a maintainer would use it, and serves the
# GPU test of that binding.  Needs fastgrid_amd/libfgmaps.so (built first).
AMD_OBJS := $(addprefix $(OBJ)/,$(filter-out electrostatics/Cpu.o CalculateGridmaps.o,$(SRCS:.cpp=.o)))

$(OBJ)/amd_backend.o: fastgrid_amd/integration/amd_backend.cpp include/fgmaps.h
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -Wall -Iinclude -c $< -o $@

$(OUT)/fastgrid4_amd: $(AMD_OBJS) $(OBJ)/amd_backend.o fastgrid_amd/libfgmaps.so
	$(CXX) -fopenmp -o $@ $(AMD_OBJS) $(OBJ)/amd_backend.o -Lfastgrid_amd -lfgmaps \
	    -Wl,-rpath,'$$ORIGIN/../../fastgrid_amd' -lpthread -lm
