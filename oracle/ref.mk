# Builds the reference CBS3D++_SI solver from its own sources, where they lie,
# into oracle/_ref/ (kept out of git).  Only the serial/OpenMP native-CG variant
# (no MPI, no PETSc) is built: that variant is exactly the time-step hot path.
#
# Flags match the reference's CMake Release build (CMakeLists.txt:69-84):
# -O3 -DNDEBUG followed by the target's -O2, C++20, OpenMP with
# CBS3D_USE_OPENMP=1.  No -march, no FMA, no fast-math.
#
#   make -f oracle/ref.mk            # from the repository root
#
# Outputs:
#   oracle/_ref/cbs3dpp_si   the reference executable (case-file driven)
#   oracle/_ref/ref_dump     oracle/ref_dump.cpp linked against the reference
#                            objects; dumps per-stage arrays for golden fixtures
REF      ?= /root/reference
HERE     := $(dir $(lastword $(MAKEFILE_LIST)))
OUT      := $(HERE)_ref
CXX      ?= g++
CXXFLAGS := -std=c++20 -O3 -DNDEBUG -O2 -fopenmp -DCBS3D_USE_OPENMP=1 -I$(REF)/include

LIB_SRCS := io/MeshIO.cpp io/Post.cpp preprocess/Preprocess.cpp \
            boundary/Boundary.cpp timestep/TimeStep.cpp solver/Convergence.cpp \
            solver/Steps.cpp solver/Solver.cpp linalg/BandedMatrix.cpp \
            linalg/BandedGaussianSolver.cpp linalg/CSRMatrix.cpp \
            linalg/MatrixVectorCalc.cpp linalg/ConjugateGradient.cpp \
            assembly/PressureAssembly.cpp assembly/MomentumAssembly.cpp \
            assembly/EnergyAssembly.cpp utils/SolverProfiler.cpp \
            parallel/PartitionMetadata.cpp parallel/HaloExchange.cpp
LIB_OBJS := $(addprefix $(OUT)/obj/,$(LIB_SRCS:.cpp=.o))

all: $(OUT)/cbs3dpp_si $(OUT)/ref_dump

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/main.o: $(REF)/src/main.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/cbs3dpp_si: $(OUT)/obj/main.o $(LIB_OBJS)
	$(CXX) $(CXXFLAGS) $^ -o $@

$(OUT)/obj/ref_dump.o: $(HERE)ref_dump.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/ref_dump: $(OUT)/obj/ref_dump.o $(LIB_OBJS)
	$(CXX) $(CXXFLAGS) $^ -o $@

.PHONY: all
