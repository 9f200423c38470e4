# Builds oracle/_ref/emfa_ref: the reference's own EFIE/CFIE/contour-edge sources, compiled
# where they lie under /root/reference (never copied), plus oracle/ref_driver.cpp.
# Test infrastructure only (golden vectors); the product never links it.
#   make -C oracle -f ref.mk
# Flags: -fpermissive accepts the sources' MSVC-isms; the reference's own
# Mathdef.h is force-included as its MSVC build gets it; Armadillo is told not
# to call an external BLAS/LAPACK (none is needed by the code exercised), and
# <experimental/filesystem> is included for EMComponent/tools.cpp.
REF ?= /root/reference
OUT := _ref
CXX ?= g++
CXXFLAGS := -std=c++14 -O2 -fpermissive -w -DARMA_DONT_USE_LAPACK -DARMA_DONT_USE_BLAS -DARMA_DONT_USE_WRAPPER \
  -I$(REF)/include -I$(REF)/include/Eigen -I$(REF)/EMComponent -I$(REF)/EMSolver -I$(REF)/EMIterative \
  -include $(REF)/EMComponent/Mathdef.h

SRCS := EMComponent/Quadrature.cpp EMComponent/Mesh.cpp EMComponent/CommonEdge.cpp EMComponent/ConfigLoader.cpp \
  EMComponent/Triangle.cpp EMComponent/Tetrahedron.cpp EMComponent/Result.cpp EMComponent/ThreadPool.cpp \
  EMComponent/tools.cpp EMComponent/ContourEdge.cpp EMIterative/iml.cpp EMSolver/EM.cpp EMSolver/EFIE.cpp EMSolver/CFIE.cpp \
  EMSolver/Miscellaneous.cpp EMSolver/MFIE.cpp
OBJS := $(addprefix $(OUT)/,$(notdir $(SRCS:.cpp=.o)))

all: $(OUT)/emfa_ref

$(OUT):
	mkdir -p $(OUT)

$(OUT)/%.o: | $(OUT)
	$(CXX) $(CXXFLAGS) -include experimental/filesystem -c $(filter %/$*.cpp,$(addprefix $(REF)/,$(SRCS))) -o $@

$(OUT)/emfa_ref: ref_driver.cpp $(OBJS)
	$(CXX) $(CXXFLAGS) ref_driver.cpp $(OBJS) -o $@ -lstdc++fs

clean:
	rm -rf $(OUT)

.PHONY: all clean
