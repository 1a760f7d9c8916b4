# Builds the reference's hot-path C++ (momokchung/polmdqc) from its own
# sources where they lie, into oracle/_ref/.  Test infrastructure only:
# the result is a checker (golden-vector generator and CPU cross-check),
# never part of the product path.  AlphaMol/GMP and QM sources are left
# out; the reference is built with clang because its OpenMP clauses need
# OpenMP 5.1 (default(private) with array-section reductions).
#
#   make -f oracle/ref.mk REF=/root/reference -j8
REF     ?= /root/reference
OUT     ?= oracle/_ref
CXX_REF ?= /opt/rocm/lib/llvm/bin/clang++
FLAGS   := -std=c++17 -O3 -finline-functions -ffast-math -fopenmp \
           -I$(REF)/include -I$(REF)/include/md -I$(REF)/include/qc -I$(REF)/include/oqc \
           -w
SKIP    := alfboxsize alfcx alfmol alforder alphamol alphamol1 alphamol2 alphavol \
           delcx initalf initalf2 initalfatm kalf multimol sosgmp spacefill vertex \
           analyzeExe testgradExe spacefillExe
MDSRC   := $(filter-out $(addprefix $(REF)/src/md/,$(addsuffix .cpp,$(SKIP))),$(wildcard $(REF)/src/md/*.cpp))
SRC     := $(REF)/src/mod.cpp $(REF)/src/promo.cpp $(REF)/src/rootdir.cpp $(REF)/src/qc/finalqm.cpp $(MDSRC)
OBJ     := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(SRC))
# these two use array-section reductions clang rejects under default(private)
NOOMP   := $(OUT)/obj/md/alterchg.o $(OUT)/obj/md/dcflux.o

all: $(OUT)/ref_dump $(OUT)/ref_analyze $(OUT)/ref_testgrad links

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX_REF) $(FLAGS) $(if $(filter $@,$(NOOMP)),-fno-openmp,) -c $< -o $@

$(OUT)/libpolmdqc_ref.a: $(OBJ)
	rm -f $@ && ar rcs $@ $^

$(OUT)/ref_dump: oracle/ref_dump.cpp $(OUT)/libpolmdqc_ref.a
	$(CXX_REF) $(FLAGS) $< $(OUT)/libpolmdqc_ref.a -o $@

$(OUT)/ref_analyze: $(REF)/src/md/analyzeExe.cpp $(OUT)/libpolmdqc_ref.a
	$(CXX_REF) $(FLAGS) $< $(OUT)/libpolmdqc_ref.a -o $@

$(OUT)/ref_testgrad: $(REF)/src/md/testgradExe.cpp $(OUT)/libpolmdqc_ref.a
	$(CXX_REF) $(FLAGS) $< $(OUT)/libpolmdqc_ref.a -o $@

# key files name parameters as ../../params/..., tests name inputs as
# ../../test/...: run the reference from $(OUT)/run/a/b
links:
	@mkdir -p $(OUT)/run/a/b
	@ln -sfn $(REF)/params $(OUT)/run/params
	@ln -sfn $(REF)/test $(OUT)/run/test

.PHONY: all links
