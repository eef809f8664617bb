# Builds the gfx950 library (hexed_amd/libhexed_amd.so) and the CPU oracle used by the tests.
HIPCC ?= /opt/rocm/bin/hipcc
ARCH ?= gfx950
# packed FP32 instructions are disabled for every HIP translation unit
HIPFLAGS = -O3 -std=c++17 --offload-arch=$(ARCH) -fPIC -Wall -Xclang -target-feature -Xclang -packed-fp32-ops
CXX_ORACLE ?= g++
# x86-64-v3 (AVX2 + FMA) rather than -march=native: the library is built here and run on other hosts
ORACLE_FLAGS = -O3 -march=x86-64-v3 -fopenmp -fPIC -shared -std=c++20 -Wall

BUILD = hexed_amd/_build
SRC = hexed_amd/csrc
HDRS = $(wildcard $(SRC)/*.hpp) $(SRC)/basis_tables.inc include/hexed_amd.h
OBJS = $(BUILD)/hx_api.o $(BUILD)/hx_kernels_nd1.o $(BUILD)/hx_kernels_nd2.o $(BUILD)/hx_kernels_nd3.o
LIB = hexed_amd/libhexed_amd.so
ORACLE = oracle/liboracle_hexed.so

all: $(LIB) $(ORACLE)

$(BUILD):
	mkdir -p $(BUILD)

$(BUILD)/%.o: $(SRC)/%.hip $(HDRS) | $(BUILD)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(LIB): $(OBJS)
	$(HIPCC) $(HIPFLAGS) -shared -o $@.tmp $^ && mv $@.tmp $@

$(ORACLE): oracle/hexed_oracle.cpp oracle/hexed_oracle.h include/hexed_amd.h $(SRC)/basis_tables.inc
	$(CXX_ORACLE) $(ORACLE_FLAGS) -o $@.tmp $< && mv $@.tmp $@

clean:
	rm -f $(HOST_LIB) $(DROP_IN_TEST)
	rm -rf $(BUILD) $(LIB) $(ORACLE)

.PHONY: all clean

# C++ drop-in adapter over the C ABI (the reference's kernel interfaces) and its test driver
HOST_INC = -Ihexed_amd/host/include
HOST_LIB = hexed_amd/libhexed_host.so
DROP_IN_TEST = tests/cpp/test_drop_in

$(HOST_LIB): hexed_amd/host/src/amd.cpp $(wildcard hexed_amd/host/include/hexed/*.hpp) include/hexed_amd.h $(LIB)
	g++ -O2 -std=c++20 -fPIC -shared -Wall $(HOST_INC) -o $@.tmp $< -Lhexed_amd -lhexed_amd -Wl,-rpath,'$$ORIGIN' && mv $@.tmp $@

$(DROP_IN_TEST): tests/cpp/test_drop_in.cpp $(HOST_LIB) $(ORACLE)
	g++ -O2 -std=c++20 -Wall $(HOST_INC) -o $@ $< -Lhexed_amd -lhexed_host -lhexed_amd -Loracle -loracle_hexed \
	  -Wl,-rpath,'$$ORIGIN/../../hexed_amd' -Wl,-rpath,'$$ORIGIN/../../oracle'

all: $(HOST_LIB) $(DROP_IN_TEST)
