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
	rm -rf $(BUILD) $(LIB) $(ORACLE)

.PHONY: all clean
