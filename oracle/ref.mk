# ref.mk -- builds the reference's own kernels from its sources under
# /root/reference (qphix-codegen), for use as a test oracle cross-check and
# as the CPU baseline.  Nothing from the reference is copied into the repo:
# everything is generated and compiled into oracle/_ref/ (git-ignored).
#
#   1. the code generator is compiled directly with g++ from
#      codegen/{codegen,data_types,dslash,dslash_common,inst_*}.cc with the
#      per-ISA preset flags of codegen/CMakeLists.txt:151-239 (the
#      generator's CMake/Make build is not run);
#   2. codegen/output_dir.h.in is configured with sed the way CMake's
#      configure_file would (its only variable is @CMAKE_BINARY_DIR@);
#   3. the generator is run; it writes the kernel bodies;
#   4. oracle/ref_driver.cpp (our QPhiX-style block driver) #includes the
#      bodies it needs (the operator bodies and the face pack / finish
#      bodies) inside functions with the parameter names of
#      codegen/jinja/*_decl.h.j2, and is compiled into a shared library.
#
# Configurations (ISA, precision, VECLEN = SOALEN, codegen/jinja/isa.js):
#   scalar_f64  scalar  FP64   1
#   avx512_f64  avx512  FP64   8
#   avx512_f32  avx512  FP32  16
#   avx2_f64    avx2    FP64   4
#   avx2_f32    avx2    FP32   8
#   avx512_f16  avx512  FP16 storage, FP32 arithmetic (ENABLE_LOW_PRECISION,
#               codegen/CMakeLists.txt:78-83)  16
# and two xy-packed ones (VECLEN = 2 SOALEN: one gauge vector spans two rows
# of y; soalens of codegen/jinja/isa.js):
#   avx512_f64_s4   avx512  FP64   VECLEN 8, SOALEN 4
#   avx512_f32_s8   avx512  FP32   VECLEN 16, SOALEN 8
REF ?= /root/reference/codegen
OUT := $(abspath _ref)
CXX ?= g++

GEN_SRCS := codegen.cc data_types.cc dslash.cc dslash_common.cc inst_dp_vec8.cc \
            inst_sp_vec16.cc inst_dp_vec4.cc inst_sp_vec8.cc inst_sp_vec4.cc \
            inst_dp_vec2.cc inst_scalar.cc
GEN_PATHS := $(addprefix $(REF)/,$(GEN_SRCS))

# yes/no flags common to every preset (CMakeLists.txt:151-239)
YES := -DSERIAL_SPIN -DUSE_LDUNPK -DUSE_PKST -DENABLE_STREAMING_STORES \
       -DUSE_PACKED_GAUGES -DUSE_PACKED_CLOVER

DEFS_scalar_f64 := $(YES) -DNO_HW_MASKING -DNO_MASKS -DPRECISION=2 -DVECLEN=1 -DSOALEN=1
DEFS_avx512_f64 := $(YES) -DAVX512 -DPRECISION=2 -DVECLEN=8 -DSOALEN=8
DEFS_avx512_f32 := $(YES) -DAVX512 -DPRECISION=1 -DVECLEN=16 -DSOALEN=16
DEFS_avx2_f64 := $(YES) -DNO_HW_MASKING -DAVX2 -DPRECISION=2 -DVECLEN=4 -DSOALEN=4
DEFS_avx2_f32 := $(YES) -DNO_HW_MASKING -DAVX2 -DPRECISION=1 -DVECLEN=8 -DSOALEN=8
DEFS_avx512_f16 := $(YES) -DAVX512 -DPRECISION=1 -DVECLEN=16 -DSOALEN=16 -DENABLE_LOW_PRECISION \
                   -DUSE_LP_SPINOR -DUSE_LP_GAUGE -DUSE_LP_CLOVER
DEFS_avx512_f64_s4 := $(YES) -DAVX512 -DPRECISION=2 -DVECLEN=8 -DSOALEN=4
DEFS_avx512_f32_s8 := $(YES) -DAVX512 -DPRECISION=1 -DVECLEN=16 -DSOALEN=8

ARCH_scalar_f64 := scalar
ARCH_avx512_f64 := avx512
ARCH_avx512_f32 := avx512
ARCH_avx2_f64 := avx2
ARCH_avx2_f32 := avx2
ARCH_avx512_f16 := avx512
ARCH_avx512_f64_s4 := avx512
ARCH_avx512_f32_s8 := avx512

MARCH_scalar_f64 := -march=x86-64
MARCH_avx512_f64 := -mavx512f -mfma
MARCH_avx512_f32 := -mavx512f -mfma
MARCH_avx2_f64 := -mavx2 -mfma -mf16c
MARCH_avx2_f32 := -mavx2 -mfma -mf16c
MARCH_avx512_f16 := -mavx512f -mfma -mf16c
MARCH_avx512_f64_s4 := -mavx512f -mfma
MARCH_avx512_f32_s8 := -mavx512f -mfma

FT_scalar_f64 := double
FT_avx512_f64 := double
FT_avx512_f32 := float
FT_avx2_f64 := double
FT_avx2_f32 := float
FT_avx512_f16 := half
FT_avx512_f64_s4 := double
FT_avx512_f32_s8 := float
VL_scalar_f64 := 1
VL_avx512_f64 := 8
VL_avx512_f32 := 16
VL_avx2_f64 := 4
VL_avx2_f32 := 8
VL_avx512_f16 := 16
VL_avx512_f64_s4 := 8
VL_avx512_f32_s8 := 16
# SOALEN = VECLEN unless set
SOA_avx512_f64_s4 := 4
SOA_avx512_f32_s8 := 8
soa = $(if $(SOA_$(1)),$(SOA_$(1)),$(VL_$(1)))

CONFIGS := scalar_f64 avx512_f64 avx512_f32 avx2_f64 avx2_f32 avx512_f16 avx512_f64_s4 avx512_f32_s8

all: $(foreach c,$(CONFIGS),$(OUT)/libref_$(c).so)

# configured header, generator, generated bodies, driver library
$(OUT)/%/output_dir.h: $(REF)/output_dir.h.in
	mkdir -p $(OUT)/$*
	sed 's|@CMAKE_BINARY_DIR@|$(OUT)/$*|' $< > $@

$(OUT)/%/codegen: $(OUT)/%/output_dir.h $(GEN_PATHS)
	$(CXX) -std=c++11 -O1 -w $(DEFS_$*) -I$(OUT)/$* -I$(REF) $(GEN_PATHS) -o $@

$(OUT)/%/generated.stamp: $(OUT)/%/codegen
	mkdir -p $(OUT)/$*/generated/$(ARCH_$*)/generated
	cd $(OUT)/$* && ./codegen > codegen.log
	touch $@

$(OUT)/%/bodies.inc: $(OUT)/%/generated.stamp ref_wrappers.py
	python3 ref_wrappers.py $@ $(FT_$*) $(VL_$*) $(call soa,$*)

$(OUT)/%/faces.inc: $(OUT)/%/generated.stamp ref_wrappers.py
	python3 ref_wrappers.py --faces $@ $(FT_$*) $(VL_$*) $(call soa,$*)

$(OUT)/libref_%.so: $(OUT)/%/bodies.inc $(OUT)/%/faces.inc ref_driver.cpp
	$(CXX) -std=c++14 -O3 -fopenmp -fPIC -shared -w $(MARCH_$*) $(DEFS_$*) \
	  -DREF_CONFIG_NAME='"$*"' -I$(OUT)/$* -I$(OUT)/$*/generated/$(ARCH_$*) -I$(REF)/include \
	  ref_driver.cpp -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
.SECONDARY:
