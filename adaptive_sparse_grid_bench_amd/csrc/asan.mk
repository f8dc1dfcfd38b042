# Host-code sanitizer build, for CPU runs only: the C++ host sources (grid tables,
# planners, C ABI, file format) with AddressSanitizer and UBSan, linked with the default
# build's device objects (build those first with the main Makefile).  Run the CPU tests
# against it with
#   make -f asan.mk
#   LD_PRELOAD=$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so) \
#   ASAN_OPTIONS=detect_leaks=0 SGRID_LIB=<csrc>/_build_asan/libsgrid_asan.so \
#   python -m pytest tests -m "not gpu"
ROCM      ?= /opt/rocm
CXX       ?= g++
HERE      := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OBJ       := $(HERE)_build
AOBJ      := $(HERE)_build_asan
SAN       := -fsanitize=address,undefined -fno-omit-frame-pointer
CXXFLAGS  := -O1 -g -std=c++17 -fPIC -Wall -D__HIP_PLATFORM_AMD__ -I$(ROCM)/include $(SAN)
HOST      := sg_grid sg_capi
DEVOBJS   := $(filter-out $(foreach h,$(HOST),$(OBJ)/$(h).o),$(wildcard $(OBJ)/*.o))

$(AOBJ)/libsgrid_asan.so: $(foreach h,$(HOST),$(AOBJ)/$(h).o) $(DEVOBJS)
	$(CXX) -shared $(SAN) -o $@ $^ -L$(ROCM)/lib -lamdhip64 -lpthread

$(AOBJ)/%.o: $(HERE)%.cpp $(wildcard $(HERE)*.hpp) $(HERE)../../include/sgrid.h
	@mkdir -p $(AOBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@
