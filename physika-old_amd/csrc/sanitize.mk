# CPU sanitizer build (make -f sanitize.mk): host code of the library and
# tests/cpp/pbf_host_check.cpp under AddressSanitizer + UndefinedBehaviorSanitizer
# (device code is built as usual); run by tests/test_sanitizers.py, no GPU
# needed. Outputs go to physika-old_amd/_build_san/.
include Makefile

SAN := ../_build_san
SANHOST := -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
  -Xarch_host -fno-sanitize-recover=all -Xarch_host -fno-omit-frame-pointer
CLANGXX ?= /opt/rocm/llvm/bin/clang++
sanitize: $(SAN)/pbf_host_check

$(SAN)/libpbf_san.so: $(SRC) $(HDR)
	@mkdir -p $(SAN)
	$(HIPCC) $(subst -O3,-O1 -g,$(HIPFLAGS)) $(SANHOST) -shared $(SRC) -o $@ $(LDLIBS)

$(SAN)/pbf_host_check: ../../tests/cpp/pbf_host_check.cpp ../../include/pbf.h $(SAN)/libpbf_san.so
	$(CLANGXX) -O1 -g -std=c++17 -Wall -I../../include -fsanitize=address,undefined -fno-sanitize-recover=all \
	  -fno-omit-frame-pointer ../../tests/cpp/pbf_host_check.cpp -L$(SAN) -lpbf_san -Wl,-rpath,'$$ORIGIN' -o $@

.DEFAULT_GOAL := sanitize
.PHONY: sanitize
