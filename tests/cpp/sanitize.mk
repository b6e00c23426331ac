# The library's host sources (proxies, loaders) and the oracle under
# AddressSanitizer and UndefinedBehaviorSanitizer: host code only, built with
# g++/gcc (no device code, no HIP runtime).  make -f sanitize.mk
ROOT    := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/../..)
OUT     := $(ROOT)/tests/cpp/_build
CSRC    := $(ROOT)/cudagsea_amd/csrc
SAN     := -fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer

all: $(OUT)/sanitize_host

$(OUT)/sanitize_host: sanitize_host.cpp $(CSRC)/gsea_proxy.cpp $(CSRC)/gsea_io.cpp \
		$(ROOT)/oracle/gsea_oracle.c $(ROOT)/oracle/oracle_impl.inc $(ROOT)/oracle/gsea_oracle.h
	@mkdir -p $(OUT)
	gcc -std=c11 -O1 -g -fopenmp -ffp-contract=off $(SAN) -c $(ROOT)/oracle/gsea_oracle.c \
		-o $(OUT)/san_oracle.o
	g++ -std=c++17 -O1 -g -Wall $(SAN) -I $(ROOT)/include -I $(CSRC) -I $(ROOT)/oracle \
		sanitize_host.cpp $(CSRC)/gsea_proxy.cpp $(CSRC)/gsea_io.cpp $(OUT)/san_oracle.o \
		-fopenmp -lm -o $@

.PHONY: all
