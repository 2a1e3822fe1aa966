# Build recipe for the CPU oracle (test infrastructure only).
#
#   make -f oracle/oracle.mk            liboracle.so (+ _ref when the reference
#                                       sources are present)
#
# liboracle.so  : this repository's C restatement (csr_oracle.c, tile_oracle.c),
#                 built with -ffp-contract=off so non-fused sums stay non-fused.
# _ref/         : the reference's own thmkl_dcsrmultcsr.c compiled from where it
#                 lies (REF_DIR), never copied.  Two builds:
#                   libthmkl_ref_fma.so   -O3 -march=x86-64-v3 (gcc contracts the
#                                         accumulate into an FMA, as the
#                                         reference's -march=native build does)
#                   libthmkl_ref_nofma.so -O3 -ffp-contract=off
#                 The reference's main() is renamed by a -D flag so the file can
#                 be a shared library.
ORACLE_DIR := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
REF_DIR ?= /root/reference
CC := gcc
CFLAGS := -O2 -fPIC -fopenmp -ffp-contract=off -Wall -Wno-unused-result

REF_SRC := $(REF_DIR)/thmkl_dcsrmultcsr.c
REF_OUT := $(ORACLE_DIR)/_ref

.PHONY: all oracle ref clean
all: oracle $(if $(wildcard $(REF_SRC)),ref,)

oracle: $(ORACLE_DIR)/liboracle.so

$(ORACLE_DIR)/liboracle.so: $(ORACLE_DIR)/csr_oracle.c $(ORACLE_DIR)/tile_oracle.c
	$(CC) $(CFLAGS) -shared -o $@ $^ -lm

ref: $(REF_OUT)/libthmkl_ref_fma.so $(REF_OUT)/libthmkl_ref_nofma.so

$(REF_OUT)/libthmkl_ref_fma.so: $(REF_SRC)
	mkdir -p $(REF_OUT)
	$(CC) -O3 -march=x86-64-v3 -fopenmp -fPIC -shared -w -Dmain=thmkl_ref_main \
	    -I$(REF_DIR) $< -o $@

$(REF_OUT)/libthmkl_ref_nofma.so: $(REF_SRC)
	mkdir -p $(REF_OUT)
	$(CC) -O3 -ffp-contract=off -fopenmp -fPIC -shared -w -Dmain=thmkl_ref_main \
	    -I$(REF_DIR) $< -o $@

clean:
	rm -f $(ORACLE_DIR)/liboracle.so $(REF_OUT)/*.so
