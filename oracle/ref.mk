# Compiles the reference's own CPU path from its sources in place
# (/root/reference: cpu.cxx, arraysizes.cxx, adler32.c) with the reference
# Makefile's optimisation flags, plus oracle/ref_harness.cxx, into
# oracle/_ref/ref_harness. Nothing is copied from the reference; outputs go
# only to oracle/_ref/. Test infrastructure only: it produces golden vectors.
#
# The reference's main() in cpu.cxx is renamed with -Dmain=... so that the
# harness can call form_beams() itself.
#
# Usage: make -f oracle/ref.mk    (no-op when /root/reference is absent)

REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
CC ?= gcc
CXX ?= g++
# reference Makefile:11-12 (optimised build)
REF_CFLAGS = -std=c11 -fopenmp -march=native -O3 -DNDEBUG
REF_CXXFLAGS = -std=c++17 -fopenmp -Drestrict=__restrict__ -march=native -O3 -DNDEBUG

ifneq ($(wildcard $(REF)/cpu.cxx),)
all: $(OUT)/ref_harness
else
all:
	@echo "reference sources not present; oracle/_ref not built"
endif

$(OUT)/adler32.o: $(REF)/adler32.c
	mkdir -p $(OUT)
	$(CC) $(REF_CFLAGS) -I$(REF) -c $< -o $@

$(OUT)/arraysizes.o: $(REF)/arraysizes.cxx
	mkdir -p $(OUT)
	$(CXX) $(REF_CXXFLAGS) -I$(REF) -c $< -o $@

$(OUT)/cpu.o: $(REF)/cpu.cxx
	mkdir -p $(OUT)
	$(CXX) $(REF_CXXFLAGS) -Dmain=bf_reference_cpu_main -I$(REF) -c $< -o $@

$(OUT)/bf_oracle.o: $(HERE)bf_oracle.c $(HERE)bf_oracle.h
	mkdir -p $(OUT)
	$(CC) -std=c11 -O3 -fopenmp -ffp-contract=off -c $< -o $@

$(OUT)/ref_harness: $(HERE)ref_harness.cxx $(OUT)/adler32.o $(OUT)/arraysizes.o $(OUT)/cpu.o $(OUT)/bf_oracle.o
	$(CXX) $(REF_CXXFLAGS) -I$(REF) -I$(HERE) -o $@ $< $(OUT)/adler32.o $(OUT)/arraysizes.o $(OUT)/cpu.o \
	  $(OUT)/bf_oracle.o -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
