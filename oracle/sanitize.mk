# CPU-only build of the restatement (wc_oracle.c) and the scene generators (wc_scenes.cpp) under
# AddressSanitizer + UndefinedBehaviorSanitizer, with the command-line driver wc_sanitize_main.c.
# Host code only, compiled by gcc/g++; nothing here is GPU code.  Output: oracle/_san/wc_sanitize.
#     make -f oracle/sanitize.mk
CC = gcc
CXX = g++
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_san
SAN = -fsanitize=address,undefined,float-cast-overflow -fno-sanitize-recover=all -fno-omit-frame-pointer -g -O1
FP = -ffp-contract=off -fno-fast-math

all: $(OUT)/wc_sanitize

$(OUT)/wc_sanitize: $(HERE)wc_oracle.c $(HERE)wc_sanitize_main.c $(HERE)../watercube_amd/csrc/wc_scenes.cpp $(HERE)../include/watercube.h
	mkdir -p $(OUT)
	$(CC) $(SAN) $(FP) -std=c11 -fopenmp -c $(HERE)wc_oracle.c -o $(OUT)/wc_oracle.o
	$(CC) $(SAN) $(FP) -std=c11 -c $(HERE)wc_sanitize_main.c -o $(OUT)/wc_sanitize_main.o
	$(CXX) $(SAN) $(FP) -std=c++17 -c $(HERE)../watercube_amd/csrc/wc_scenes.cpp -o $(OUT)/wc_scenes.o
	$(CXX) $(SAN) -fopenmp -o $@ $(OUT)/wc_sanitize_main.o $(OUT)/wc_oracle.o $(OUT)/wc_scenes.o -lm

.PHONY: all
