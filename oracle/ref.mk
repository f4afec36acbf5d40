# Builds pieces of the reference from its own sources under /root/reference
# into oracle/_ref/ (the reference's build system is not used).  Only the
# sources that compile as shipped are built: the pwl interpolation C files and
# the InterpolationScheme class.  The force methods do not compile as shipped
# (direct_force_method.h opens a comment at :58 that swallows the rest of the
# header) and are not built.  Test infrastructure only.
REF ?= /root/reference
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT := $(HERE)_ref
CFLAGS := -O2 -fPIC -w

all: $(OUT)/libref_pwl.so $(OUT)/libref_efield.so

$(OUT)/r8lib.o: $(REF)/r8lib.c
	mkdir -p $(OUT)
	gcc $(CFLAGS) -c $< -o $@

$(OUT)/pwl.o: $(REF)/pwl_interp_2d_scattered.c
	mkdir -p $(OUT)
	gcc $(CFLAGS) -I$(REF) -c $< -o $@

$(OUT)/libref_pwl.so: $(OUT)/pwl.o $(OUT)/r8lib.o
	gcc -shared -o $@ $^ -lm

$(OUT)/libref_efield.so: $(HERE)ref_efield_driver.cpp $(REF)/interpolation_scheme.cpp $(REF)/auxiliary_functions.cpp $(REF)/timer.cpp $(OUT)/pwl.o $(OUT)/r8lib.o
	g++ $(CFLAGS) -std=c++14 -I$(REF) -shared -o $@ $(HERE)ref_efield_driver.cpp $(REF)/interpolation_scheme.cpp $(REF)/auxiliary_functions.cpp $(REF)/timer.cpp $(OUT)/pwl.o $(OUT)/r8lib.o -lm -lpthread

clean:
	rm -rf $(OUT)
.PHONY: all clean
