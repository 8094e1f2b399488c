# Builds the reference's own CPU c_baseline (agurvich/WIND) from its sources,
# unmodified and as shipped (Katz96 network in cuda/ode_system/ode_gold.c,
# NUM_CONST = 2 from cuda/include/ode_gold.h), into oracle/_ref/.
# Test infrastructure only: the _ref libraries pin the oracle and can serve as
# the CPU baseline; they are never part of the product.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Flags follow cuda/c_baseline/Makefile:5-29 (-DADAPTIVE_TIMESTEP, -DSIE/-DRK2
# on common_gold.c), compiled by gcc directly with plain x86-64 code (no
# -march, so no FMA) and -ffp-contract=off.
REF ?= /root/reference/cuda
OUT ?= oracle/_ref
CC ?= gcc
CFLAGS = -O2 -fPIC -ffp-contract=off -DADAPTIVE_TIMESTEP
INC = -I$(REF)/c_baseline/linear_algebra -I$(REF)/c_baseline -I$(REF)/ode_system \
      -I$(REF)/c_baseline/include -I$(REF)/c_baseline/common -I$(REF)/include
CB = $(REF)/c_baseline

all: $(OUT)/sie_gold.so $(OUT)/rk2_gold.so

$(OUT)/sie_gold.so: $(CB)/common/common_gold.c $(CB)/sie/sie_gold.c $(CB)/linear_algebra/gje_gold.c \
                    $(CB)/linear_algebra/matrix_ops.c $(REF)/ode_system/ode_gold.c
	mkdir -p $(OUT)
	$(CC) -shared $(CFLAGS) -DSIE $(INC) -o $@ $^ -lm

$(OUT)/rk2_gold.so: $(CB)/common/common_gold.c $(CB)/rk2/rk2_gold.c $(REF)/ode_system/ode_gold.c
	mkdir -p $(OUT)
	$(CC) -shared $(CFLAGS) -DRK2 $(INC) -o $@ $^ -lm

clean:
	rm -f $(OUT)/sie_gold.so $(OUT)/rk2_gold.so

.PHONY: all clean
