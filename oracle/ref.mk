# ref.mk -- build the REAL SKID reference from its sources where they lie.
#
# Test infrastructure only: the outputs go to oracle/_ref/ (git-ignored) and
# are used as a checker for the oracle's restatement and as the CPU baseline.
# Sources are read from $(REF) and never copied.  Flags follow the reference
# Makefile (-O3, no -march: no FMA contraction on x86-64).
#
#   make -f oracle/ref.mk            # from the repo root
#
REF    ?= /root/reference
OUT    ?= oracle/_ref
CC     ?= gcc
CFLAGS  = -O3 -I/usr/include/tirpc -I$(REF) -fPIC -w
LIBS    = -lm -ltirpc

OBJS = $(OUT)/kd.o $(OUT)/smooth1.o $(OUT)/grav.o $(OUT)/cosmo.o \
       $(OUT)/romberg.o $(OUT)/runge.o

all: $(OUT)/skid $(OUT)/totipnat $(OUT)/libskidref.so

$(OUT):
	mkdir -p $(OUT)

$(OUT)/%.o: $(REF)/%.c | $(OUT)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/main.o: $(REF)/main.c | $(OUT)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/skid: $(OUT)/main.o $(OBJS)
	$(CC) $(CFLAGS) -o $@ $^ $(LIBS)

$(OUT)/totipnat: $(REF)/totipnat.c | $(OUT)
	$(CC) $(CFLAGS) -o $@ $< $(LIBS)

$(OUT)/ref_harness.o: oracle/ref_harness.c | $(OUT)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/libskidref.so: $(OUT)/ref_harness.o $(OBJS)
	$(CC) -shared -o $@ $^ $(LIBS)

clean:
	rm -rf $(OUT)

.PHONY: all clean
