# Builds the reference's own CPU path (SEAL-4.0 port, src/*.cpp, src/utils/*.{cpp,c})
# straight from the sources where they lie under /root/reference, together with
# oracle/ref_driver.cpp (this repository's own extern "C" shim over the reference's
# public CPU API), into oracle/_ref/libtroyref.so.
#
# The result is a checker only: tests use it to validate the oracle's restatement,
# and bench.py may time it as the "reference" CPU baseline.  Nothing here copies
# reference sources into the repository; objects and the library go to oracle/_ref/.
#
#   make -f oracle/ref.mk -j8            (from the repository root)

REF      ?= /root/reference
OUT      := oracle/_ref
CXX      ?= g++
CC       ?= gcc
CXXFLAGS := -std=c++17 -O2 -fPIC -w -I$(REF)/src
CFLAGS   := -O2 -fPIC -w -I$(REF)/src/utils

REF_CPP  := $(filter-out %_cuda.cpp,$(wildcard $(REF)/src/*.cpp) $(wildcard $(REF)/src/utils/*.cpp))
REF_C    := $(wildcard $(REF)/src/utils/*.c)

OBJS := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(REF_CPP)) \
        $(patsubst $(REF)/src/%.c,$(OUT)/obj/%.o,$(REF_C)) \
        $(OUT)/obj/ref_driver.o

all: $(OUT)/libtroyref.so

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/%.o: $(REF)/src/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/obj/ref_driver.o: oracle/ref_driver.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libtroyref.so: $(OBJS)
	$(CXX) -shared -Wl,--no-undefined -o $@ $^

clean:
	rm -rf $(OUT)

.PHONY: all clean
