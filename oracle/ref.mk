# Builds the reference's two single-threaded CPU counters straight from their
# sources under /root/reference into oracle/_ref/ (git-ignored). They only
# check results: tests/golden/make_golden.py runs them to write fixtures, and
# bench.py may time them as the "reference" CPU baseline. Flags follow the
# reference notebooks (RunBenchmark.ipynb cells 5-6: g++ -O3 -std=c++17).
#
#   make -f oracle/ref.mk            (from the repo root)
#
# Nothing here is copied into the repository; the sources are compiled where
# they lie.

REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
CXXFLAGS ?= -O3 -std=c++17

REWRITTEN_SRC := $(REF)/Rewritten_CPP/main.cpp $(REF)/Rewritten_CPP/RdfIndexes.cpp \
                 $(REF)/Rewritten_CPP/RuleParser.cpp $(REF)/Rewritten_CPP/FinalRule.cpp \
                 $(REF)/Rewritten_CPP/SupportCounting.cpp
CPU01_SRC     := $(REF)/CPU_01_nonparallel/main.cpp $(REF)/CPU_01_nonparallel/RdfIndexes.cpp \
                 $(REF)/CPU_01_nonparallel/RuleParser.cpp $(REF)/CPU_01_nonparallel/FinalRule.cpp \
                 $(REF)/CPU_01_nonparallel/SupportCounting.cpp

FRONT_SRC     := $(REF)/GPU_03/RdfIndexes.cpp $(REF)/GPU_03/RuleParser.cpp $(REF)/GPU_03/FinalRule.cpp

all: $(OUT)/rewritten $(OUT)/cpu01 $(OUT)/frontend_probe

$(OUT)/rewritten: $(REWRITTEN_SRC)
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF)/Rewritten_CPP $(REWRITTEN_SRC) -o $@

$(OUT)/cpu01: $(CPU01_SRC)
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF)/CPU_01_nonparallel $(CPU01_SRC) -o $@

# the reference's Turtle reader and rule parser behind a probe main of this
# project (oracle/ref_frontend_probe.cpp)
$(OUT)/frontend_probe: oracle/ref_frontend_probe.cpp $(FRONT_SRC)
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -I$(REF)/GPU_03 oracle/ref_frontend_probe.cpp $(FRONT_SRC) -o $@

.PHONY: all
