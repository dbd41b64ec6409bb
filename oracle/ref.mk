# Build the reference simulator from its own sources (read-only, under
# /root/reference/code) with the reference's own compiler flags
# (code/Makefile:3-5).  Objects and the binary go to oracle/_ref/ only.
# Usage (from the repo root):  make -f oracle/ref.mk
REF_SRC ?= /root/reference/code
OUT     ?= oracle/_ref
CXX     ?= g++
CXXFLAGS = -Wall -std=c++11 -pthread -O3 -w
UNITS = Info receptor StiExt neuron_affi neuronGnl2 EqsGnl2 parser_affi parser \
        NetSim NetSimGnl2 NetSimGnl2_info NetSimGnl2_BulPp main
OBJS  = $(addprefix $(OUT)/,$(addsuffix .o,$(UNITS)))

all: $(OUT)/flysim_ref

$(OUT)/flysim_ref: $(OBJS)
	$(CXX) $(OBJS) -o $@ $(CXXFLAGS)

$(OUT)/%.o: $(REF_SRC)/%.cpp | $(OUT)
	$(CXX) -c $< -o $@ -I$(REF_SRC) $(CXXFLAGS)

$(OUT):
	mkdir -p $(OUT)

.PHONY: all
