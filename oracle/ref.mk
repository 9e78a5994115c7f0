# Build recipe for the upstream ECOGEN CPU code, used only as a test oracle.
#
# Compiles the reference sources where they lie (read-only tree) with the
# system g++ and links MPICH 3.3.2 as shipped in this image under /opt/conda.
# Objects and the executable go to oracle/_ref/ (git-ignored).  Nothing from
# the reference tree is copied.  The reference's own Makefile is not used.
#
#   make -f oracle/ref.mk -j8            # builds oracle/_ref/ECOGEN
#
# The image's conda libstdc++ is older than the system one, so the runtime
# search path points at a private directory holding links to just the three
# MPICH runtime libraries (libmpi, libgfortran, libquadmath).

REF      ?= /root/reference
MPIROOT  ?= /opt/conda
OUT      := $(dir $(lastword $(MAKEFILE_LIST)))_ref
OBJDIR   := $(OUT)/obj
LIBDIR   := $(OUT)/mpilib

SRCS     := $(shell find $(REF)/src -name '*.cpp' | sort)
OBJS     := $(patsubst $(REF)/src/%.cpp,$(OBJDIR)/%.o,$(SRCS))

CXX      := g++
# Same optimisation level as the upstream Makefile; no -march, so no FMA
# contraction (x86-64 baseline has no FMA instructions).
CXXFLAGS := -O3 -w -I$(MPIROOT)/include

all: $(OUT)/ECOGEN

$(OBJDIR)/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(LIBDIR)/.stamp:
	@mkdir -p $(LIBDIR)
	ln -sf $(MPIROOT)/lib/libmpi.so.12 $(LIBDIR)/libmpi.so.12
	ln -sf $(MPIROOT)/lib/libgfortran.so.4 $(LIBDIR)/libgfortran.so.4
	ln -sf $(MPIROOT)/lib/libquadmath.so.0 $(LIBDIR)/libquadmath.so.0
	touch $@

$(OUT)/ECOGEN: $(OBJS) $(LIBDIR)/.stamp
	$(CXX) -O3 $(OBJS) -o $@ $(LIBDIR)/libmpi.so.12 -Wl,-rpath,'$$ORIGIN/mpilib'

clean:
	rm -rf $(OUT)

.PHONY: all clean
