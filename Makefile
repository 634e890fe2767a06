# vite_amd build: HIP library for gfx950, CLI, CPU oracle and (when the reference tree is
# present) the compiled reference under oracle/_ref.
HIPCC    ?= /opt/rocm/bin/hipcc
ARCH     ?= gfx950
# -ffp-contract=off: the reference forms no FMA; a contracted gain can flip a near-tie.
# packed-fp32 ops are disabled on this platform.
HIPFLAGS := --offload-arch=$(ARCH) -O3 -std=c++17 -fPIC -ffp-contract=off \
            -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-result
CXXFLAGS := -O2 -std=c++17 -fPIC -ffp-contract=off -Wall
LIB      := vite_amd/libvite_amd.so
SRCS     := vite_amd/csrc/vt_core.hip vite_amd/csrc/vt_louvain.hip vite_amd/csrc/vt_rebuild.hip vite_amd/csrc/vt_color.hip vite_amd/csrc/vt_gen.hip vite_amd/csrc/vt_driver.hip vite_amd/csrc/vt_batch.hip vite_amd/csrc/vt_server.hip vite_amd/csrc/vt_compare.hip vite_amd/csrc/vt_dsort.hip
OBJS     := $(SRCS:.hip=.o)
HDRS     := include/vite_amd.h vite_amd/csrc/vt_internal.hpp

HOSTLIB  := vite_amd/libvite_host.so
CLI      := vite_amd/bin/vite_amd_cluster

all: $(LIB) $(HOSTLIB) $(CLI) oracle/liboracle.so

vite_amd/csrc/%.o: vite_amd/csrc/%.hip $(HDRS)
	$(HIPCC) $(HIPFLAGS) -c -o $@ $<

$(LIB): $(OBJS)
	$(HIPCC) --offload-arch=$(ARCH) -shared -o $@ $(OBJS)

# Host C++ drop-in layer (reference entry-point signatures over the C ABI) and the CLI.
$(HOSTLIB): vite_amd/host/vite_louvain.cpp vite_amd/host/vite_louvain.hpp vite_amd/host/vite_graph.hpp include/vite_amd.h $(LIB)
	g++ $(CXXFLAGS) -shared -o $@ vite_amd/host/vite_louvain.cpp -Lvite_amd -lvite_amd -Wl,-rpath,'$$ORIGIN'

$(CLI): vite_amd/host/main.cpp $(HOSTLIB)
	mkdir -p vite_amd/bin
	g++ $(CXXFLAGS) -o $@ vite_amd/host/main.cpp -Lvite_amd -lvite_host -lvite_amd -Wl,-rpath,'$$ORIGIN/..'

oracle/liboracle.so: oracle/vite_oracle.c oracle/vite_oracle.h
	gcc -O2 -std=gnu99 -Wall -ffp-contract=off -fPIC -shared -o $@ oracle/vite_oracle.c -lm

ref:
	@if [ -d /root/reference ]; then $(MAKE) -f oracle/Makefile.ref; else echo "reference tree absent: skipping oracle/_ref"; fi

clean:
	rm -f $(OBJS) $(LIB) $(HOSTLIB) $(CLI) oracle/liboracle.so

.PHONY: all ref clean
