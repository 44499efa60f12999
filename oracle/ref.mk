# Builds oracle/_ref/libref_minimd.so (the oracle's second source) from the
# reference's own miniMD sources where they lie (never copied), plus the
# driver oracle/ref_minimd.cpp.  Needs /root/reference and MPICH's mpi.h
# (setup.cpp/atom.cpp include it; no MPI call is linked in the path used).
# Built without OpenMP: miniMD's serial path (its OpenMP pragmas use a
# default(none) form current GCC rejects).
REF ?= /root/reference/miniMD
MPI_INC ?= /opt/conda/include
SRC = $(REF)/atom.cpp $(REF)/force.cpp $(REF)/neighbor.cpp $(REF)/setup.cpp $(REF)/thermo.cpp
OUT = _ref/libref_minimd.so

$(OUT): ref_minimd.cpp $(SRC)
	mkdir -p _ref
	g++ -O2 -fPIC -shared -w -I$(REF) -I$(MPI_INC) -o $@ ref_minimd.cpp $(SRC) -L/opt/conda/lib -lmpi -Wl,-rpath,/opt/conda/lib

clean:
	rm -rf _ref
