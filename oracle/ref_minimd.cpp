// Entry point of oracle/_ref/libref_minimd.so: links miniMD's own Atom, Neighbor,
// Force and setup code (compiled from /root/reference/miniMD by
// oracle/ref.mk) and exposes one C entry point, so the tests can compare the
// oracle's LJ force, pair generator and FCC lattice with the reference's
// serial path.  TEST INFRASTRUCTURE ONLY.
#include <cstring>

#include "atom.h"
#include "force.h"
#include "neighbor.h"
#include "thermo.h"

void create_box(Atom &, int, int, int, double);
int create_atoms(Atom &, int, int, int, double);

extern "C" {

// Builds an n x n x n FCC system at density rho (miniMD create_box /
// create_atoms, miniMD/setup.cpp:44-147), its binned half neighbour list
// with cutoff cutneigh on nbin^3 bins (Neighbor::setup/build,
// miniMD/neighbor.cpp:76-427) and the LJ forces with cutoff cutforce
// (Force::compute, miniMD/force.cpp:48-137).
// Outputs (capacities in *natoms / *npairs on input, counts on output):
//   pos[3*natoms], force[3*natoms] (the reference's force, no 48 factor),
//   pairs[2*npairs] flattened as miniMD/ljs.cu:104-117 does.
int ref_minimd_system(int n, double rho, int nbin, double cutneigh, double cutforce, double *pos, double *force,
                      int *natoms, int *pairs, long *npairs)
{
    Atom atom;
    Neighbor neighbor;
    Force force_;
    neighbor.nbinx = neighbor.nbiny = neighbor.nbinz = nbin;
    neighbor.cutneigh = cutneigh;
    neighbor.every = 20;
    force_.cutforce = cutforce;
    create_box(atom, n, n, n, rho);
    if (create_atoms(atom, n, n, n, rho)) return -1;
    if (atom.nlocal > *natoms) return -2;
    *natoms = atom.nlocal;
    force_.setup();
    neighbor.setup(atom);
    neighbor.build(atom);
    force_.compute(atom, neighbor, 0);
    for (int i = 0; i < atom.nlocal; i++)
        for (int c = 0; c < 3; c++) {
            pos[3 * i + c] = atom.x[i][c];
            force[3 * i + c] = atom.f[i][c];
        }
    long cnt = 0;
    for (int i = 0; i < atom.nlocal; i++)
        for (int k = 0; k < neighbor.numneigh[i]; k++) {
            if (cnt < *npairs) {
                pairs[2 * cnt] = i;
                pairs[2 * cnt + 1] = neighbor.firstneigh[i][k];
            }
            cnt++;
        }
    *npairs = cnt;
    return 0;
}
}
