// LJatoms.cpp -- the shape of the reference's example driver
// (src/bin/LJatoms.cpp), written against the drop-in headers: the same
// classes and calls (NListed<EpsSigCutAtom, LennardJonesCutPair>, a
// NeighborList out to 1.4 cutoffs, random insertion with an energy check,
// CollectionVerlet, xyz frames), running on the MI355X.  The frames go
// through the device's asynchronous snapshots, so writing one does not stall
// the steps; the run length and file name come from the command line.
//
//   usage: LJatoms [natoms] [outer] [inner] [xyz file, "-" for none]
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "collection.hpp"

int main(int argc, char **argv) {
    const uint Natoms = argc > 1 ? (uint)atoi(argv[1]) : 400;
    const uint outer = argc > 2 ? (uint)atoi(argv[2]) : 500;
    const uint inner = argc > 3 ? (uint)atoi(argv[3]) : 1000;
    const std::string xyz = argc > 4 ? argv[4] : "LJatoms.xyz";
    const flt sigma = 1.0, epsilon = 1.0, sigcut = 2.5, phi = 0.3, dt = 0.0001;
    seed(12345);

    // a box holding Natoms spheres of diameter sigma at packing fraction phi
    const flt Vs = Natoms * std::pow(sigma, NDIM) * M_PI_2 / NDIM;
    const flt L = std::pow(Vs / phi, OVERNDIM);
    std::cout << "Using L = " << L << "\n";

    sptr<OriginBox> obox(new OriginBox(L));
    sptr<AtomVec> atomptr(new AtomVec(Natoms, 1.0));
    AtomVec &atoms = *atomptr;
    // cut at 2.5 sigma, neighbour list 40 % further out
    sptr<NListed<EpsSigCutAtom, LennardJonesCutPair> > LJ(
        new NListed<EpsSigCutAtom, LennardJonesCutPair>(obox, atomptr, 0.4 * (sigcut * sigma)));
    sptr<NeighborList> nl = LJ->neighbor_list();

    // insert the atoms one at a time at random places, retrying a place
    // whose energy is more than epsilon / 2 above the energy before
    for (uint i = 0; i < atoms.size(); i++) {
        flt E0 = LJ->energy(*obox);
        atoms[i].x = obox->rand_loc();
        atoms[i].v = rand_vec();
        atoms[i].f = Vec::Zero();
        atoms[i].a = Vec::Zero();
        LJ->add(EpsSigCutAtom(atoms.get_id(i), epsilon, sigma, sigcut));
        nl->update_list(true);
        while (LJ->energy(*obox) > E0 + epsilon / 2.0) {
            atoms[i].x = obox->rand_loc();
            nl->update_list(true);
        }
    }
    std::cout << "Starting. Neighborlist contains " << nl->numpairs() << " / "
              << (atoms.size() * (atoms.size() - 1)) / 2 << " pairs.\n";

    CollectionVerlet collec(std::static_pointer_cast<Box>(obox), atomptr, dt);
    collec.add_tracker(nl);
    collec.add_interaction(LJ);
    collec.reset_com_velocity();
    collec.scale_velocities_to_energy(Natoms / 4.0);

    const bool write = xyz != "-";
    if (write) atoms.write_xyz(atoms.snapshot(), xyz, false);
    flt E0 = collec.energy();
    std::printf("E0 %.12e K %.12e U %.12e\n", E0, collec.kinetic_energy(), LJ->energy(*obox));
    for (uint i = 0; i < outer; i++) {
        collec.timesteps(inner);
        // queued behind the steps; written while the next steps run
        const int64_t t = write ? atoms.snapshot() : -1;
        std::printf("step %u E %.12e K %.12e U %.12e T %.6f which %u\n", (i + 1) * inner,
                    collec.energy(), collec.kinetic_energy(), LJ->energy(*obox),
                    collec.temp(), nl->which());
        if (write) atoms.write_xyz(t, xyz, true);
    }
    std::cout << "Done. Neighborlist contains " << nl->numpairs() << " / "
              << (atoms.size() * (atoms.size() - 1)) / 2 << " pairs.\n";
    return 0;
}
