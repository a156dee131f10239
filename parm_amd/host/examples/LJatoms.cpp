// LJatoms.cpp -- the reference's example driver (src/bin/LJatoms.cpp), written
// against the drop-in headers: the same calls, now running on the MI355X.
// Differences: the pair type is LJAttractRepulsePair (the one on the device
// path), atoms are inserted on a jittered lattice instead of one random trial
// at a time, and the run length comes from the command line.
//
//   usage: LJatoms [natoms] [outer] [inner]
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "collection.hpp"

int main(int argc, char **argv) {
    const uint Natoms = argc > 1 ? (uint)atoi(argv[1]) : 4000;
    const uint outer = argc > 2 ? (uint)atoi(argv[2]) : 10;
    const uint inner = argc > 3 ? (uint)atoi(argv[3]) : 100;
    const flt sigma = 1.0, epsilon = 1.0, sigcut = 2.5, rho = 0.8, dt = 0.002;
    seed(12345);
    const flt L = std::pow(Natoms / rho, OVERNDIM);
    std::cout << "Using L = " << L << "\n";

    sptr<OriginBox> obox(new OriginBox(L));
    sptr<AtomVec> atomptr(new AtomVec(Natoms, 1.0));
    AtomVec &atoms = *atomptr;
    sptr<NListed<IEpsSigCutAtom, LJAttractRepulsePair> > LJ(
        new NListed<IEpsSigCutAtom, LJAttractRepulsePair>(obox, atomptr, 0.3));
    sptr<NeighborList> nl = LJ->neighbor_list();

    uint m = (uint)std::ceil(std::pow((flt)Natoms, OVERNDIM) - 1e-9);
    vector<flt> eps(1, epsilon);
    for (uint i = 0; i < atoms.size(); i++) {
        uint q = i;
        Vec x;
        for (int k = 0; k < NDIM; k++) {
            x[k] = ((q % m) + 0.5) * (L / m) + 0.05 * gauss01();
            q /= m;
        }
        atoms[i].x = x;
        atoms[i].v = rand_vec();
        LJ->add(IEpsSigCutAtom(atoms.get_id(i), eps, 0, sigma, sigcut));
    }
    nl->update_list(true);
    std::cout << "Starting. Neighborlist contains " << nl->numpairs() << " pairs.\n";

    CollectionVerlet collec(std::static_pointer_cast<Box>(obox), atomptr, dt);
    collec.add_tracker(nl);
    collec.add_interaction(LJ);
    collec.reset_com_velocity();
    collec.scale_velocities_to_temp(1.0);

    flt E0 = collec.energy();
    std::printf("E0 %.12e K %.12e U %.12e\n", E0, collec.kinetic_energy(),
                LJ->energy(*obox));
    for (uint i = 0; i < outer; i++) {
        for (uint j = 0; j < inner; j++) collec.timestep();
        std::printf("step %u E %.12e K %.12e U %.12e T %.6f which %u\n", (i + 1) * inner,
                    collec.energy(), collec.kinetic_energy(), LJ->energy(*obox),
                    collec.temp(), nl->which());
    }
    std::cout << "Done. Neighborlist contains " << nl->numpairs() << " pairs.\n";
    return 0;
}
