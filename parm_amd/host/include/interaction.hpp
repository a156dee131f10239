// interaction.hpp -- drop-in for ParM's Interaction, IEpsSigCutAtom /
// LJAttractRepulsePair, EpsSigExpAtom / RepulsionPair and NListed<A,P>
// (src/interaction.hpp:59-104, 989-1018, 1251-1299, 1454-1466, 1528-1590,
// 1876-1945, 2101-2291) on the MI355X device path.
//
// NListed<A,P> keeps the reference's interface.  Its per-atom parameters are
// folded into a type table (one type per distinct parameter tuple) that
// configures the device; energy, forces, pressure, stress, contacts and
// overlaps then run on the device over the device list.  Only the two pair
// types of the device path are provided.
#ifndef INTERACTION_H
#define INTERACTION_H

#include <map>
#include <tuple>
#include <typeinfo>

#include "trackers.hpp"

PARM_NS_BEGIN

class Interaction {
   public:
    virtual flt energy(Box &box) = 0;
    virtual void set_forces(Box &box) = 0;
    virtual flt set_forces_get_pressure(Box &box) {
        throw std::runtime_error(std::string("set_forces_get_pressure not defined for class ") +
                                 typeid(*this).name());
    }
    virtual flt pressure(Box &box) = 0;
    virtual Matrix stress(Box &box) {
        (void)box;
        throw std::runtime_error(std::string("stress not defined for class ") +
                                 typeid(*this).name());
    }
    //! Device-path hooks: the AtomVec and NeighborList this interaction runs
    //! on (nullptr for interactions that are not on the device path).
    virtual AtomVec *device_atoms() { return nullptr; }
    virtual NeighborList *device_list() { return nullptr; }
    virtual void device_prepare() {}
    virtual ~Interaction() {}
};

// ------------------------------------------------------ per-atom types
struct IEpsSigCutAtom : public AtomID {
    vector<flt> epsilons;
    uint indx;
    flt sigma;
    flt sigcut;
    IEpsSigCutAtom() {}
    IEpsSigCutAtom(AtomID a, vector<flt> epsilons, uint indx, flt sigma, flt cut)
        : AtomID(a), epsilons(epsilons), indx(indx), sigma(sigma), sigcut(cut) {}
    IEpsSigCutAtom(AtomID a, IEpsSigCutAtom other)
        : AtomID(a), epsilons(other.epsilons), indx(other.indx), sigma(other.sigma),
          sigcut(other.sigcut) {}
    flt get_epsilon(IEpsSigCutAtom &other) {
        if (other.indx >= epsilons.size())
            throw std::invalid_argument("IEpsSigCutAtom: epsilon index out of range");
        return epsilons[other.indx];
    }
    flt max_size() { return sigma * sigcut; }
};

struct EpsSigExpAtom : public AtomID {
    flt eps, sigma, exponent;
    EpsSigExpAtom() {}
    EpsSigExpAtom(AtomID a, flt eps, flt sigma, flt exponent)
        : AtomID(a), eps(eps), sigma(sigma), exponent(exponent) {}
    EpsSigExpAtom(AtomID a, EpsSigExpAtom other)
        : AtomID(a), eps(other.eps), sigma(other.sigma), exponent(other.exponent) {}
    flt max_size() { return sigma; }
};

// ------------------------------------------------------------- pairs
//! Single-pair evaluation (the reference's pair structs); the bulk sums of
//! NListed run on the device, these serve NListed::energy(box, pair) and
//! direct callers.
struct LJAttractRepulsePair {
    flt eps, sig;
    flt cut_distance, cut_energy;
    AtomID atom1, atom2;
    LJAttractRepulsePair(IEpsSigCutAtom a1, IEpsSigCutAtom a2)
        : eps(a1.get_epsilon(a2)),
          sig((a1.sigma + a2.sigma) / 2.0),
          cut_distance(std::max(a1.sigcut, a2.sigcut)),
          atom1(a1),
          atom2(a2) {
        if (eps <= 0) {
            cut_distance = 1;
            cut_energy = 0;
            eps = std::fabs(eps);
            return;
        }
        flt mid = (1 - std::pow(cut_distance, -6));
        cut_energy = eps * (mid * mid);
    }
    flt energy(Box &box) {
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt rsq = rij.squaredNorm() / (sig * sig);
        if (rsq > cut_distance * cut_distance) return 0;
        flt mid = (1 - std::pow(rsq, -3));
        return eps * (mid * mid) - cut_energy;
    }
    Vec forces(Box &box) {
        if (eps == 0) return Vec::Zero();
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt dsq = rij.squaredNorm();
        flt rsq = dsq / (sig * sig);
        if (rsq > (cut_distance * cut_distance)) return Vec::Zero();
        flt rsix = std::pow(rsq, -3);
        flt fmagTimesR = 12 * eps * rsix * (rsix - 1);
        return rij * (fmagTimesR / dsq);
    }
};

struct RepulsionPair {
    flt eps, sig, exponent;
    AtomID atom1, atom2;
    RepulsionPair(EpsSigExpAtom a1, EpsSigExpAtom a2)
        : eps(std::sqrt(a1.eps * a2.eps)),
          sig((a1.sigma + a2.sigma) / 2.0),
          exponent((a1.exponent + a2.exponent) / 2.0),
          atom1(a1),
          atom2(a2) {}
    flt energy(Box &box) {
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt dsq = rij.squaredNorm();
        if (dsq > sig * sig) return 0.0;
        flt R = std::sqrt(dsq);
        return eps * std::pow(1.0 - (R / sig), exponent) / exponent;
    }
    Vec forces(Box &box) {
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt dsq = rij.squaredNorm();
        if (dsq > sig * sig) return Vec::Zero();
        flt R = std::sqrt(dsq);
        return rij * (eps * std::pow(1.0 - (R / sig), exponent - 1) / sig / R);
    }
};

// --------------------------------------------- device parameter tables
template <class A>
struct DeviceAtomTraits;

template <>
struct DeviceAtomTraits<IEpsSigCutAtom> {
    static const int kind = PARM_PAIR_LJ_ATTRACT_REPULSE;
    typedef std::tuple<vector<flt>, uint, flt, flt> Key;
    static Key key(const IEpsSigCutAtom &a) {
        return Key(a.epsilons, a.indx, a.sigma, a.sigcut);
    }
    static void table(const vector<Key> &types, DeviceState &d) {
        const size_t nt = types.size();
        d.t_eps.assign(nt * nt, 0.0);
        d.t_sigma.resize(nt);
        d.t_third.resize(nt);
        for (size_t t1 = 0; t1 < nt; t1++) {
            const vector<flt> &row = std::get<0>(types[t1]);
            for (size_t t2 = 0; t2 < nt; t2++) {
                uint ix = std::get<1>(types[t2]);
                if (ix >= row.size())
                    throw std::invalid_argument("IEpsSigCutAtom: epsilon index out of range");
                d.t_eps[t1 * nt + t2] = row[ix];
            }
            d.t_sigma[t1] = std::get<2>(types[t1]);
            d.t_third[t1] = std::get<3>(types[t1]);
        }
    }
};

template <>
struct DeviceAtomTraits<EpsSigExpAtom> {
    static const int kind = PARM_PAIR_REPULSION;
    typedef std::tuple<flt, flt, flt> Key;
    static Key key(const EpsSigExpAtom &a) { return Key(a.eps, a.sigma, a.exponent); }
    static void table(const vector<Key> &types, DeviceState &d) {
        const size_t nt = types.size();
        d.t_eps.resize(nt);
        d.t_sigma.resize(nt);
        d.t_third.resize(nt);
        for (size_t t = 0; t < nt; t++) {
            d.t_eps[t] = std::get<0>(types[t]);
            d.t_sigma[t] = std::get<1>(types[t]);
            d.t_third[t] = std::get<2>(types[t]);
        }
    }
};

// ----------------------------------------------------------- NListed
template <class A, class P>
class NListed : public Interaction {
   protected:
    vector<A> atoms;  // indexed by AtomVec index; filled for added atoms
    vector<bool> added;
    sptr<AtomVec> vecp;
    sptr<NeighborList> neighbors;
    bool params_dirty = true;

    DeviceState &dev() { return vecp->device(); }
    void claim() {
        DeviceState &d = dev();
        if (d.pair_owner && d.pair_owner != this)
            throw std::runtime_error(
                "the device path supports one NListed interaction per AtomVec");
        d.pair_owner = this;
    }
    //! Folds the per-atom parameters into the device type table.
    void push_params() {
        claim();
        if (!params_dirty) return;
        typedef DeviceAtomTraits<A> T;
        std::map<typename T::Key, int32_t> ids;
        vector<typename T::Key> types;
        DeviceState &d = dev();
        d.atom_type.assign(vecp->size(), 0);
        for (size_t i = 0; i < atoms.size(); i++) {
            if (!added[i]) continue;
            typename T::Key k = T::key(atoms[i]);
            auto it = ids.find(k);
            int32_t t;
            if (it == ids.end()) {
                t = (int32_t)types.size();
                ids[k] = t;
                types.push_back(k);
            } else {
                t = it->second;
            }
            d.atom_type[i] = t;
        }
        if (types.empty()) types.push_back(T::key(A()));
        if (types.size() > 64)
            throw std::runtime_error("the device path supports at most 64 distinct atom "
                                     "parameter sets per interaction");
        T::table(types, d);
        d.ntypes = (int)types.size();
        d.pair_kind = T::kind;
        d.pair_dirty = true;
        params_dirty = false;
    }
    void ready() {
        device_prepare();
        neighbors->push_config();
        dev().ensure();
    }
    void sums(double *U, double *W, double *S, uint64_t *nc, uint64_t *no) {
        ready();
        parm_check(parm_pair_sums(dev().sys, U, W, S, nc, no));
    }

   public:
    NListed(sptr<AtomVec> vec, sptr<NeighborList> neighbors)
        : atoms(vec->size()), added(vec->size(), false), vecp(vec), neighbors(neighbors) {
        claim();
    }
    NListed(sptr<Box> box, sptr<AtomVec> atomv, const flt skin)
        : atoms(atomv->size()),
          added(atomv->size(), false),
          vecp(atomv),
          neighbors(new NeighborList(box, atomv, skin)) {
        claim();
    }
    void add(A atm) {
        if (atm.n() >= atoms.size())
            throw std::invalid_argument("NListed::add: atom not in this AtomVec");
        neighbors->add(atm, atm.max_size());
        atoms[atm.n()] = atm;
        added[atm.n()] = true;
        params_dirty = true;
    }
    P get_pair(IDPair &pair) { return P(atoms[pair.first().n()], atoms[pair.last().n()]); }
    A &getatom(uint n) { return atoms[n]; }
    flt energy(Box &box, IDPair &pair) {
        P p = get_pair(pair);
        return p.energy(box);
    }
    flt energy(Box &) {
        double U = 0;
        sums(&U, nullptr, nullptr, nullptr, nullptr);
        return U;
    }
    unsigned long long contacts(Box &) {
        uint64_t nc = 0;
        sums(nullptr, nullptr, nullptr, &nc, nullptr);
        return nc;
    }
    unsigned long long overlaps(Box &) {
        uint64_t no = 0;
        sums(nullptr, nullptr, nullptr, nullptr, &no);
        return no;
    }
    flt pressure(Box &) {
        double W = 0;
        sums(nullptr, &W, nullptr, nullptr, nullptr);
        return W;
    }
    uint size() { return (uint)atoms.size(); }
    //! f += pair forces (src/interaction.hpp:2165-2175).
    void set_forces(Box &) {
        ready();
        parm_check(parm_add_pair_forces(dev().sys));
        dev().dev_newer = true;
    }
    flt set_forces_get_pressure(Box &box) {
        set_forces(box);
        return pressure(box);
    }
    Matrix stress(Box &) {
        Matrix S;
        double s[NDIM * NDIM];
        sums(nullptr, nullptr, s, nullptr, nullptr);
        for (int a = 0; a < NDIM; a++)
            for (int b = 0; b < NDIM; b++) S(a, b) = s[a * NDIM + b];
        return S;
    }
    Matrix set_forces_get_stress(Box &box) {
        set_forces(box);
        return stress(box);
    }
    vector<A> &atom_list() { return atoms; }
    sptr<NeighborList> neighbor_list() { return neighbors; }

    AtomVec *device_atoms() { return vecp.get(); }
    NeighborList *device_list() { return neighbors.get(); }
    void device_prepare() { push_params(); }
    ~NListed() {
        if (dev().pair_owner == this) dev().pair_owner = nullptr;
    }
};

PARM_NS_END
#endif
