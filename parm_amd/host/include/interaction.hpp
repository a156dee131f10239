// interaction.hpp -- drop-in for ParM's Interaction, NListed<A,P> and the
// neighbour-listed pair types the device path runs (src/interaction.hpp:
// 59-104, 119-147, 238-281, 857-987, 989-1018, 1251-1299, 1454-1590,
// 1876-1945, 2101-2291):
//   NListed<IEpsSigCutAtom, LJAttractRepulsePair>  (SWIG LJAttractRepulse)
//   NListed<EpsSigExpAtom, RepulsionPair>          (SWIG Repulsion)
//   NListed<EpsSigCutAtom, LennardJonesCutPair>    (src/bin/LJatoms.cpp)
//   NListed<IEpsISigCutAtom, LennardJonesCutPair>  (SWIG LJIICut)
//   NListed<EpsSigAtom, LJRepulsePair>             (SWIG LJRepulse)
//   NListed<IEpsISigExpAtom, RepulsionPair>        (SWIG RepulsionII)
//
// NListed<A,P> keeps the reference's interface.  Its per-atom parameters are
// folded into a type table (one type per distinct parameter tuple) that
// configures the device; energy, forces, pressure, stress, contacts and
// overlaps then run on the device over the device list.
#ifndef INTERACTION_H
#define INTERACTION_H

#include <algorithm>
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

struct EpsSigAtom : public AtomID {
    flt epsilon, sigma;
    EpsSigAtom() {}
    EpsSigAtom(AtomID a, flt epsilon, flt sigma) : AtomID(a), epsilon(epsilon), sigma(sigma) {}
    EpsSigAtom(AtomID a, EpsSigAtom other)
        : AtomID(a), epsilon(other.epsilon), sigma(other.sigma) {}
    flt max_size() { return sigma; }
};

struct EpsSigCutAtom : public EpsSigAtom {
    flt sigcut;
    EpsSigCutAtom() {}
    EpsSigCutAtom(AtomID a, flt epsilon, flt sigma, flt cut)
        : EpsSigAtom(a, epsilon, sigma), sigcut(cut) {}
    EpsSigCutAtom(AtomID a, EpsSigCutAtom other) : EpsSigAtom(a, other), sigcut(other.sigcut) {}
    flt max_size() { return sigma * sigcut; }
};

//! Row-indexed epsilons and sigmas: the pair (a1, a2) takes
//! a1.epsilons[a2.indx] and a1.sigmas[a2.indx].
struct IEpsISigCutAtom : public AtomID {
    vector<flt> epsilons, sigmas;
    uint indx;
    flt sigcut;
    IEpsISigCutAtom() {}
    IEpsISigCutAtom(AtomID a, vector<flt> epsilons, vector<flt> sigmas, uint indx, flt cut)
        : AtomID(a), epsilons(epsilons), sigmas(sigmas), indx(indx), sigcut(cut) {
        if (sigmas.size() != epsilons.size())
            throw std::invalid_argument("IEpsISigCutAtom: epsilons and sigmas differ in size");
    }
    IEpsISigCutAtom(AtomID a, IEpsISigCutAtom other)
        : AtomID(a), epsilons(other.epsilons), sigmas(other.sigmas), indx(other.indx),
          sigcut(other.sigcut) {}
    flt get_epsilon(IEpsISigCutAtom &other) { return row_at(epsilons, other.indx); }
    flt get_sigma(IEpsISigCutAtom &other) { return row_at(sigmas, other.indx); }
    flt max_size() { return *std::max_element(sigmas.begin(), sigmas.end()) * sigcut; }
    static flt row_at(const vector<flt> &row, uint i) {
        if (i >= row.size()) throw std::invalid_argument("parameter index out of range");
        return row[i];
    }
};

struct IEpsISigExpAtom : public AtomID {
    vector<flt> epsilons, sigmas;
    flt exponent;
    uint indx;
    IEpsISigExpAtom() {}
    IEpsISigExpAtom(AtomID a, vector<flt> epsilons, vector<flt> sigmas, uint indx,
                    flt exponent = 2.5)
        : AtomID(a), epsilons(epsilons), sigmas(sigmas), exponent(exponent), indx(indx) {
        if (sigmas.size() != epsilons.size())
            throw std::invalid_argument("IEpsISigExpAtom: epsilons and sigmas differ in size");
    }
    IEpsISigExpAtom(AtomID a, IEpsISigExpAtom other)
        : AtomID(a), epsilons(other.epsilons), sigmas(other.sigmas),
          exponent(other.exponent), indx(other.indx) {}
    flt get_epsilon(IEpsISigExpAtom &other) {
        return IEpsISigCutAtom::row_at(epsilons, other.indx);
    }
    flt get_sigma(IEpsISigExpAtom &other) { return IEpsISigCutAtom::row_at(sigmas, other.indx); }
    flt max_size() { return *std::max_element(sigmas.begin(), sigmas.end()); }
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
    RepulsionPair(IEpsISigExpAtom a1, IEpsISigExpAtom a2)
        : eps(a1.get_epsilon(a2)),
          sig(a1.get_sigma(a2)),
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

//! LennardJonesCut (src/interaction.hpp:238-281): eps((1 - r^-6)^2 - 1),
//! shifted to zero at the cut, the force not shifted.
struct LennardJonesCutPair {
    flt epsilon, sigma, cut_distance, cut_energy;
    AtomID atom1, atom2;
    LennardJonesCutPair(EpsSigCutAtom a1, EpsSigCutAtom a2)
        : epsilon(std::sqrt(a1.epsilon * a2.epsilon)),
          sigma((a1.sigma + a2.sigma) / 2),
          cut_distance(std::max(a1.sigcut, a2.sigcut)),
          atom1(a1),
          atom2(a2) {
        set_cut();
    }
    LennardJonesCutPair(IEpsISigCutAtom a1, IEpsISigCutAtom a2)
        : epsilon(a1.get_epsilon(a2)),
          sigma(a1.get_sigma(a2)),
          cut_distance(std::max(a1.sigcut, a2.sigcut)),
          atom1(a1),
          atom2(a2) {
        set_cut();
    }
    void set_cut() {
        flt rsix = std::pow(cut_distance, 6);
        flt mid = (1 - 1 / rsix);
        cut_energy = epsilon * (mid * mid - 1);
    }
    flt energy(Box &box) {
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt rsq = rij.squaredNorm() / (sigma * sigma);
        if (rsq > (cut_distance * cut_distance)) return 0;
        flt rsix = rsq * rsq * rsq;
        flt mid = (1 - 1 / rsix);
        return epsilon * (mid * mid - 1) - cut_energy;
    }
    Vec forces(Box &box) {
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt dsq = rij.squaredNorm();
        flt rsq = dsq / (sigma * sigma);
        if (rsq > (cut_distance * cut_distance)) return Vec::Zero();
        flt rsix = rsq * rsq * rsq;
        flt fmagTimesR = 12 * epsilon / rsix * (1 / rsix - 1);
        return rij * (fmagTimesR / dsq);
    }
};

//! Repulsive LJ cut at sigma (LJRepulsive, src/interaction.hpp:119-147).
struct LJRepulsePair {
    flt epsilon, sigma;
    AtomID atom1, atom2;
    LJRepulsePair(EpsSigAtom a1, EpsSigAtom a2)
        : epsilon(std::sqrt(a1.epsilon * a2.epsilon)),
          sigma((a1.sigma + a2.sigma) / 2),
          atom1(a1),
          atom2(a2) {}
    flt energy(Box &box) {
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt rsq = rij.squaredNorm() / (sigma * sigma);
        if (rsq > 1) return 0;
        flt rsix = rsq * rsq * rsq;
        flt mid = (1 - 1 / rsix);
        return epsilon * (mid * mid);
    }
    Vec forces(Box &box) {
        Vec rij = box.diff(atom1.peek().x, atom2.peek().x);
        flt dsq = rij.squaredNorm();
        flt rsq = dsq / (sigma * sigma);
        if (rsq > 1) return Vec::Zero();
        flt rsix = rsq * rsq * rsq;
        flt fmagTimesR = 12 * epsilon / rsix * (1 / rsix - 1);
        return rij * (fmagTimesR / dsq);
    }
};

// --------------------------------------------- device parameter tables
//! How an NListed<A, P> folds its atoms into the device's type table: the
//! C ABI pair form and, per distinct parameter tuple (a "type"), the
//! parameter rows it passes.
template <class A, class P>
struct DevicePairTraits;

//! Row-indexed parameters of types t1, t2 (the rows of t1 at the index of t2).
template <class K>
static void fill_rows(const vector<K> &types, vector<double> &out, int row_field) {
    const size_t nt = types.size();
    out.assign(nt * nt, 0.0);
    for (size_t t1 = 0; t1 < nt; t1++)
        for (size_t t2 = 0; t2 < nt; t2++) {
            const vector<flt> &row =
                row_field == 0 ? std::get<0>(types[t1]) : std::get<1>(types[t1]);
            out[t1 * nt + t2] = IEpsISigCutAtom::row_at(row, std::get<2>(types[t2]));
        }
}

template <>
struct DevicePairTraits<IEpsSigCutAtom, LJAttractRepulsePair> {
    static const int form = PARM_PAIR_LJ_ATTRACT_REPULSE;
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
struct DevicePairTraits<EpsSigExpAtom, RepulsionPair> {
    static const int form = PARM_PAIR_REPULSION;
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

template <>
struct DevicePairTraits<EpsSigCutAtom, LennardJonesCutPair> {
    static const int form = PARM_PAIR_LJ_CUT;
    typedef std::tuple<flt, flt, flt> Key;
    static Key key(const EpsSigCutAtom &a) { return Key(a.epsilon, a.sigma, a.sigcut); }
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

template <>
struct DevicePairTraits<IEpsISigCutAtom, LennardJonesCutPair> {
    static const int form = PARM_PAIR_LJ_CUT_INDEXED;
    typedef std::tuple<vector<flt>, vector<flt>, uint, flt> Key;
    static Key key(const IEpsISigCutAtom &a) {
        return Key(a.epsilons, a.sigmas, a.indx, a.sigcut);
    }
    static void table(const vector<Key> &types, DeviceState &d) {
        fill_rows(types, d.t_eps, 0);
        fill_rows(types, d.t_sigma, 1);
        d.t_third.resize(types.size());
        for (size_t t = 0; t < types.size(); t++) d.t_third[t] = std::get<3>(types[t]);
    }
};

template <>
struct DevicePairTraits<EpsSigAtom, LJRepulsePair> {
    static const int form = PARM_PAIR_LJ_REPULSE;
    typedef std::tuple<flt, flt> Key;
    static Key key(const EpsSigAtom &a) { return Key(a.epsilon, a.sigma); }
    static void table(const vector<Key> &types, DeviceState &d) {
        const size_t nt = types.size();
        d.t_eps.resize(nt);
        d.t_sigma.resize(nt);
        d.t_third.assign(nt, 0.0);
        for (size_t t = 0; t < nt; t++) {
            d.t_eps[t] = std::get<0>(types[t]);
            d.t_sigma[t] = std::get<1>(types[t]);
        }
    }
};

template <>
struct DevicePairTraits<IEpsISigExpAtom, RepulsionPair> {
    static const int form = PARM_PAIR_REPULSION_INDEXED;
    typedef std::tuple<vector<flt>, vector<flt>, uint, flt> Key;
    static Key key(const IEpsISigExpAtom &a) {
        return Key(a.epsilons, a.sigmas, a.indx, a.exponent);
    }
    static void table(const vector<Key> &types, DeviceState &d) {
        fill_rows(types, d.t_eps, 0);
        fill_rows(types, d.t_sigma, 1);
        d.t_third.resize(types.size());
        for (size_t t = 0; t < types.size(); t++) d.t_third[t] = std::get<3>(types[t]);
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
        typedef DevicePairTraits<A, P> T;
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
        d.pair_form = T::form;
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
