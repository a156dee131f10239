// box.hpp -- drop-in for ParM's src/box.hpp on the MI355X device path:
// Box, OriginBox, Atom, AtomID, IDPair, AtomGroup, AtomVec, SubGroup.
//
// Same names and signatures as the reference (src/box.hpp:47-506), with
// std::shared_ptr for sptr.  An AtomVec keeps a host copy of its atoms and a
// device copy (libparm_hip.so, include/parm_hip.h) and moves state between
// them on demand: element access through operator[], get_id or an AtomID
// first brings the host copy up to date, then marks it as the newer copy, so
// the next device call uploads it.  Calls that run on the device (list
// updates, forces, timesteps) leave the device copy newer.
#ifndef BOX_H
#define BOX_H

#include <algorithm>
#include <cstdlib>
#include <climits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "parm_hip.h"
#include "vecrand.hpp"

PARM_NS_BEGIN

#define sptr std::shared_ptr

typedef const unsigned int cuint;

class AtomGroup;
class AtomVec;

//! Throws the C++ exception the reference would raise for a failed device call.
inline void parm_check(int rc) {
    if (rc == PARM_OK) return;
    std::string msg = parm_last_error();
    if (rc == PARM_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

class Box {
   public:
    virtual Vec diff(Vec r1, Vec r2) = 0;
    virtual flt V() = 0;
    virtual ~Box() {}
};

//! The modulus function for Vec: remainder per component (src/box.hpp:68-78).
inline Vec vec_mod(Vec r1, Vec r2) {
    Vec r;
    for (int k = 0; k < NDIM; k++) r[k] = std::remainder(r1[k], r2[k]);
    return r;
}

//! A rectilinear Box with periodic boundary conditions (src/box.hpp:97-157).
class OriginBox : public Box {
   protected:
    Vec boxsize;

   public:
    OriginBox(Vec size) : boxsize(size) {}
    OriginBox(flt L) {
        for (int k = 0; k < NDIM; k++) boxsize[k] = L;
    }
    Vec diff(Vec r1, Vec r2) { return vec_mod((r1 - r2), boxsize); }
    flt V() {
        flt v = 1;
        for (int k = 0; k < NDIM; k++) v *= boxsize[k];
        return v;
    }
    flt L() {
        flt s = 0;
        for (int k = 0; k < NDIM; k++) s += boxsize[k];
        return s / NDIM;
    }
    Vec rand_loc() {
        Vec v = rand_vec_boxed();
        for (int k = 0; k < NDIM; k++) v[k] *= boxsize[k];
        return diff(v, Vec::Zero());
    }
    Vec box_shape() { return boxsize; }
    //! src/box.cpp:3-70 (box only; the atoms are not scaled)
    flt resize(flt factor) {
        boxsize *= factor;
        return V();
    }
    flt resize_to(Vec newsize) {
        boxsize = newsize;
        return V();
    }
    flt resize_to_V(flt newV) { return resize(std::pow(newV / V(), OVERNDIM)); }
    flt resize_to_L(flt newL) { return resize(newL / std::pow(V(), OVERNDIM)); }
};

struct Atom {
    Vec x;  //!< location
    Vec v;  //!< velocity
    Vec a;  //!< acceleration
    Vec f;  //!< forces
    flt m;  //!< mass
};

//! Host/device coherence of one AtomVec.  Owned by the AtomVec; the
//! NeighborList and the NListed interaction attached to it push their
//! configuration here.
class DeviceState {
   public:
    parm_system *sys = nullptr;
    int precision = PARM_FP64;
    bool host_dirty = true;  // host copy newer than the device copy
    bool dev_newer = false;  // device copy newer than the host copy
    // configuration pushed on the next ensure()
    bool box_dirty = false, list_dirty = false, pair_dirty = false;
    Vec box_L;
    bool have_box = false;
    const void *list_owner = nullptr;
    const void *pair_owner = nullptr;
    double skin = 0;
    std::vector<int32_t> members;
    std::vector<double> diameters;
    // NeighborList::ignore pairs, pushed after the list configuration
    std::vector<int32_t> ign_a, ign_b;
    size_t ign_pushed = 0;
    int pair_form = PARM_PAIR_NONE;  // enum parm_pair_kind
    int ntypes = 0;
    std::vector<double> t_eps, t_sigma, t_third;
    std::vector<int32_t> atom_type;
    Atom *atoms = nullptr;
    uint n = 0;

    DeviceState(Atom *a, uint n_) : atoms(a), n(n_) {
        if (const char *e = std::getenv("PARM_PRECISION"))
            if (std::string(e) == "mixed") precision = PARM_MIXED;
    }
    ~DeviceState() {
        if (sys) parm_system_destroy(sys);
    }
    DeviceState(const DeviceState &) = delete;
    DeviceState &operator=(const DeviceState &) = delete;

    //! Device copy complete and current; configuration pushed.
    void ensure() {
        if (!sys) {
            std::vector<double> m(n);
            for (uint i = 0; i < n; i++) m[i] = atoms[i].m;
            parm_check(parm_system_create(NDIM, precision, n, m.data(), &sys));
            host_dirty = true;
            box_dirty = have_box;
            list_dirty = list_owner != nullptr;
            pair_dirty = pair_owner != nullptr && ntypes > 0;
        }
        if (box_dirty) {
            parm_check(parm_set_box(sys, box_L.data()));
            box_dirty = false;
        }
        if (list_dirty) {
            parm_check(parm_nlist_set(sys, skin, (int64_t)members.size(), members.data(),
                                      diameters.data()));
            list_dirty = false;
            ign_pushed = 0;  // a new device list starts without exclusions
        }
        if (ign_pushed < ign_a.size()) {
            parm_check(parm_nlist_ignore(sys, (int64_t)(ign_a.size() - ign_pushed),
                                         ign_a.data() + ign_pushed, ign_b.data() + ign_pushed));
            ign_pushed = ign_a.size();
        }
        if (pair_dirty && ntypes > 0) {
            atom_type.resize(n, 0);
            const double *e = t_eps.data(), *sg = t_sigma.data(), *th = t_third.data();
            const int32_t *ty = atom_type.data();
            switch (pair_form) {
                case PARM_PAIR_LJ_ATTRACT_REPULSE:
                    parm_check(parm_pair_lj_attract_repulse(sys, ntypes, e, sg, th, ty));
                    break;
                case PARM_PAIR_REPULSION:
                    parm_check(parm_pair_repulsion(sys, ntypes, e, sg, th, ty));
                    break;
                case PARM_PAIR_LJ_CUT:
                    parm_check(parm_pair_lj_cut(sys, ntypes, e, sg, th, ty));
                    break;
                case PARM_PAIR_LJ_CUT_INDEXED:
                    parm_check(parm_pair_lj_cut_indexed(sys, ntypes, e, sg, th, ty));
                    break;
                case PARM_PAIR_LJ_REPULSE:
                    parm_check(parm_pair_lj_repulse(sys, ntypes, e, sg, ty));
                    break;
                case PARM_PAIR_REPULSION_INDEXED:
                    parm_check(parm_pair_repulsion_indexed(sys, ntypes, e, sg, th, ty));
                    break;
                default:
                    throw std::runtime_error("unknown pair form");
            }
            pair_dirty = false;
        }
        if (host_dirty) {
            std::vector<double> x(n * NDIM), v(n * NDIM), a(n * NDIM), f(n * NDIM), m(n);
            for (uint i = 0; i < n; i++) {
                for (int k = 0; k < NDIM; k++) {
                    x[i * NDIM + k] = atoms[i].x[k];
                    v[i * NDIM + k] = atoms[i].v[k];
                    a[i * NDIM + k] = atoms[i].a[k];
                    f[i * NDIM + k] = atoms[i].f[k];
                }
                m[i] = atoms[i].m;
            }
            parm_check(parm_set_masses(sys, m.data()));
            parm_check(parm_set_state(sys, x.data(), v.data(), a.data(), f.data()));
            host_dirty = false;
        }
    }
    //! Host copy current (downloads after device work).
    void pull() {
        if (!dev_newer || !sys) return;
        std::vector<double> x(n * NDIM), v(n * NDIM), a(n * NDIM), f(n * NDIM);
        parm_check(parm_get_state(sys, x.data(), v.data(), a.data(), f.data()));
        for (uint i = 0; i < n; i++) {
            for (int k = 0; k < NDIM; k++) {
                atoms[i].x[k] = x[i * NDIM + k];
                atoms[i].v[k] = v[i * NDIM + k];
                atoms[i].a[k] = a[i * NDIM + k];
                atoms[i].f[k] = f[i * NDIM + k];
            }
        }
        dev_newer = false;
    }
    //! Host copy current and about to be modified.
    void touch() {
        pull();
        host_dirty = true;
    }
};

class AtomRef {
   protected:
    Atom *ptr;
    DeviceState *dev;

   public:
    AtomRef() : ptr(NULL), dev(NULL) {}
    AtomRef(Atom *a, DeviceState *d = NULL) : ptr(a), dev(d) {}
    Atom &operator*() const {
        if (dev) dev->touch();
        return *ptr;
    }
    Atom *operator->() const {
        if (dev) dev->touch();
        return ptr;
    }
    bool operator==(const AtomRef &o) const { return o.ptr == ptr; }
    bool operator==(const Atom *o) const { return o == ptr; }
    bool operator!=(const AtomRef &o) const { return o.ptr != ptr; }
    bool operator<(const AtomRef &o) const { return ptr < o.ptr; }
    bool operator<=(const AtomRef &o) const { return ptr <= o.ptr; }
    bool operator>=(const AtomRef &o) const { return ptr >= o.ptr; }
    bool operator>(const AtomRef &o) const { return ptr > o.ptr; }
    bool is_null() { return ptr == NULL; }
    //! Read-only access that does not mark the host copy modified.
    const Atom &peek() const {
        if (dev) dev->pull();
        return *ptr;
    }
};

//! A pointer to an Atom that also knows its index in its AtomVec
//! (src/box.hpp:288-298).
class AtomID : public AtomRef {
   private:
    uint num;

   public:
    AtomID() : AtomRef(), num(UINT_MAX) {}
    AtomID(Atom *a, uint n) : AtomRef(a), num(n) {}
    AtomID(Atom *a, uint n, DeviceState *d) : AtomRef(a, d), num(n) {}
    uint n() const { return num; }
};

class IDPair {
   private:
    AtomID id1, id2;

   public:
    IDPair() : id1(), id2() {}
    IDPair(AtomID a, AtomID b) : id1(a), id2(b) {}
    AtomID first() const { return id1; }
    AtomID last() const { return id2; }
};

//! A group of atoms (src/box.hpp:311-431).  The observables below are the
//! reference's definitions; AtomVec computes them on the device.
class AtomGroup {
   public:
    virtual AtomVec &vec() = 0;
    virtual Atom &operator[](cuint n) = 0;
    virtual Atom &operator[](cuint n) const = 0;
    virtual Atom &get(cuint n) { return ((*this)[n]); }
    virtual AtomID get_id(cuint n) = 0;
    virtual uint size() const = 0;

    virtual Vec com() const {
        Vec v = Vec::Zero();
        for (uint i = 0; i < size(); i++) {
            Atom &a = (*this)[i];
            if (a.m <= 0 || std::isinf(a.m)) continue;
            v += a.x * a.m;
        }
        return v / mass();
    }
    virtual flt mass() const {
        flt m = 0;
        for (uint i = 0; i < size(); i++) {
            Atom &a = (*this)[i];
            if (a.m <= 0 || std::isinf(a.m)) continue;
            m += a.m;
        }
        return m;
    }
    virtual Vec momentum() const {
        Vec tot = Vec::Zero();
        for (uint i = 0; i < size(); i++) {
            Atom &a = (*this)[i];
            if (a.m <= 0 || std::isinf(a.m)) continue;
            tot += a.v * a.m;
        }
        return tot;
    }
    Vec com_velocity() const { return momentum() / mass(); }
    virtual flt kinetic_energy(const Vec originvelocity = Vec::Zero()) const {
        flt totE = 0;
        for (uint i = 0; i < size(); i++) {
            Atom &a = (*this)[i];
            if (a.m == 0 || std::isinf(a.m)) continue;
            Vec c = a.v - originvelocity;
            totE += a.m / 2 * c.dot(c);
        }
        return totE;
    }
    void add_velocity(Vec v) {
        for (uint i = 0; i < size(); i++) (*this)[i].v += v;
    }
    void reset_com_velocity() { add_velocity(-com_velocity()); }
    void randomize_velocities(flt T) {
        for (uint i = 0; i < size(); i++) {
            Atom &a = (*this)[i];
            if (a.m == 0 || std::isinf(a.m)) continue;
            a.v = rand_vec() * std::sqrt(T / a.m);
        }
    }
    void reset_forces() {
        for (uint i = 0; i < size(); i++) (*this)[i].f = Vec::Zero();
    }
    virtual ~AtomGroup() {}
};

//! The main class for representing particles (src/box.hpp:441-479).
class AtomVec : public virtual AtomGroup {
   private:
    Atom *atoms;
    uint sz;
    std::unique_ptr<DeviceState> dev_;
    void zero() {
        for (uint i = 0; i < sz; i++) {
            atoms[i].x = Vec::Zero();
            atoms[i].v = Vec::Zero();
            atoms[i].f = Vec::Zero();
            atoms[i].a = Vec::Zero();
        }
    }

   public:
    AtomVec(vector<double> masses) : sz((uint)masses.size()) {
        atoms = new Atom[sz > 0 ? sz : 1];
        for (uint i = 0; i < sz; i++) atoms[i].m = masses[i];
        zero();
        dev_.reset(new DeviceState(atoms, sz));
    }
    AtomVec(uint N, flt mass) : sz(N) {
        atoms = new Atom[sz > 0 ? sz : 1];
        for (uint i = 0; i < sz; i++) atoms[i].m = mass;
        zero();
        dev_.reset(new DeviceState(atoms, sz));
    }
    AtomVec(AtomVec &other) : sz(other.size()) {
        atoms = new Atom[sz > 0 ? sz : 1];
        other.dev_->pull();
        for (uint i = 0; i < sz; i++) atoms[i] = other.atoms[i];
        dev_.reset(new DeviceState(atoms, sz));
    }
    AtomVec &vec() { return *this; }
    Atom &operator[](cuint n) {
        dev_->touch();
        return atoms[n];
    }
    Atom &operator[](cuint n) const {
        dev_->touch();
        return atoms[n];
    }
    AtomID get_id(cuint n) {
        if (n > sz) return AtomID();
        return AtomID(atoms + n, n, dev_.get());
    }
    uint size() const { return sz; }

    // Observables on the device (AtomGroup's definitions).
    flt kinetic_energy(const Vec originvelocity = Vec::Zero()) const {
        dev_->ensure();
        double K = 0;
        parm_check(parm_kinetic_energy(dev_->sys, originvelocity.data(), &K));
        return K;
    }
    Vec momentum() const {
        dev_->ensure();
        Vec p;
        double m = 0;
        parm_check(parm_momentum(dev_->sys, p.data(), &m));
        return p;
    }
    flt mass() const {
        dev_->ensure();
        Vec p;
        double m = 0;
        parm_check(parm_momentum(dev_->sys, p.data(), &m));
        return m;
    }

    //! Device-path extension (no reference counterpart): PARM_FP64 (default)
    //! or PARM_MIXED; takes effect when the device copy is first created.
    void set_device_precision(int precision) {
        if (dev_->sys) throw std::runtime_error("device precision is fixed once created");
        dev_->precision = precision;
    }
    DeviceState &device() const { return *dev_; }

    //! Device-path extension (no reference counterpart): queues an
    //! asynchronous copy of the device positions (and velocities) for output
    //! and returns its ticket; the copy travels while later steps run.
    int64_t snapshot(bool positions = true, bool velocities = false) const {
        dev_->ensure();
        int64_t t = 0;
        parm_check(parm_snapshot_begin(dev_->sys, (positions ? 1 : 0) | (velocities ? 2 : 0),
                                       &t));
        return t;
    }
    //! Writes snapshot `ticket` as one xyz frame (the format of the
    //! reference's drivers): elements holds one letter per atom, or is empty
    //! for all 'C'; precision as an ostream's (6 by default).
    void write_xyz(int64_t ticket, const std::string &path, bool append = true,
                   const std::string &comment = "", const std::string &elements = "",
                   int precision = 6) const {
        if (!elements.empty() && elements.size() != sz)
            throw std::invalid_argument("write_xyz: one element letter per atom");
        parm_check(parm_snapshot_write_xyz(dev_->sys, ticket, path.c_str(), append ? 1 : 0,
                                           comment.c_str(),
                                           elements.empty() ? nullptr : elements.c_str(),
                                           precision));
    }

    ~AtomVec() { delete[] atoms; }
};

//! Any grouping of atoms of one AtomVec (src/box.hpp:486-506).
//! The duplicate check is O(1) (a membership bitmap) instead of the
//! reference's linear std::find, with the same exception.
class SubGroup : public AtomGroup {
   protected:
    sptr<AtomVec> atoms;
    vector<AtomID> ids;
    vector<bool> member;

   public:
    SubGroup(sptr<AtomVec> atoms) : atoms(atoms), member(atoms->size(), false) {}
    AtomVec &vec() { return *atoms; }
    Atom &operator[](cuint n) { return *ids[n]; }
    Atom &operator[](cuint n) const { return *ids[n]; }
    Atom &get(cuint n) { return *ids[n]; }
    void add(AtomID a) {
        if (a.n() >= member.size())
            throw std::invalid_argument("Cannot add AtomID to SubGroup: not in its AtomVec.");
        if (member[a.n()])
            throw std::invalid_argument("Cannot add AtomID to SubGroup: it already exists.");
        member[a.n()] = true;
        return ids.push_back(a);
    }
    AtomID get_id(cuint n) { return ids[n]; }
    uint size() const { return (uint)ids.size(); }
};

PARM_NS_END
#endif
