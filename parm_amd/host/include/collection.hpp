// collection.hpp -- drop-in for ParM's Collection, CollectionVerlet and
// CollectionDamped (src/collection.hpp:23-130, 273-295, 360-374;
// src/collection.cpp:3-208, 324-381, 442-469) on the MI355X device path.
//
// The device path runs a collection whose atoms are one AtomVec, whose
// interactions are at most one NListed on that AtomVec and whose trackers
// are that interaction's NeighborList; it has no constraints.  Anything else
// raises std::runtime_error (there is no CPU fallback).
#ifndef COLLECTION_H
#define COLLECTION_H

#include "interaction.hpp"

PARM_NS_BEGIN

//! Constraints exist so signatures match; the device path accepts none.
class Constraint {
   public:
    virtual ~Constraint() {}
};

class Collection {
   protected:
    sptr<Box> box;
    sptr<AtomGroup> atoms;
    vector<sptr<Interaction> > interactions;
    vector<sptr<StateTracker> > trackers;
    vector<sptr<Constraint> > constraints;

    void update_trackers() {
        for (auto &t : trackers) t->update(*box);
    }
    void update_constraint_positions() {}
    void update_constraint_velocities() {}
    void update_constraint_forces() {}

    AtomVec &atomvec() {
        AtomVec *av = dynamic_cast<AtomVec *>(atoms.get());
        if (!av) throw std::runtime_error("the device path needs an AtomVec as the atoms");
        return *av;
    }
    //! The single device interaction (nullptr when there is none).
    Interaction *device_interaction() {
        if (!constraints.empty())
            throw std::runtime_error("constraints are not supported by the device path");
        if (interactions.size() > 1)
            throw std::runtime_error(
                "the device path supports one interaction per collection");
        if (interactions.empty()) return nullptr;
        Interaction *I = interactions[0].get();
        if (I->device_atoms() != &atomvec())
            throw std::runtime_error(
                "the interaction is not a device NListed on this collection's atoms");
        return I;
    }
    //! Checks the configuration for a device timestep and returns the system.
    parm_system *device_step_system() {
        Interaction *I = device_interaction();
        if (!I) throw std::runtime_error("the device path needs an interaction to step");
        bool list_tracked = false;
        for (auto &t : trackers) {
            if (t.get() != I->device_list())
                throw std::runtime_error(
                    "the device path supports only the interaction's NeighborList as a "
                    "tracker");
            list_tracked = true;
        }
        if (!list_tracked)
            throw std::runtime_error(
                "the device path needs the interaction's NeighborList among the trackers");
        I->device_prepare();
        I->device_list()->push_config();
        DeviceState &d = atomvec().device();
        d.ensure();
        return d.sys;
    }

   public:
    Collection(sptr<Box> box, sptr<AtomGroup> atoms,
               vector<sptr<Interaction> > interactions = vector<sptr<Interaction> >(),
               vector<sptr<StateTracker> > trackers = vector<sptr<StateTracker> >(),
               vector<sptr<Constraint> > constraints = vector<sptr<Constraint> >(),
               bool should_initialize = true)
        : box(box),
          atoms(atoms),
          interactions(interactions),
          trackers(trackers),
          constraints(constraints) {
        if (should_initialize) initialize();
    }

    //! src/collection.cpp:13-19
    virtual void initialize() {
        update_trackers();
        update_constraint_positions();
        update_constraint_velocities();
        set_forces(true);
        update_trackers();
    }

    //! src/collection.cpp:159-179: f = sum of interaction forces; if
    //! constraints_and_a, a = f / m (0 for fixed atoms).
    virtual void set_forces(bool constraints_and_a = true) {
        Interaction *I = device_interaction();
        if (!I) {
            AtomVec &av = atomvec();
            for (uint i = 0; i < av.size(); i++) {
                Atom &a = av[i];
                a.f = Vec::Zero();
                if (constraints_and_a) a.a = Vec::Zero();
            }
            return;
        }
        I->device_prepare();
        I->device_list()->push_config();
        DeviceState &d = atomvec().device();
        d.ensure();
        parm_check(parm_set_forces(d.sys, constraints_and_a ? 1 : 0));
        d.dev_newer = true;
    }

    virtual void timestep() = 0;

    flt degrees_of_freedom() {
        int ndof = 0;
        AtomVec &av = atomvec();
        for (uint i = 0; i < av.size(); i++) {
            const Atom &a = av.get_id(i).peek();
            if (a.m <= 0 || std::isinf(a.m)) continue;
            ndof += NDIM;
        }
        return ndof;
    }
    flt potential_energy() {
        flt E = 0;
        for (auto &I : interactions) E += I->energy(*box);
        return E;
    }
    flt energy() { return potential_energy() + kinetic_energy(); }
    virtual flt temp(bool minuscomv = true) {
        Vec v = Vec::Zero();
        if (minuscomv) v = com_velocity();
        int ndof = (int)degrees_of_freedom();
        if (minuscomv) ndof -= NDIM;
        return atoms->kinetic_energy(v) * 2 / ndof;
    }
    virtual flt kinetic_energy() { return atoms->kinetic_energy(); }
    virtual flt virial() {
        flt E = 0;
        for (auto &I : interactions) E += I->pressure(*box);
        return E;
    }
    //! (2K + sum r.f) / (NDIM V)  (src/collection.cpp:85-96)
    virtual flt pressure() {
        flt V = box->V();
        flt E = 2.0 * kinetic_energy();
        for (auto &I : interactions) E += I->pressure(*box);
        return E / V / flt(NDIM);
    }
    sptr<Box> get_box() { return box; }
    Vec com() { return atoms->com(); }
    Vec com_velocity() { return atoms->momentum() / atoms->mass(); }
    void reset_com_velocity() { atoms->add_velocity(-com_velocity()); }
    void scale_velocities(flt scaleby) {
        AtomVec &av = atomvec();
        for (uint i = 0; i < av.size(); i++) {
            Atom &a = av[i];
            if (a.m <= 0 || std::isinf(a.m)) continue;
            a.v *= scaleby;
        }
    }
    void scale_velocities_to_temp(flt T, bool minuscomv = true) {
        flt t = temp(minuscomv);
        scale_velocities(std::sqrt(T / t));
    }
    void scale_velocities_to_energy(flt E) {
        flt E0 = energy();
        flt k0 = kinetic_energy();
        flt goalkinetic = k0 + (E - E0);
        scale_velocities(std::sqrt(goalkinetic / k0));
    }
    virtual void add_interaction(sptr<Interaction> inter) {
        interactions.push_back(inter);
        update_trackers();
    }
    virtual void add_tracker(sptr<StateTracker> track) {
        trackers.push_back(track);
        update_trackers();
    }
    virtual void add_constraint(sptr<Constraint> c) {
        constraints.push_back(c);
        update_trackers();
    }
    void add(sptr<Interaction> a) { add_interaction(a); }
    void add(sptr<StateTracker> a) { add_tracker(a); }
    void add(sptr<Constraint> a) { add_constraint(a); }
    vector<sptr<Interaction> > get_interactions() { return interactions; }
    virtual ~Collection() {}
};

//! NVE velocity Verlet (src/collection.hpp:360-374, collection.cpp:442-469).
class CollectionVerlet : public Collection {
   protected:
    flt dt;

   public:
    CollectionVerlet(sptr<Box> box, sptr<AtomGroup> atoms, const flt dt,
                     vector<sptr<Interaction> > interactions = vector<sptr<Interaction> >(),
                     vector<sptr<StateTracker> > trackers = vector<sptr<StateTracker> >(),
                     vector<sptr<Constraint> > constraints = vector<sptr<Constraint> >())
        : Collection(box, atoms, interactions, trackers, constraints), dt(dt) {}
    void timestep() { timesteps(1); }
    //! Device-path extension: n timesteps in one call (same result as n calls
    //! of timestep(), fused on the device).
    void timesteps(uint64_t n) {
        parm_system *s = device_step_system();
        parm_check(parm_verlet_steps(s, dt, (int64_t)n));
        atomvec().device().dev_newer = true;
    }
    void set_dt(flt newdt) { dt = newdt; }
};

//! Damped Verlet without noise (src/collection.hpp:273-295,
//! collection.cpp:324-381).
class CollectionDamped : public Collection {
   protected:
    flt dt;
    flt damping;

   public:
    CollectionDamped(sptr<Box> box, sptr<AtomGroup> atoms, const flt dt, const flt damping,
                     vector<sptr<Interaction> > interactions = vector<sptr<Interaction> >(),
                     vector<sptr<StateTracker> > trackers = vector<sptr<StateTracker> >(),
                     vector<sptr<Constraint> > constraints = vector<sptr<Constraint> >())
        : Collection(box, atoms, interactions, trackers, constraints, false),
          dt(dt),
          damping(damping) {
        if (dt <= 0) throw std::invalid_argument("Collection::CollectionSol: dt >= 0");
        initialize();
    }
    void change_damping(const flt damp) { damping = damp; }
    void set_dt(const flt newdt) { dt = newdt; }
    void timestep() { timesteps(1); }
    void timesteps(uint64_t n) {
        parm_system *s = device_step_system();
        parm_check(parm_damped_steps(s, dt, damping, (int64_t)n));
        atomvec().device().dev_newer = true;
    }
};

//! Pressure-controlled nonlinear conjugate-gradient energy minimiser
//! (src/collection.hpp:400-474, src/collection.cpp:494-855), minimising
//! H = U + P0 V over the positions and the box: the device runs it
//! (parm_nlcg_*); the public members mirror the reference's and are
//! exchanged with the device around every call.
class CollectionNLCG : public Collection {
   public:
    // parameters
    flt dt;
    flt seceps;
    uint secmax;
    flt kappa;
    flt alphamax, afrac, dxmax, stepmax, kmax;
    flt P0;
    // kept between iterations
    flt Knew, k, vl, fl, al;
    // for tracking
    flt alpha, beta, betaused, dxsum, alphavmax, maxdV;
    uint sec;

   protected:
    parm_nlcg *dev_c = nullptr;
    bool fixedL = false;

    OriginBox &obox() {
        OriginBox *ob = dynamic_cast<OriginBox *>(box.get());
        if (!ob) throw std::runtime_error("CollectionNLCG needs an OriginBox");
        return *ob;
    }
    //! The device minimiser, created on first use (the collection's
    //! interaction and list are fixed from then on); members pushed.
    parm_nlcg *device_nlcg() {
        parm_system *s = device_step_system();
        if (!dev_c)
            parm_check(parm_nlcg_create(s, dt, P0, kappa, kmax, (int)secmax, seceps,
                                        fixedL ? 1 : 0, &dev_c));
        parm_check(parm_nlcg_set_params(dev_c, dt, P0, kappa, kmax, (int)secmax, seceps));
        parm_check(parm_nlcg_set_limits(dev_c, alphamax, afrac, dxmax, stepmax, maxdV));
        return dev_c;
    }
    //! Members and the box back from the device.
    void pull() {
        double o[12];
        parm_check(parm_nlcg_info(dev_c, o));
        alpha = o[0];
        beta = o[1];
        betaused = o[2];
        dxsum = o[3];
        alphavmax = o[4];
        sec = (uint)o[5];
        k = o[6];
        vl = o[7];
        fl = o[8];
        al = o[9];
        Knew = o[10];
        DeviceState &d = atomvec().device();
        Vec L;
        parm_check(parm_get_box(d.sys, L.data()));
        obox().resize_to(L);
        d.box_L = L;  // the device already has it
        d.dev_newer = true;
    }

   public:
    CollectionNLCG(sptr<OriginBox> box, sptr<AtomGroup> atoms, const flt dt, const flt P0,
                   vector<sptr<Interaction> > interactions = vector<sptr<Interaction> >(),
                   vector<sptr<StateTracker> > trackers = vector<sptr<StateTracker> >(),
                   vector<sptr<Constraint> > constraints = vector<sptr<Constraint> >(),
                   const flt kappa = 10.0, const flt kmax = 1000, const uint secmax = 40,
                   const flt seceps = 1e-20)
        : Collection(box, atoms, interactions, trackers, constraints),
          dt(dt),
          seceps(seceps),
          secmax(secmax),
          kappa(kappa),
          alphamax(2.0),
          afrac(0),
          dxmax(100),
          stepmax(1e-3),
          kmax(kmax),
          P0(P0),
          Knew(0),
          k(0),
          vl(0),
          fl(0),
          al(0),
          alpha(0),
          beta(0),
          betaused(0),
          dxsum(0),
          alphavmax(0),
          maxdV(0),
          sec(0) {}
    ~CollectionNLCG() {
        if (dev_c) parm_nlcg_destroy(dev_c);
    }

    flt kinetic_energy() {
        double K = 0;
        parm_check(parm_nlcg_observables(device_nlcg(), nullptr, nullptr, &K));
        return K;
    }
    flt pressure() {
        double P = 0;
        parm_check(parm_nlcg_observables(device_nlcg(), nullptr, &P, nullptr));
        return P;
    }
    flt hamiltonian() {
        double H = 0;
        parm_check(parm_nlcg_observables(device_nlcg(), &H, nullptr, nullptr));
        return H;
    }
    void set_forces(bool constraints_and_a = true) { set_forces(constraints_and_a, true); }
    void set_forces(bool constraints_and_a, bool setV) {
        if (!device_interaction()) {
            Collection::set_forces(constraints_and_a);
            return;
        }
        parm_check(parm_nlcg_set_forces(device_nlcg(), constraints_and_a, setV));
        pull();
    }
    void timestep() { timesteps(1); }
    //! Device-path extension: n timesteps in one call.
    void timesteps(uint64_t n) {
        parm_check(parm_nlcg_timestep(device_nlcg(), (int64_t)n));
        pull();
    }
    void descend() {
        parm_check(parm_nlcg_descend(device_nlcg()));
        pull();
    }
    void reset() {
        parm_check(parm_nlcg_reset(device_nlcg()));
        pull();
    }
    void set_dt(flt newdt) {
        dt = newdt;
        reset();
    }
    void set_pressure_goal(flt P) {
        P0 = P;
        reset();
    }
    flt get_pressure_goal() { return P0; }
    void set_kappa(flt kk) {
        kappa = kk;
        reset();
    }
    void set_max_alpha(flt a) { alphamax = a; }
    void set_max_alpha_fraction(flt a) { afrac = a; }
    void set_max_dx(flt d) { dxmax = d; }
    void set_max_step(flt m) { stepmax = m; }
};

//! Like CollectionNLCG, but only the first box edge changes
//! (src/collection.hpp:476-488, src/collection.cpp:855-870).
class CollectionNLCGFixedL : public CollectionNLCG {
   public:
    CollectionNLCGFixedL(
        sptr<OriginBox> box, sptr<AtomGroup> atoms, const flt dt, const flt P0,
        vector<sptr<Interaction> > interactions = vector<sptr<Interaction> >(),
        vector<sptr<StateTracker> > trackers = vector<sptr<StateTracker> >(),
        vector<sptr<Constraint> > constraints = vector<sptr<Constraint> >(),
        const flt kappa = 10.0, const flt kmax = 1000, const uint secmax = 40,
        const flt seceps = 1e-20)
        : CollectionNLCG(box, atoms, dt, P0, interactions, trackers, constraints, kappa, kmax,
                         secmax, seceps) {
        fixedL = true;
    }
};

PARM_NS_END
#endif
