// trackers.hpp -- drop-in for ParM's StateTracker and NeighborList
// (src/trackers.hpp:18-31, 157-214; src/trackers.cpp:10-85) on the device.
//
// NeighborList keeps the reference's interface; its members, diameters and
// skin configure the device list of the AtomVec, and update_list / which /
// numpairs / get run on the device.  Pair export (get, begin, end,
// __getitem__) returns the reference's order: first by add order, then last
// by add order, last added before first.
#ifndef TRACKERS_H
#define TRACKERS_H

#include <algorithm>
#include <cassert>
#include <set>

#include "box.hpp"

PARM_NS_BEGIN

class StateTracker {
   public:
    virtual void update(Box &box) = 0;
    virtual bool every_collision() { return false; }
    virtual void update_collision(Box &, AtomID, AtomID, flt, Vec) {}
    virtual ~StateTracker() {}
};

class NeighborList : public StateTracker {
   protected:
    sptr<Box> box;
    flt skin;
    sptr<AtomVec> vecp;
    SubGroup atoms;
    vector<flt> diameters;
    vector<IDPair> curpairs;  // export cache
    uint cached_which = UINT_MAX;
    std::set<std::pair<uint, uint> > ignored;  // (larger, smaller) atom index

    DeviceState &dev() const { return vecp->device(); }
    void claim() {
        DeviceState &d = dev();
        if (d.list_owner && d.list_owner != this)
            throw std::runtime_error(
                "the device path supports one NeighborList per AtomVec");
        d.list_owner = this;
    }
   public:
    //! Pushes the box to the device state (the box may be shared and resized).
    void push_config() {
        DeviceState &d = dev();
        claim();
        OriginBox *ob = dynamic_cast<OriginBox *>(box.get());
        if (!ob) throw std::runtime_error("the device path supports OriginBox only");
        if (!d.have_box || !(d.box_L == ob->box_shape())) {
            d.box_L = ob->box_shape();
            d.have_box = true;
            d.box_dirty = true;
        }
    }

   protected:
    void refresh_pairs() {
        uint w = which();
        if (w == cached_which) return;
        curpairs.clear();
        int64_t np = numpairs();
        vector<int32_t> first(np > 0 ? np : 1), last(np > 0 ? np : 1);
        int64_t got = 0;
        if (np > 0)
            parm_check(parm_nlist_pairs(dev().sys, np, first.data(), last.data(), &got));
        curpairs.reserve(np);
        for (int64_t p = 0; p < np; p++)
            curpairs.push_back(IDPair(vecp->get_id(first[p]), vecp->get_id(last[p])));
        cached_which = w;
    }

   public:
    NeighborList(sptr<Box> box, sptr<AtomVec> atomv, const flt skin)
        : box(box), skin(skin), vecp(atomv), atoms(atomv) {
        claim();
        dev().skin = skin;
        dev().list_dirty = true;
        push_config();
    }
    //! Calls update_list(false) (src/trackers.hpp:173-176).
    void update(Box &newbox) {
        assert(&newbox == box.get());
        (void)newbox;
        update_list(false);
    }
    //! src/trackers.cpp:19-85, on the device.
    bool update_list(bool force = true) {
        push_config();
        DeviceState &d = dev();
        d.ensure();
        int rebuilt = 0;
        parm_check(parm_nlist_update(d.sys, force ? 1 : 0, &rebuilt));
        return rebuilt != 0;
    }
    AtomVec &vec() { return atoms.vec(); }
    uint which() {
        if (!dev().sys) return 0;
        uint32_t w = 0;
        parm_check(parm_nlist_which(dev().sys, &w));
        return w;
    }
    uint numpairs() {
        if (!dev().sys) return 0;
        int64_t n = 0;
        parm_check(parm_nlist_numpairs(dev().sys, &n));
        return (uint)n;
    }
    //! src/trackers.hpp:190-193: the pair never enters the list; the next
    //! update rebuilds.  Stored once per unordered pair (PairList).
    void ignore(AtomID a, AtomID b) {
        uint hi = std::max(a.n(), b.n()), lo = std::min(a.n(), b.n());
        ignored.insert(std::make_pair(hi, lo));
        DeviceState &d = dev();
        d.ign_a.push_back((int32_t)a.n());
        d.ign_b.push_back((int32_t)b.n());
    }
    void add(AtomID a, flt diameter) {
        atoms.add(a);
        diameters.push_back(diameter);
        DeviceState &d = dev();
        d.members.push_back((int32_t)a.n());
        d.diameters.push_back(diameter);
        d.list_dirty = true;  // like ignorechanged: rebuild at the next update
    }
    uint ignore_size() const { return (uint)ignored.size(); }
    uint size() const { return atoms.size(); }
    vector<IDPair>::iterator begin() {
        refresh_pairs();
        return curpairs.begin();
    }
    vector<IDPair>::iterator end() {
        refresh_pairs();
        return curpairs.end();
    }
    IDPair get(uint i) {
        refresh_pairs();
        return curpairs.at(i);
    }
    sptr<Box> get_box() { return box; }
    ~NeighborList() {
        if (dev().list_owner == this) dev().list_owner = nullptr;
    }
};

PARM_NS_END
#endif
