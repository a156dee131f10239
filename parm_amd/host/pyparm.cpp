// pyparm.cpp -- Python bindings of the drop-in C++ API, with the names the
// reference's SWIG modules export (src/sim.i:1-14, 494-670): built twice,
// as parm_amd._d3 (-DVEC3D) and parm_amd._d2 (-DVEC2D).
//
// Vec and Matrix convert to and from numpy arrays, as the reference's
// typemaps do (src/sim.i:159-428).  Container and iteration helpers follow
// the reference's %extend blocks: AtomVec.__getitem__ returns an AtomID,
// AtomID proxies x/v/a/f/m, NeighborList.__getitem__ raises IndexError past
// the end.  Errors arrive as ValueError (std::invalid_argument) or
// RuntimeError (std::runtime_error), like the SWIG exception map.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#ifdef VEC2D
#define PARM_NS_BEGIN namespace parm_d2 {
#else
#define PARM_NS_BEGIN namespace parm_d3 {
#endif
#define PARM_NS_END }
#include "collection.hpp"

#ifdef VEC2D
using namespace parm_d2;
#else
using namespace parm_d3;
#endif

namespace py = pybind11;

namespace pybind11 {
namespace detail {
template <>
struct type_caster<Vec> {
   public:
    PYBIND11_TYPE_CASTER(Vec, _("Vec"));
    bool load(handle src, bool) {
        if (!src) return false;
        py::array_t<double, py::array::forcecast> a =
            py::array_t<double, py::array::forcecast>::ensure(src);
        if (!a || a.ndim() != 1 || a.shape(0) != NDIM) return false;
        for (int k = 0; k < NDIM; k++) value[k] = a.at(k);
        return true;
    }
    static handle cast(const Vec &v, return_value_policy, handle) {
        py::array_t<double> a(NDIM);
        for (int k = 0; k < NDIM; k++) a.mutable_at(k) = v[k];
        return a.release();
    }
};
template <>
struct type_caster<Matrix> {
   public:
    PYBIND11_TYPE_CASTER(Matrix, _("Matrix"));
    bool load(handle, bool) { return false; }
    static handle cast(const Matrix &m, return_value_policy, handle) {
        py::array_t<double> a({NDIM, NDIM});
        for (int i = 0; i < NDIM; i++)
            for (int j = 0; j < NDIM; j++) a.mutable_at(i, j) = m(i, j);
        return a.release();
    }
};
}  // namespace detail
}  // namespace pybind11

typedef NListed<EpsSigExpAtom, RepulsionPair> Repulsion;
typedef NListed<IEpsSigCutAtom, LJAttractRepulsePair> LJAttractRepulse;
typedef NListed<EpsSigAtom, LJRepulsePair> LJRepulse;
typedef NListed<IEpsISigCutAtom, LennardJonesCutPair> LJIICut;
typedef NListed<IEpsISigExpAtom, RepulsionPair> RepulsionII;

// Bulk state access (extension): natoms x NDIM arrays in AtomVec order.
static py::array_t<double> get_field(AtomVec &av, int which) {
    py::array_t<double> out({(py::ssize_t)av.size(), (py::ssize_t)NDIM});
    auto o = out.mutable_unchecked<2>();
    for (uint i = 0; i < av.size(); i++) {
        const Atom &a = av.get_id(i).peek();
        const Vec &v = which == 0 ? a.x : which == 1 ? a.v : which == 2 ? a.a : a.f;
        for (int k = 0; k < NDIM; k++) o(i, k) = v[k];
    }
    return out;
}
static void set_field(AtomVec &av, int which,
                      py::array_t<double, py::array::c_style | py::array::forcecast> in) {
    if (in.ndim() != 2 || in.shape(0) != (py::ssize_t)av.size() || in.shape(1) != NDIM)
        throw std::invalid_argument("expected an array of shape (natoms, NDIM)");
    auto q = in.unchecked<2>();
    for (uint i = 0; i < av.size(); i++) {
        Atom &a = av[i];
        Vec &v = which == 0 ? a.x : which == 1 ? a.v : which == 2 ? a.a : a.f;
        for (int k = 0; k < NDIM; k++) v[k] = q(i, k);
    }
}

// Atom state as the reference's pickling stores it (src/sim.i:494-503):
// (tuple(x), tuple(v), tuple(f), tuple(a)).
static py::tuple vec_tuple(const Vec &v) {
    py::tuple t(NDIM);
    for (int k = 0; k < NDIM; k++) t[k] = v[k];
    return t;
}
static Vec tuple_vec(py::handle h) {
    py::sequence q = py::reinterpret_borrow<py::sequence>(h);
    if (py::len(q) != NDIM) throw std::invalid_argument("state vector of the wrong length");
    Vec v;
    for (int k = 0; k < NDIM; k++) v[k] = q[k].cast<double>();
    return v;
}
static py::tuple atom_state(const Atom &a) {
    return py::make_tuple(vec_tuple(a.x), vec_tuple(a.v), vec_tuple(a.f), vec_tuple(a.a));
}
static void set_atom_state(Atom &a, py::handle st) {
    py::sequence q = py::reinterpret_borrow<py::sequence>(st);
    if (py::len(q) != 4) throw std::invalid_argument("an atom state holds x, v, f, a");
    a.x = tuple_vec(q[0]);
    a.v = tuple_vec(q[1]);
    a.f = tuple_vec(q[2]);
    a.a = tuple_vec(q[3]);
}

template <class NL, class A>
static void bind_nlisted(py::module_ &m, const char *name) {
    py::class_<NL, Interaction, sptr<NL> >(m, name)
        .def(py::init<sptr<Box>, sptr<AtomVec>, flt>(), py::arg("box"), py::arg("atoms"),
             py::arg("skin"))
        .def(py::init<sptr<AtomVec>, sptr<NeighborList> >())
        .def("add", &NL::add)
        .def("energy", py::overload_cast<Box &>(&NL::energy))
        .def("contacts", &NL::contacts)
        .def("overlaps", &NL::overlaps)
        .def("pressure", &NL::pressure)
        .def("set_forces", &NL::set_forces)
        .def("set_forces_get_pressure", &NL::set_forces_get_pressure)
        .def("stress", &NL::stress)
        .def("set_forces_get_stress", &NL::set_forces_get_stress)
        .def("neighbor_list", &NL::neighbor_list)
        .def("size", &NL::size)
        .def("getatom", &NL::getatom, py::return_value_policy::reference_internal);
}

#ifdef VEC2D
#define MODNAME _d2
#else
#define MODNAME _d3
#endif

PYBIND11_MODULE(MODNAME, m) {
    m.doc() = "ParM-compatible MD hot path on the MI355X device (NDIM = " +
              std::to_string(NDIM) + ")";
    m.attr("NDIM") = NDIM;
    m.attr("FP64") = (int)PARM_FP64;
    m.attr("MIXED") = (int)PARM_MIXED;
    m.def("seed", &seed);
    m.def("rand01", &rand01);
    m.def("rand_vec", &rand_vec);
    m.def("vec", []() { return Vec::Zero(); });

    py::class_<Box, sptr<Box> >(m, "Box")
        .def("diff", &Box::diff)
        .def("V", &Box::V);
    py::class_<OriginBox, Box, sptr<OriginBox> >(m, "OriginBox")
        .def(py::init<flt>())
        .def(py::init<Vec>())
        .def("diff", &OriginBox::diff)
        .def("V", &OriginBox::V)
        .def("L", &OriginBox::L)
        .def("rand_loc", &OriginBox::rand_loc)
        .def("box_shape", &OriginBox::box_shape);

    py::class_<Atom>(m, "Atom")
        .def(py::init([]() {
            Atom a;
            a.x = a.v = a.a = a.f = Vec::Zero();
            a.m = 0;
            return a;
        }))
        .def_readwrite("x", &Atom::x)
        .def_readwrite("v", &Atom::v)
        .def_readwrite("a", &Atom::a)
        .def_readwrite("f", &Atom::f)
        .def_readwrite("m", &Atom::m)
        .def(py::pickle([](const Atom &a) { return atom_state(a); },
                        [](py::tuple st) {
                            Atom a;
                            a.m = 0;
                            set_atom_state(a, st);
                            return a;
                        }));

    py::class_<AtomID>(m, "AtomID")
        .def("n", &AtomID::n)
        .def("is_null", &AtomID::is_null)
        .def_property(
            "x", [](const AtomID &a) { return a.peek().x; },
            [](AtomID &a, Vec v) { a->x = v; })
        .def_property(
            "v", [](const AtomID &a) { return a.peek().v; },
            [](AtomID &a, Vec v) { a->v = v; })
        .def_property(
            "a", [](const AtomID &a) { return a.peek().a; },
            [](AtomID &a, Vec v) { a->a = v; })
        .def_property(
            "f", [](const AtomID &a) { return a.peek().f; },
            [](AtomID &a, Vec v) { a->f = v; })
        .def_property(
            "m", [](const AtomID &a) { return a.peek().m; },
            [](AtomID &a, flt m) { a->m = m; })
        .def("__getstate__", [](const AtomID &a) { return atom_state(a.peek()); })
        .def("__eq__", [](const AtomID &a, const AtomID &b) { return a == b; })
        .def("__hash__", [](const AtomID &a) { return a.n(); })
        .def("__repr__", [](const AtomID &a) {
            return "<AtomID " + std::to_string(a.n()) + ">";
        });

    py::class_<IDPair>(m, "IDPair")
        .def("first", &IDPair::first)
        .def("last", &IDPair::last)
        .def("__iter__", [](const IDPair &p) {
            return py::iter(py::make_tuple(p.first(), p.last()));
        });

    py::class_<AtomGroup, sptr<AtomGroup> >(m, "AtomGroup")
        .def("size", &AtomGroup::size)
        .def("get_id", &AtomGroup::get_id)
        .def("com", &AtomGroup::com)
        .def("com_velocity", &AtomGroup::com_velocity)
        .def("mass", &AtomGroup::mass)
        .def("momentum", &AtomGroup::momentum)
        .def("kinetic_energy", &AtomGroup::kinetic_energy,
             py::arg("originvelocity") = Vec::Zero())
        .def("add_velocity", &AtomGroup::add_velocity)
        .def("reset_com_velocity", &AtomGroup::reset_com_velocity)
        .def("randomize_velocities", &AtomGroup::randomize_velocities)
        .def("reset_forces", &AtomGroup::reset_forces);

    py::class_<AtomVec, AtomGroup, sptr<AtomVec> >(m, "AtomVec")
        .def(py::init<vector<double> >())
        .def(py::init<uint, flt>())
        .def("get_id", &AtomVec::get_id)
        .def("size", &AtomVec::size)
        .def("N", &AtomVec::size)
        .def("get", [](AtomVec &av, uint i) { return av.get_id(i); })
        .def("getmass", [](AtomVec &av, uint i) { return av.get_id(i).peek().m; })
        .def("__len__", &AtomVec::size)
        .def("__getitem__",
             [](AtomVec &av, int i) {
                 if (i < 0) i += (int)av.size();
                 if (i < 0 || (uint)i >= av.size()) throw py::index_error();
                 return av.get_id((uint)i);
             })
        .def("__iter__",
             [](py::object self) {
                 AtomVec &av = self.cast<AtomVec &>();
                 py::list l;
                 for (uint i = 0; i < av.size(); i++) l.append(av.get_id(i));
                 return py::iter(l);
             })
        .def("kinetic_energy", &AtomVec::kinetic_energy,
             py::arg("originvelocity") = Vec::Zero())
        .def("momentum", &AtomVec::momentum)
        .def("mass", &AtomVec::mass)
        .def("set_device_precision", &AtomVec::set_device_precision)
        .def("get_positions", [](AtomVec &av) { return get_field(av, 0); })
        .def("get_velocities", [](AtomVec &av) { return get_field(av, 1); })
        .def("get_accelerations", [](AtomVec &av) { return get_field(av, 2); })
        .def("get_forces", [](AtomVec &av) { return get_field(av, 3); })
        .def("set_positions", [](AtomVec &av, py::array_t<double> a) { set_field(av, 0, a); })
        .def("set_velocities",
             [](AtomVec &av, py::array_t<double> a) { set_field(av, 1, a); })
        .def("set_accelerations",
             [](AtomVec &av, py::array_t<double> a) { set_field(av, 2, a); })
        .def("set_forces", [](AtomVec &av, py::array_t<double> a) { set_field(av, 3, a); })
        .def("snapshot", &AtomVec::snapshot, py::arg("positions") = true,
             py::arg("velocities") = false)
        .def("write_xyz", &AtomVec::write_xyz, py::arg("ticket"), py::arg("path"),
             py::arg("append") = true, py::arg("comment") = "", py::arg("elements") = "",
             py::arg("precision") = 6)
        // (masses, [atom states]) as the reference pickles an AtomVec
        // (src/sim.i:546-576)
        .def(py::pickle(
            [](AtomVec &av) {
                py::list masses, states;
                for (uint i = 0; i < av.size(); i++) {
                    const Atom &a = av.get_id(i).peek();
                    masses.append(a.m);
                    states.append(atom_state(a));
                }
                return py::make_tuple(masses, states);
            },
            [](py::tuple st) {
                if (py::len(st) != 2) throw std::invalid_argument("AtomVec state");
                vector<double> masses = st[0].cast<vector<double> >();
                py::sequence states = st[1].cast<py::sequence>();
                if (py::len(states) != masses.size())
                    throw std::invalid_argument("AtomVec state: one state per mass");
                auto av = std::make_shared<AtomVec>(masses);
                for (uint i = 0; i < masses.size(); i++) set_atom_state((*av)[i], states[i]);
                return av;
            }));

    py::class_<StateTracker, sptr<StateTracker> >(m, "StateTracker")
        .def("update", &StateTracker::update);
    py::class_<NeighborList, StateTracker, sptr<NeighborList> >(m, "NeighborList")
        .def(py::init<sptr<Box>, sptr<AtomVec>, flt>())
        .def("update", &NeighborList::update)
        .def("update_list", &NeighborList::update_list, py::arg("force") = true)
        .def("which", &NeighborList::which)
        .def("numpairs", &NeighborList::numpairs)
        .def("ignore", &NeighborList::ignore)
        .def("add", &NeighborList::add)
        .def("ignore_size", &NeighborList::ignore_size)
        .def("size", &NeighborList::size)
        .def("get", &NeighborList::get)
        .def("__len__", &NeighborList::numpairs)
        .def("__getitem__",
             [](NeighborList &nl, size_t i) {
                 if (i >= nl.numpairs()) throw py::index_error("Index out of bounds");
                 return nl.get((uint)i);
             })
        .def("__iter__", [](NeighborList &nl) {
            py::list l;
            for (uint i = 0; i < nl.numpairs(); i++) l.append(nl.get(i));
            return py::iter(l);
        });

    py::class_<Interaction, sptr<Interaction> >(m, "Interaction")
        .def("energy", &Interaction::energy)
        .def("set_forces", &Interaction::set_forces)
        .def("set_forces_get_pressure", &Interaction::set_forces_get_pressure)
        .def("pressure", &Interaction::pressure)
        .def("stress", &Interaction::stress);

    py::class_<EpsSigExpAtom, AtomID>(m, "EpsSigExpAtom")
        .def(py::init<AtomID, flt, flt, flt>())
        .def_readwrite("eps", &EpsSigExpAtom::eps)
        .def_readwrite("sigma", &EpsSigExpAtom::sigma)
        .def_readwrite("exponent", &EpsSigExpAtom::exponent)
        .def("max_size", &EpsSigExpAtom::max_size);
    py::class_<IEpsSigCutAtom, AtomID>(m, "IEpsSigCutAtom")
        .def(py::init<AtomID, vector<flt>, uint, flt, flt>())
        .def_readwrite("epsilons", &IEpsSigCutAtom::epsilons)
        .def_readwrite("indx", &IEpsSigCutAtom::indx)
        .def_readwrite("sigma", &IEpsSigCutAtom::sigma)
        .def_readwrite("sigcut", &IEpsSigCutAtom::sigcut)
        .def("max_size", &IEpsSigCutAtom::max_size);

    py::class_<EpsSigAtom, AtomID>(m, "EpsSigAtom")
        .def(py::init<AtomID, flt, flt>())
        .def_readwrite("epsilon", &EpsSigAtom::epsilon)
        .def_readwrite("sigma", &EpsSigAtom::sigma)
        .def("max_size", &EpsSigAtom::max_size);
    py::class_<EpsSigCutAtom, EpsSigAtom>(m, "EpsSigCutAtom")
        .def(py::init<AtomID, flt, flt, flt>())
        .def_readwrite("sigcut", &EpsSigCutAtom::sigcut)
        .def("max_size", &EpsSigCutAtom::max_size);
    py::class_<IEpsISigCutAtom, AtomID>(m, "IEpsISigCutAtom")
        .def(py::init<AtomID, vector<flt>, vector<flt>, uint, flt>())
        .def_readwrite("epsilons", &IEpsISigCutAtom::epsilons)
        .def_readwrite("sigmas", &IEpsISigCutAtom::sigmas)
        .def_readwrite("indx", &IEpsISigCutAtom::indx)
        .def_readwrite("sigcut", &IEpsISigCutAtom::sigcut)
        .def("max_size", &IEpsISigCutAtom::max_size);
    py::class_<IEpsISigExpAtom, AtomID>(m, "IEpsISigExpAtom")
        .def(py::init<AtomID, vector<flt>, vector<flt>, uint, flt>(), py::arg("a"),
             py::arg("epsilons"), py::arg("sigmas"), py::arg("indx"),
             py::arg("exponent") = 2.5)
        .def_readwrite("epsilons", &IEpsISigExpAtom::epsilons)
        .def_readwrite("sigmas", &IEpsISigExpAtom::sigmas)
        .def_readwrite("indx", &IEpsISigExpAtom::indx)
        .def_readwrite("exponent", &IEpsISigExpAtom::exponent)
        .def("max_size", &IEpsISigExpAtom::max_size);

    bind_nlisted<Repulsion, EpsSigExpAtom>(m, "Repulsion");
    bind_nlisted<LJAttractRepulse, IEpsSigCutAtom>(m, "LJAttractRepulse");
    bind_nlisted<LJRepulse, EpsSigAtom>(m, "LJRepulse");
    bind_nlisted<LJIICut, IEpsISigCutAtom>(m, "LJIICut");
    bind_nlisted<RepulsionII, IEpsISigExpAtom>(m, "RepulsionII");

    py::class_<Constraint, sptr<Constraint> >(m, "Constraint");
    py::class_<Collection, sptr<Collection> >(m, "Collection")
        .def("timestep", &Collection::timestep)
        .def("set_forces", &Collection::set_forces, py::arg("constraints_and_a") = true)
        .def("initialize", &Collection::initialize)
        .def("energy", &Collection::energy)
        .def("potential_energy", &Collection::potential_energy)
        .def("kinetic_energy", &Collection::kinetic_energy)
        .def("temp", &Collection::temp, py::arg("minuscomv") = true)
        .def("virial", &Collection::virial)
        .def("pressure", &Collection::pressure)
        .def("degrees_of_freedom", &Collection::degrees_of_freedom)
        .def("com", &Collection::com)
        .def("com_velocity", &Collection::com_velocity)
        .def("reset_com_velocity", &Collection::reset_com_velocity)
        .def("scale_velocities", &Collection::scale_velocities)
        .def("scale_velocities_to_temp", &Collection::scale_velocities_to_temp,
             py::arg("T"), py::arg("minuscomv") = true)
        .def("scale_velocities_to_energy", &Collection::scale_velocities_to_energy)
        .def("add_interaction", &Collection::add_interaction)
        .def("add_tracker", &Collection::add_tracker)
        .def("get_interactions", &Collection::get_interactions);
    py::class_<CollectionVerlet, Collection, sptr<CollectionVerlet> >(m, "CollectionVerlet")
        .def(py::init<sptr<Box>, sptr<AtomGroup>, flt, vector<sptr<Interaction> >,
                      vector<sptr<StateTracker> >, vector<sptr<Constraint> > >(),
             py::arg("box"), py::arg("atoms"), py::arg("dt"),
             py::arg("interactions") = vector<sptr<Interaction> >(),
             py::arg("trackers") = vector<sptr<StateTracker> >(),
             py::arg("constraints") = vector<sptr<Constraint> >())
        .def("timestep", &CollectionVerlet::timestep)
        .def("timesteps", &CollectionVerlet::timesteps)
        .def("set_dt", &CollectionVerlet::set_dt);
    py::class_<CollectionDamped, Collection, sptr<CollectionDamped> >(m, "CollectionDamped")
        .def(py::init<sptr<Box>, sptr<AtomGroup>, flt, flt, vector<sptr<Interaction> >,
                      vector<sptr<StateTracker> >, vector<sptr<Constraint> > >(),
             py::arg("box"), py::arg("atoms"), py::arg("dt"), py::arg("damping"),
             py::arg("interactions") = vector<sptr<Interaction> >(),
             py::arg("trackers") = vector<sptr<StateTracker> >(),
             py::arg("constraints") = vector<sptr<Constraint> >())
        .def("timestep", &CollectionDamped::timestep)
        .def("timesteps", &CollectionDamped::timesteps)
        .def("change_damping", &CollectionDamped::change_damping)
        .def("set_dt", &CollectionDamped::set_dt);
    py::class_<CollectionNLCG, Collection, sptr<CollectionNLCG> >(m, "CollectionNLCG")
        .def(py::init<sptr<OriginBox>, sptr<AtomGroup>, flt, flt, vector<sptr<Interaction> >,
                      vector<sptr<StateTracker> >, vector<sptr<Constraint> >, flt, flt, uint,
                      flt>(),
             py::arg("box"), py::arg("atoms"), py::arg("dt"), py::arg("P0"),
             py::arg("interactions") = vector<sptr<Interaction> >(),
             py::arg("trackers") = vector<sptr<StateTracker> >(),
             py::arg("constraints") = vector<sptr<Constraint> >(), py::arg("kappa") = 10.0,
             py::arg("kmax") = 1000.0, py::arg("secmax") = 40u, py::arg("seceps") = 1e-20)
        .def("timestep", &CollectionNLCG::timestep)
        .def("timesteps", &CollectionNLCG::timesteps)
        .def("descend", &CollectionNLCG::descend)
        .def("reset", &CollectionNLCG::reset)
        .def("set_forces",
             py::overload_cast<bool, bool>(&CollectionNLCG::set_forces),
             py::arg("constraints_and_a") = true, py::arg("setV") = true)
        .def("kinetic_energy", &CollectionNLCG::kinetic_energy)
        .def("pressure", &CollectionNLCG::pressure)
        .def("hamiltonian", &CollectionNLCG::hamiltonian)
        .def("set_dt", &CollectionNLCG::set_dt)
        .def("set_pressure_goal", &CollectionNLCG::set_pressure_goal)
        .def("get_pressure_goal", &CollectionNLCG::get_pressure_goal)
        .def("set_kappa", &CollectionNLCG::set_kappa)
        .def("set_max_alpha", &CollectionNLCG::set_max_alpha)
        .def("set_max_alpha_fraction", &CollectionNLCG::set_max_alpha_fraction)
        .def("set_max_dx", &CollectionNLCG::set_max_dx)
        .def("set_max_step", &CollectionNLCG::set_max_step)
        .def_readwrite("dt", &CollectionNLCG::dt)
        .def_readwrite("seceps", &CollectionNLCG::seceps)
        .def_readwrite("secmax", &CollectionNLCG::secmax)
        .def_readwrite("kappa", &CollectionNLCG::kappa)
        .def_readwrite("alphamax", &CollectionNLCG::alphamax)
        .def_readwrite("afrac", &CollectionNLCG::afrac)
        .def_readwrite("dxmax", &CollectionNLCG::dxmax)
        .def_readwrite("stepmax", &CollectionNLCG::stepmax)
        .def_readwrite("kmax", &CollectionNLCG::kmax)
        .def_readwrite("P0", &CollectionNLCG::P0)
        .def_readwrite("maxdV", &CollectionNLCG::maxdV)
        .def_readonly("Knew", &CollectionNLCG::Knew)
        .def_readonly("k", &CollectionNLCG::k)
        .def_readonly("vl", &CollectionNLCG::vl)
        .def_readonly("fl", &CollectionNLCG::fl)
        .def_readonly("al", &CollectionNLCG::al)
        .def_readonly("alpha", &CollectionNLCG::alpha)
        .def_readonly("beta", &CollectionNLCG::beta)
        .def_readonly("betaused", &CollectionNLCG::betaused)
        .def_readonly("dxsum", &CollectionNLCG::dxsum)
        .def_readonly("alphavmax", &CollectionNLCG::alphavmax)
        .def_readonly("sec", &CollectionNLCG::sec);
    py::class_<CollectionNLCGFixedL, CollectionNLCG, sptr<CollectionNLCGFixedL> >(
        m, "CollectionNLCGFixedL")
        .def(py::init<sptr<OriginBox>, sptr<AtomGroup>, flt, flt, vector<sptr<Interaction> >,
                      vector<sptr<StateTracker> >, vector<sptr<Constraint> >, flt, flt, uint,
                      flt>(),
             py::arg("box"), py::arg("atoms"), py::arg("dt"), py::arg("P0"),
             py::arg("interactions") = vector<sptr<Interaction> >(),
             py::arg("trackers") = vector<sptr<StateTracker> >(),
             py::arg("constraints") = vector<sptr<Constraint> >(), py::arg("kappa") = 10.0,
             py::arg("kmax") = 1000.0, py::arg("secmax") = 40u, py::arg("seceps") = 1e-20);
}
