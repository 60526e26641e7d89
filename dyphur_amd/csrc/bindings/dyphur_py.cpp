// dyphur_py — the reference's Python module (bindings/python/dyphur_py.cpp:29-163)
// over the gfx950 C++ layer: same class names, methods and argument names
// (Device, Stream, AABB, ShapeType, ShapeParams, ShapeStore, BodyParams,
// BodyStore, Broadphase, Narrowphase, ContactSensor), with zero-copy numpy
// views of the sensor arrays.  Additions: Device.default_gpu / preferred,
// IntegratorParams + integrate, JointType / JointParams / JointStore,
// XpbdSolver (contacts and joints), step (one frame of substeps), state_hash,
// BodyStore.state (host copy of the body arrays).  This build has no CPU
// device: Device.default_cpu raises.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <compute/device.hpp>
#include <core/aabb.hpp>
#include <core/body_store.hpp>
#include <core/broadphase.hpp>
#include <core/contact_sensor.hpp>
#include <core/integrator.hpp>
#include <core/joint_store.hpp>
#include <core/narrowphase.hpp>
#include <core/shape_store.hpp>
#include <core/step.hpp>
#include <core/xpbd_solver.hpp>

#include <string>
#include <vector>

namespace py = pybind11;
using namespace pybind11::literals;
using namespace dyphur;

namespace {

// Zero-copy numpy view of a host array owned by `owner`.
template <typename T>
py::array_t<T> view_of(const T *ptr, size_t n, py::handle owner) {
    return py::array_t<T>({static_cast<py::ssize_t>(n)}, {static_cast<py::ssize_t>(sizeof(T))}, ptr, owner);
}

template <typename T>
py::array_t<T> download(Stream &s, const T *d, size_t n) {
    py::array_t<T> out(static_cast<py::ssize_t>(n));
    if (n) s.queue().memcpy(out.mutable_data(), d, n * sizeof(T)).wait();
    return out;
}

}  // namespace

PYBIND11_MODULE(dyphur_py, m) {
    m.doc() = "dyphur rigid-body frame step on gfx950 (Python bindings)";

    py::register_exception<Error>(m, "Error", PyExc_RuntimeError);

    py::class_<Device>(m, "Device")
        .def_static("default_cpu", &Device::default_cpu)
        .def_static("default_gpu", &Device::default_gpu)
        .def_static("preferred", &Device::preferred)
        .def("name", &Device::name)
        .def("is_gpu", &Device::is_gpu)
        .def("is_cpu", &Device::is_cpu)
        .def("max_compute_units", &Device::max_compute_units)
        .def("global_mem_bytes", &Device::global_mem_bytes)
        .def("make_stream", &Device::make_stream);

    py::class_<Stream>(m, "Stream").def("wait", &Stream::wait);

    py::class_<AABB>(m, "AABB")
        .def(py::init([](float a, float b, float c, float d, float e, float f) { return AABB{a, b, c, d, e, f}; }),
             "min_x"_a, "min_y"_a, "min_z"_a, "max_x"_a, "max_y"_a, "max_z"_a)
        .def_readwrite("min_x", &AABB::min_x)
        .def_readwrite("min_y", &AABB::min_y)
        .def_readwrite("min_z", &AABB::min_z)
        .def_readwrite("max_x", &AABB::max_x)
        .def_readwrite("max_y", &AABB::max_y)
        .def_readwrite("max_z", &AABB::max_z)
        .def("overlaps", &AABB::overlaps);

    py::enum_<ShapeType>(m, "ShapeType")
        .value("Box", ShapeType::Box)
        .value("Sphere", ShapeType::Sphere)
        .export_values();

    py::class_<ShapeParams>(m, "ShapeParams")
        .def(py::init<>())
        .def_readwrite("type", &ShapeParams::type)
        .def_readwrite("half_x", &ShapeParams::half_x)
        .def_readwrite("half_y", &ShapeParams::half_y)
        .def_readwrite("half_z", &ShapeParams::half_z);

    py::class_<ShapeView>(m, "ShapeView").def_readonly("n", &ShapeView::n);

    py::class_<ShapeStore>(m, "ShapeStore")
        .def(py::init<Stream &, uint32_t>(), "s"_a, "capacity"_a, py::keep_alive<1, 2>())
        .def("add", &ShapeStore::add)
        .def("upload", &ShapeStore::upload)
        .def("count", &ShapeStore::count)
        .def("view", &ShapeStore::view);

    py::class_<BodyParams>(m, "BodyParams")
        .def(py::init<>())
        .def("set_position", [](BodyParams &p, float x, float y, float z) { p.position = Vec3f{x, y, z}; },
             "x"_a, "y"_a, "z"_a)
        .def("set_velocity", [](BodyParams &p, float vx, float vy, float vz) { p.linear_velocity = Vec3f{vx, vy, vz}; },
             "vx"_a, "vy"_a, "vz"_a)
        .def("set_angular_velocity",
             [](BodyParams &p, float wx, float wy, float wz) { p.angular_velocity = Vec3f{wx, wy, wz}; }, "wx"_a,
             "wy"_a, "wz"_a)
        .def("set_rotation", [](BodyParams &p, float w, float x, float y, float z) { p.rotation = Quatf{w, x, y, z}; },
             "w"_a, "x"_a, "y"_a, "z"_a)
        .def("set_inertia", [](BodyParams &p, const std::vector<float> &m9) {
                 if (m9.size() != 9) throw py::value_error("inertia: 9 row-major values");
                 for (size_t i = 0; i < 9; ++i) p.inertia.data_[i] = m9[i];
             }, "m"_a)
        .def_readwrite("mass", &BodyParams::mass)
        .def_readwrite("shape_handle", &BodyParams::shape_handle)
        .def_readwrite("flags", &BodyParams::flags);

    py::class_<BodyView>(m, "BodyView").def_readonly("n", &BodyView::n);

    py::class_<BodyStore>(m, "BodyStore")
        .def(py::init<Stream &, uint32_t>(), "s"_a, "capacity"_a, py::keep_alive<1, 2>())
        .def("add", &BodyStore::add)
        .def("upload", &BodyStore::upload)
        .def("count", &BodyStore::count)
        .def("view", &BodyStore::view)
        .def("body_shapes", &BodyStore::body_shapes)
        .def("state", [](BodyStore &b, Stream &s) {
                 BodyView v = b.view();
                 py::dict d;
                 const char *names[13] = {"pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z",
                                          "vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z"};
                 float *ptrs[13] = {v.pos_x, v.pos_y, v.pos_z, v.rot_w, v.rot_x, v.rot_y, v.rot_z,
                                    v.vel_x, v.vel_y, v.vel_z, v.ang_x, v.ang_y, v.ang_z};
                 for (int k = 0; k < 13; ++k) d[names[k]] = download(s, ptrs[k], v.n);
                 return d;
             }, "s"_a, "Host copy of the body state arrays (blocking).");

    py::class_<Broadphase>(m, "Broadphase")
        .def(py::init<Stream &, uint32_t, uint32_t>(), "s"_a, "max_bodies"_a, "max_pairs"_a)
        .def("build_and_query", &Broadphase::build_and_query, "s"_a, "bodies"_a, "shapes"_a, "scene_bounds"_a)
        .def("sort_pairs", &Broadphase::sort_pairs, "s"_a, "n_pairs"_a)
        .def("download_count", &Broadphase::download_count)
        .def("max_pairs", &Broadphase::max_pairs)
        .def("pairs", [](Broadphase &bp, Stream &s, uint32_t n) {
                 const uint32_t k = n < bp.max_pairs() ? n : bp.max_pairs();
                 auto flat = download(s, reinterpret_cast<const uint32_t *>(bp.pairs_ptr()), 2 * static_cast<size_t>(k));
                 return flat.reshape({static_cast<py::ssize_t>(k), static_cast<py::ssize_t>(2)});
             }, "s"_a, "n_pairs"_a, "The first n_pairs (a, b) pairs as an (n, 2) uint32 array (blocking).");

    py::class_<Narrowphase>(m, "Narrowphase")
        .def(py::init<Stream &, uint32_t>(), "s"_a, "max_contacts"_a)
        .def("run", [](Narrowphase &np, Stream &s, const Broadphase &bp, uint32_t n_pairs, BodyView bv,
                       ShapeView sv) { np.run(s, bp.pairs_ptr(), n_pairs, bv, sv); },
             "s"_a, "bp"_a, "n_pairs"_a, "bodies"_a, "shapes"_a)
        .def("download_count", &Narrowphase::download_count)
        .def("last_contact_count", &Narrowphase::last_contact_count)
        .def("overflowed", &Narrowphase::overflowed);

    py::class_<ContactSensor>(m, "ContactSensor")
        .def(py::init<uint32_t>(), "shape_idx"_a, "Subscribe to contacts involving the given shape index.")
        .def("query", &ContactSensor::query, "s"_a, "np"_a, "body_shapes"_a,
             "Download and filter contacts from the narrowphase. Returns contact count.")
        .def("count", &ContactSensor::count)
        .def("shape_idx", &ContactSensor::shape_idx)
        .def("body_a", [](py::object self) {
                 const auto &cs = self.cast<const ContactSensor &>();
                 return view_of(cs.body_a(), cs.count(), self);
             })
        .def("body_b", [](py::object self) {
                 const auto &cs = self.cast<const ContactSensor &>();
                 return view_of(cs.body_b(), cs.count(), self);
             })
#define DY_SENSOR_FIELD(f)                                          \
        .def(#f, [](py::object self) {                              \
            const auto &cs = self.cast<const ContactSensor &>();    \
            return view_of(cs.f(), cs.count(), self);               \
        })
        DY_SENSOR_FIELD(pos_x) DY_SENSOR_FIELD(pos_y) DY_SENSOR_FIELD(pos_z)
        DY_SENSOR_FIELD(norm_x) DY_SENSOR_FIELD(norm_y) DY_SENSOR_FIELD(norm_z)
        DY_SENSOR_FIELD(depth);
#undef DY_SENSOR_FIELD

    py::class_<IntegratorParams>(m, "IntegratorParams")
        .def(py::init<>())
        .def("set_gravity", [](IntegratorParams &p, float x, float y, float z) { p.gravity = Vec3f{x, y, z}; },
             "x"_a, "y"_a, "z"_a)
        .def_readwrite("dt", &IntegratorParams::dt)
        .def_readwrite("n_substeps", &IntegratorParams::n_substeps);

    m.def("integrate", &integrate, "s"_a, "bodies"_a, "params"_a);

    py::enum_<JointType>(m, "JointType")
        .value("Fixed", JointType::Fixed)
        .value("Revolute", JointType::Revolute)
        .value("Prismatic", JointType::Prismatic)
        .value("Ball", JointType::Ball);

    py::class_<JointParams>(m, "JointParams")
        .def(py::init<>())
        .def_readwrite("body_parent", &JointParams::body_parent)
        .def_readwrite("body_child", &JointParams::body_child)
        .def("set_anchor_parent", [](JointParams &p, float x, float y, float z) { p.anchor_parent = Vec3f{x, y, z}; },
             "x"_a, "y"_a, "z"_a)
        .def("set_anchor_child", [](JointParams &p, float x, float y, float z) { p.anchor_child = Vec3f{x, y, z}; },
             "x"_a, "y"_a, "z"_a)
        .def("set_axis", [](JointParams &p, float x, float y, float z) { p.axis = Vec3f{x, y, z}; }, "x"_a, "y"_a,
             "z"_a)
        .def_readwrite("limit_lo", &JointParams::limit_lo)
        .def_readwrite("limit_hi", &JointParams::limit_hi)
        .def_readwrite("target_pos", &JointParams::target_pos)
        .def_readwrite("target_vel", &JointParams::target_vel)
        .def_readwrite("stiffness", &JointParams::stiffness)
        .def_readwrite("damping", &JointParams::damping)
        .def_readwrite("compliance_pos", &JointParams::compliance_pos)
        .def_readwrite("compliance_ang", &JointParams::compliance_ang)
        .def_readwrite("type", &JointParams::type);

    py::class_<JointStore>(m, "JointStore")
        .def(py::init<Stream &, uint32_t>(), "s"_a, "capacity"_a)
        .def("add", &JointStore::add, "p"_a)
        .def("upload", &JointStore::upload)
        .def("set_targets", [](JointStore &js, const std::vector<float> &pos, const std::vector<float> &vel) {
                 if (pos.size() != js.count() || vel.size() != js.count())
                     throw py::value_error("set_targets: one target per joint");
                 js.set_targets(pos.data(), vel.data());
             }, "pos"_a, "vel"_a)
        .def("count", &JointStore::count)
        .def("capacity", &JointStore::capacity);

    py::class_<XpbdSolver>(m, "XpbdSolver")
        .def(py::init<Stream &, int, uint32_t, float>(), "s"_a, "n_iters"_a = 10, "max_contacts"_a = 16384,
             "friction"_a = 0.5f)
        .def("solve", [](XpbdSolver &x, Stream &s, Narrowphase &np, BodyView bv, float dt, JointStore *joints) {
                 x.solve(s, np.contacts(), joints ? joints->view() : JointView{}, bv, dt);
             }, "s"_a, "np"_a, "bodies"_a, "dt"_a, "joints"_a = nullptr,
             "Solve the narrowphase's contacts and the joints of a JointStore (None: contacts only).")
        .def("set_force_serial", &XpbdSolver::set_force_serial)
        .def("last_stats", [](const XpbdSolver &x) {
                 const dy_solve_stats st = x.last_stats();
                 py::dict d;
                 d["n_constraints"] = st.n_constraints;
                 d["used_serial"] = st.used_serial;
                 d["n_classes"] = st.n_classes;
                 d["rounds"] = st.rounds;
                 d["overflow"] = st.overflow;
                 return d;
             })
        .def_property_readonly_static("kSerialThreshold", [](py::object) { return XpbdSolver::kSerialThreshold; });

    m.def("step", &step, "s"_a, "bodies"_a, "shapes"_a, "bp"_a, "np"_a, "solver"_a, "bounds"_a, "params"_a,
          "One frame: params.n_substeps x (integrate, broadphase, narrowphase, solve) at dt / n_substeps.");
    m.def("state_hash", &state_hash, "s"_a, "bodies"_a, "FNV-1a-64 of pos/rot of all bodies (blocking).");
}
