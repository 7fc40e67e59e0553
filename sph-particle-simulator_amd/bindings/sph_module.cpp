// sph_module.cpp -- pybind11 module `sph`, the Python surface of the reference's
// python/bindings.cpp:8-166 over this build's SPHEngine.
//
// Kept: module name, class and function names, argument names and defaults, numpy float32
// copies from the getters, __version__ = "1.0.0".  Added: vec3 <-> 3-sequence conversion
// (the reference declares Particle.position etc. as glm::vec3 without a converter), the
// remaining SPHParameters fields (bounds, CFL_factor, neighbor_search_radius, formulation),
// and a public PerformanceStats.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "export.h"
#include "kernels.h"
#include "sph_engine.h"

namespace py = pybind11;

namespace pybind11 {
namespace detail {
template <>
struct type_caster<sph::vec3> {
    PYBIND11_TYPE_CASTER(sph::vec3, _("numpy.ndarray[float32[3]]"));
    bool load(handle src, bool) {
        if (!src) return false;
        try {
            auto arr = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(src);
            if (!arr || arr.size() != 3) return false;
            const float* d = arr.data();
            value = sph::vec3(d[0], d[1], d[2]);
            return true;
        } catch (...) {
            return false;
        }
    }
    static handle cast(const sph::vec3& v, return_value_policy, handle) {
        py::array_t<float> a(3);
        float* d = a.mutable_data();
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        return a.release();
    }
};
}  // namespace detail
}  // namespace pybind11

static py::array_t<float> vec3s_to_numpy(const std::vector<sph::vec3>& v) {
    py::array_t<float> a({v.size(), (size_t)3});
    if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(sph::vec3));
    return a;
}

static py::array_t<float> floats_to_numpy(const std::vector<float>& v) {
    py::array_t<float> a(v.size());
    if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(float));
    return a;
}

PYBIND11_MODULE(sph, m) {
    m.doc() = "SPH Particle Simulator - Python bindings (MI355X)";

    py::register_exception<sph::SPHError>(m, "SPHError", PyExc_RuntimeError);

    py::class_<sph::SPHParameters::Boundaries>(m, "Boundaries")
        .def(py::init<>())
        .def_readwrite("xmin", &sph::SPHParameters::Boundaries::xmin)
        .def_readwrite("xmax", &sph::SPHParameters::Boundaries::xmax)
        .def_readwrite("ymin", &sph::SPHParameters::Boundaries::ymin)
        .def_readwrite("ymax", &sph::SPHParameters::Boundaries::ymax)
        .def_readwrite("zmin", &sph::SPHParameters::Boundaries::zmin)
        .def_readwrite("zmax", &sph::SPHParameters::Boundaries::zmax);

    py::class_<sph::SPHParameters>(m, "SPHParameters")
        .def(py::init<>())
        .def_readwrite("rest_density", &sph::SPHParameters::rest_density)
        .def_readwrite("gas_constant", &sph::SPHParameters::gas_constant)
        .def_readwrite("viscosity", &sph::SPHParameters::viscosity)
        .def_readwrite("smoothing_length", &sph::SPHParameters::smoothing_length)
        .def_readwrite("particle_mass", &sph::SPHParameters::particle_mass)
        .def_readwrite("timestep", &sph::SPHParameters::timestep)
        .def_readwrite("gravity", &sph::SPHParameters::gravity)
        .def_readwrite("damping", &sph::SPHParameters::damping)
        .def_readwrite("CFL_factor", &sph::SPHParameters::CFL_factor)
        .def_readwrite("bounds", &sph::SPHParameters::bounds)
        .def_readwrite("neighbor_search_radius", &sph::SPHParameters::neighbor_search_radius)
        .def_readwrite("formulation", &sph::SPHParameters::formulation);

    py::enum_<sph::ParticleType>(m, "ParticleType")
        .value("FLUID", sph::ParticleType::FLUID)
        .value("BOUNDARY", sph::ParticleType::BOUNDARY)
        .value("SOLID", sph::ParticleType::SOLID);

    py::class_<sph::Particle>(m, "Particle")
        .def(py::init<>())
        .def(py::init<const sph::vec3&, float, sph::ParticleType>(), py::arg("position"),
             py::arg("mass") = 1.0f, py::arg("type") = sph::ParticleType::FLUID)
        .def_readwrite("position", &sph::Particle::position)
        .def_readwrite("velocity", &sph::Particle::velocity)
        .def_readwrite("acceleration", &sph::Particle::acceleration)
        .def_readwrite("density", &sph::Particle::density)
        .def_readwrite("pressure", &sph::Particle::pressure)
        .def_readwrite("mass", &sph::Particle::mass)
        .def_readwrite("type", &sph::Particle::type)
        .def_readwrite("temperature", &sph::Particle::temperature)
        .def_readwrite("viscosity", &sph::Particle::viscosity)
        .def_readwrite("color", &sph::Particle::color)
        .def_readonly("id", &sph::Particle::id);

    py::class_<sph::ParticleSystem>(m, "ParticleSystem")
        .def("size", &sph::ParticleSystem::size)
        .def("capacity", &sph::ParticleSystem::capacity)
        .def("empty", &sph::ParticleSystem::empty)
        .def("__len__", &sph::ParticleSystem::size)
        .def("__getitem__",
             [](const sph::ParticleSystem& ps, size_t i) {
                 if (i >= ps.size()) throw py::index_error();
                 return ps[i];
             })
        .def("get_positions", [](const sph::ParticleSystem& ps) { return vec3s_to_numpy(ps.get_positions()); })
        .def("get_velocities", [](const sph::ParticleSystem& ps) { return vec3s_to_numpy(ps.get_velocities()); })
        .def("get_densities", [](const sph::ParticleSystem& ps) { return floats_to_numpy(ps.get_densities()); })
        .def("get_pressures", [](const sph::ParticleSystem& ps) { return floats_to_numpy(ps.get_pressures()); });

    py::class_<sph::SPHEngine::PerformanceStats>(m, "PerformanceStats")
        .def_readonly("total_time", &sph::SPHEngine::PerformanceStats::total_time)
        .def_readonly("neighbor_search_time", &sph::SPHEngine::PerformanceStats::neighbor_search_time)
        .def_readonly("density_computation_time", &sph::SPHEngine::PerformanceStats::density_computation_time)
        .def_readonly("force_computation_time", &sph::SPHEngine::PerformanceStats::force_computation_time)
        .def_readonly("integration_time", &sph::SPHEngine::PerformanceStats::integration_time)
        .def_readonly("max_neighbors", &sph::SPHEngine::PerformanceStats::max_neighbors)
        .def_readonly("total_neighbor_queries", &sph::SPHEngine::PerformanceStats::total_neighbor_queries);

    py::class_<sph::SPHEngine>(m, "Simulator")
        .def(py::init<size_t>(), py::arg("max_particles") = 1000000)
        .def("initialize", &sph::SPHEngine::initialize, py::arg("params"))
        .def("initialize_dam_break", &sph::SPHEngine::initialize_dam_break)
        .def("initialize_fluid_drop", &sph::SPHEngine::initialize_fluid_drop)
        .def("initialize_granular_flow", &sph::SPHEngine::initialize_granular_flow)
        .def("add_particles", &sph::SPHEngine::add_particles)
        .def("clear_particles", &sph::SPHEngine::clear_particles)
        .def("step", &sph::SPHEngine::step, py::arg("dt") = 0.0f)
        .def("run_steps", &sph::SPHEngine::run_steps, py::arg("num_steps"),
             py::arg("adaptive_timestep") = true)
        .def("get_particles", &sph::SPHEngine::get_particles, py::return_value_policy::reference_internal)
        .def("get_parameters", &sph::SPHEngine::get_parameters)
        .def("get_current_time", &sph::SPHEngine::get_current_time)
        .def("get_step_count", &sph::SPHEngine::get_step_count)
        .def("set_parameters", &sph::SPHEngine::set_parameters)
        .def("set_gravity", &sph::SPHEngine::set_gravity)
        .def("set_viscosity", &sph::SPHEngine::set_viscosity)
        .def("set_smoothing_length", &sph::SPHEngine::set_smoothing_length)
        .def("set_boundaries", &sph::SPHEngine::set_boundaries)
        .def("get_performance_stats", &sph::SPHEngine::get_performance_stats,
             py::return_value_policy::copy)
        .def("reset_performance_stats", &sph::SPHEngine::reset_performance_stats)
        .def("compute_conservation_errors",
             [](sph::SPHEngine& e) {
                 float me, ee;
                 e.compute_conservation_errors(me, ee);
                 return py::make_tuple(me, ee);
             })
        .def("get_total_mass", &sph::SPHEngine::get_total_mass)
        .def("get_total_energy", &sph::SPHEngine::get_total_energy)
        .def("validate_simulation", &sph::SPHEngine::validate_simulation)
        .def("is_initialized", &sph::SPHEngine::is_initialized)
        .def("get_positions", [](const sph::SPHEngine& e) { return vec3s_to_numpy(e.get_positions()); })
        .def("get_velocities", [](const sph::SPHEngine& e) { return vec3s_to_numpy(e.get_velocities()); })
        .def("get_densities", [](const sph::SPHEngine& e) { return floats_to_numpy(e.get_densities()); })
        .def("get_pressures", [](const sph::SPHEngine& e) { return floats_to_numpy(e.get_pressures()); })
        .def("synchronize", &sph::SPHEngine::synchronize)
        .def("capacity", &sph::SPHEngine::capacity);

    m.def(
        "create_fluid_block",
        [](const py::array_t<float>& center, const py::array_t<float>& size, float spacing,
           float mass) {
            if (center.size() != 3 || size.size() != 3)
                throw std::invalid_argument("center and size must be 3D vectors");
            sph::vec3 c(center.at(0), center.at(1), center.at(2));
            sph::vec3 s(size.at(0), size.at(1), size.at(2));
            return sph::create_fluid_block(c, s, spacing, mass);
        },
        py::arg("center"), py::arg("size"), py::arg("spacing"), py::arg("mass") = 1.0f);

    m.def(
        "create_boundary_box",
        [](const py::array_t<float>& center, const py::array_t<float>& size, float spacing,
           float mass) {
            if (center.size() != 3 || size.size() != 3)
                throw std::invalid_argument("center and size must be 3D vectors");
            sph::vec3 c(center.at(0), center.at(1), center.at(2));
            sph::vec3 s(size.at(0), size.at(1), size.at(2));
            return sph::create_boundary_box(c, s, spacing, mass);
        },
        py::arg("center"), py::arg("size"), py::arg("spacing"), py::arg("mass") = 1.0f);

    // Additions of this build (no binding in the reference module): the kernel classes of
    // src/kernels.h and the CSV time series of examples/dam_break.cpp.
    py::enum_<sph::KernelType>(m, "KernelType")
        .value("CUBIC_SPLINE", sph::KernelType::CUBIC_SPLINE)
        .value("WENDLAND_C2", sph::KernelType::WENDLAND_C2)
        .value("GAUSSIAN", sph::KernelType::GAUSSIAN);
    py::class_<sph::Kernel>(m, "Kernel")
        .def("W", &sph::Kernel::W)
        .def("gradW", &sph::Kernel::gradW)
        .def("laplacianW", &sph::Kernel::laplacianW)
        .def("get_smoothing_length", &sph::Kernel::get_smoothing_length)
        .def("set_smoothing_length", &sph::Kernel::set_smoothing_length)
        .def("get_support_radius", &sph::Kernel::get_support_radius);
    m.def("create_kernel", &sph::create_kernel, py::arg("type"), py::arg("smoothing_length"));
    m.def(
        "cubic_W_2D",
        [](float h, float x, float y) { return sph::CubicSplineKernel(h).W_2D(sph::vec2(x, y)); });
    m.def("cubic_W_1D", [](float h, float r) { return sph::CubicSplineKernel(h).W_1D(r); });
    m.def("estimate_smoothing_length", &sph::kernel_utils::estimate_smoothing_length,
          py::arg("particle_spacing"), py::arg("neighbor_count") = 50.0f);
    py::class_<sph::utils::CsvTimeSeries>(m, "CsvTimeSeries")
        .def(py::init<const std::string&>())
        .def("is_open", &sph::utils::CsvTimeSeries::is_open)
        .def("record", &sph::utils::CsvTimeSeries::record);

    m.attr("__version__") = "1.0.0";
}
