// sph_engine.cpp -- SPHEngine over the C ABI of libsphk.
#include "sph_engine.h"

#include <iostream>
#include <numeric>

#include "../../include/sphk.h"

namespace sph {

namespace {

sphk_params to_c(const SPHParameters& p) {
    sphk_params c{};
    c.rest_density = p.rest_density;
    c.gas_constant = p.gas_constant;
    c.viscosity = p.viscosity;
    c.smoothing_length = p.smoothing_length;
    c.particle_mass = p.particle_mass;
    c.timestep = p.timestep;
    c.gravity = p.gravity;
    c.damping = p.damping;
    c.CFL_factor = p.CFL_factor;
    c.xmin = p.bounds.xmin;
    c.xmax = p.bounds.xmax;
    c.ymin = p.bounds.ymin;
    c.ymax = p.bounds.ymax;
    c.zmin = p.bounds.zmin;
    c.zmax = p.bounds.zmax;
    c.neighbor_search_radius = p.neighbor_search_radius;
    c.formulation = p.formulation;
    return c;
}

std::string last_error() {
    char buf[1024];
    sphk_last_error(buf, sizeof(buf));
    return std::string(buf);
}

}  // namespace

void SPHEngine::check(int rc) const {
    if (rc != SPHK_OK) throw SPHError(rc, "sphk: " + last_error());
}

// src/sph_engine.cpp:13-18
SPHEngine::SPHEngine(size_t max_particles) : capacity_(max_particles), particles_(max_particles) {
    check(sphk_create(max_particles, &engine_));
    push_params();
}

SPHEngine::~SPHEngine() { sphk_destroy(engine_); }

void SPHEngine::push_params() {
    sphk_params c = to_c(params_);
    check(sphk_set_params(engine_, &c));
}

// src/sph_engine.cpp:20-33: the hash cell is neighbor_search_radius (not recomputed from h)
void SPHEngine::initialize(const SPHParameters& params) {
    params_ = params;
    push_params();
    initialized_ = true;
    reset_performance_stats();
}

// Uploads mirror particles [first, size) to the device at the same offset; the device
// count becomes size (0 when the mirror is empty).
void SPHEngine::upload_from(size_t first) {
    const size_t n = particles_.size();
    const size_t m = n > first ? n - first : 0;
    std::vector<float> pos(3 * m + 3), vel(3 * m + 3), mass(m + 1);
    for (size_t k = 0; k < m; ++k) {
        const Particle& p = particles_[first + k];
        pos[3 * k] = p.position.x;
        pos[3 * k + 1] = p.position.y;
        pos[3 * k + 2] = p.position.z;
        vel[3 * k] = p.velocity.x;
        vel[3 * k + 1] = p.velocity.y;
        vel[3 * k + 2] = p.velocity.z;
        mass[k] = p.mass;
    }
    check(sphk_upload(engine_, first, m, pos.data(), vel.data(), mass.data()));
    mirror_stale_ = false;
}

void SPHEngine::replace_particles(const std::vector<Particle>& ps) {
    particles_.clear();
    particles_.add_particles(ps);
    upload_from(0);
}

// src/sph_engine.cpp:35-49
void SPHEngine::initialize_dam_break() {
    if (!initialized_) initialize(SPHParameters{});
    replace_particles(utils::create_dam_break_setup(vec3(0.4f, 0.6f, 0.8f), vec3(0.2f, 0.4f, 0.8f),
                                                    0.01f, params_));
}

// src/sph_engine.cpp:51-65
void SPHEngine::initialize_fluid_drop() {
    if (!initialized_) initialize(SPHParameters{});
    replace_particles(utils::create_fluid_drop_setup(vec3(0.0f, 0.5f, 0.0f), 0.1f, 0.008f, params_));
}

// src/sph_engine.cpp:67-81
void SPHEngine::initialize_granular_flow() {
    if (!initialized_) initialize(SPHParameters{});
    replace_particles(utils::create_granular_flow_setup(vec3(0.3f, 0.4f, 0.8f),
                                                        vec3(1.0f, 1.0f, 1.0f), 0.012f, params_));
}

// src/sph_engine.cpp:83-85 -> ParticleSystem::add_particles: appends within capacity
void SPHEngine::add_particles(const std::vector<Particle>& particles) {
    refresh_mirror();
    const size_t first = particles_.size();
    particles_.add_particles(particles);
    if (particles_.size() > first) upload_from(first);
}

// src/sph_engine.cpp:87-91
void SPHEngine::clear_particles() {
    particles_.clear();
    check(sphk_clear(engine_));
    step_count_ = 0;
    mirror_stale_ = false;
}

// src/sph_engine.cpp:93-144: silent no-op when uninitialised or empty
void SPHEngine::step(float dt) {
    if (!initialized_ || particles_.size() == 0) return;
    check(sphk_step(engine_, dt));
    step_count_++;
    mirror_stale_ = true;
}

// src/sph_engine.cpp:146-150
void SPHEngine::run_steps(size_t num_steps, bool adaptive_timestep) {
    for (size_t i = 0; i < num_steps; ++i) step(adaptive_timestep ? 0.0f : params_.timestep);
}

void SPHEngine::refresh_mirror() const {
    if (!mirror_stale_) return;
    const size_t n = particles_.size();
    std::vector<float> pos(3 * n), vel(3 * n);
    std::vector<float> dens(capacity_), prs(capacity_);
    check(sphk_download(engine_, SPHK_POSITIONS, pos.data(), pos.size()));
    check(sphk_download(engine_, SPHK_VELOCITIES, vel.data(), vel.size()));
    check(sphk_download(engine_, SPHK_DENSITIES, dens.data(), dens.size()));
    check(sphk_download(engine_, SPHK_PRESSURES, prs.data(), prs.size()));
    for (size_t i = 0; i < n; ++i) {
        Particle& p = particles_[i];
        p.position = vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        p.velocity = vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        p.density = dens[i];
        p.pressure = prs[i];
    }
    mirror_stale_ = false;
}

const ParticleSystem& SPHEngine::get_particles() const {
    refresh_mirror();
    return particles_;
}

float SPHEngine::get_current_time() const {
    float t = 0.0f;
    check(sphk_get_time(engine_, &t));
    return t;
}

// src/sph_engine.cpp:152-156: set_parameters re-derives the search radius from h
void SPHEngine::set_parameters(const SPHParameters& params) {
    params_ = params;
    set_smoothing_length(params.smoothing_length);
}

void SPHEngine::set_gravity(float gravity) {
    params_.gravity = gravity;
    push_params();
}

void SPHEngine::set_viscosity(float viscosity) {
    params_.viscosity = viscosity;
    push_params();
}

// src/sph_engine.cpp:158-163
void SPHEngine::set_smoothing_length(float h) {
    params_.smoothing_length = h;
    params_.neighbor_search_radius = 2.0f * h;
    push_params();
}

// src/sph_engine.cpp:165-172
void SPHEngine::set_boundaries(float xmin, float xmax, float ymin, float ymax, float zmin,
                               float zmax) {
    params_.bounds.xmin = xmin;
    params_.bounds.xmax = xmax;
    params_.bounds.ymin = ymin;
    params_.bounds.ymax = ymax;
    params_.bounds.zmin = zmin;
    params_.bounds.zmax = zmax;
    push_params();
}

const SPHEngine::PerformanceStats& SPHEngine::get_performance_stats() const {
    sphk_stats s{};
    check(sphk_get_stats(engine_, &s));
    perf_stats_.total_time = s.total_time;
    perf_stats_.neighbor_search_time = s.neighbor_search_time;
    perf_stats_.density_computation_time = s.density_computation_time;
    perf_stats_.force_computation_time = s.force_computation_time;
    perf_stats_.integration_time = s.integration_time;
    perf_stats_.max_neighbors = (size_t)s.max_neighbors;
    perf_stats_.total_neighbor_queries = (size_t)s.total_neighbor_queries;
    return perf_stats_;
}

void SPHEngine::reset_performance_stats() {
    check(sphk_reset_stats(engine_));
    check(sphk_set_timing(engine_, 1));
    perf_stats_ = PerformanceStats{};
}

void SPHEngine::compute_conservation_errors(float& mass_error, float& energy_error) const {
    check(sphk_conservation_errors(engine_, &mass_error, &energy_error));
}

float SPHEngine::get_total_mass() const {
    float m = 0.0f;
    check(sphk_total_mass(engine_, &m));
    return m;
}

float SPHEngine::get_total_energy() const {
    float e = 0.0f;
    check(sphk_total_energy(engine_, &e));
    return e;
}

std::vector<vec3> SPHEngine::get_positions() const {
    const size_t n = particles_.size();
    std::vector<vec3> out(n);
    check(sphk_download(engine_, SPHK_POSITIONS, reinterpret_cast<float*>(out.data()), 3 * n));
    return out;
}

std::vector<vec3> SPHEngine::get_velocities() const {
    const size_t n = particles_.size();
    std::vector<vec3> out(n);
    check(sphk_download(engine_, SPHK_VELOCITIES, reinterpret_cast<float*>(out.data()), 3 * n));
    return out;
}

std::vector<float> SPHEngine::get_densities() const {
    std::vector<float> out(capacity_);
    check(sphk_download(engine_, SPHK_DENSITIES, out.data(), out.size()));
    return out;
}

std::vector<float> SPHEngine::get_pressures() const {
    std::vector<float> out(capacity_);
    check(sphk_download(engine_, SPHK_PRESSURES, out.data(), out.size()));
    return out;
}

// src/sph_engine.cpp:355-365
void SPHEngine::validate_simulation() const {
    refresh_mirror();
    if (!utils::validate_particle_setup(std::vector<Particle>(particles_.begin(), particles_.end())))
        std::cerr << "Warning: Particle setup validation failed\n";
    float avg = utils::compute_average_neighbors(*this);
    if (avg < 20.0f || avg > 80.0f)
        std::cerr << "Warning: Average neighbor count (" << avg
                  << ") may affect simulation stability\n";
}

void SPHEngine::synchronize() const { check(sphk_synchronize(engine_)); }

namespace utils {

// src/sph_engine.cpp:450-487
std::vector<Particle> create_dam_break_setup(const vec3& dam_size, const vec3& fluid_size,
                                             float spacing, const SPHParameters& params) {
    std::vector<Particle> particles;
    std::vector<Particle> boundaries =
        create_boundary_box(vec3(0.0f, dam_size.y / 2.0f, 0.0f), dam_size, spacing,
                            params.particle_mass);
    for (auto& p : boundaries) {
        p.type = ParticleType::BOUNDARY;
        p.velocity = vec3(0.0f);
        p.color = vec3(0.5f, 0.5f, 0.5f);
    }
    std::vector<Particle> fluid = create_fluid_block(
        vec3(-dam_size.x / 2.0f + fluid_size.x / 2.0f, fluid_size.y / 2.0f, 0.0f), fluid_size,
        spacing, params.particle_mass);
    for (auto& p : fluid) {
        p.type = ParticleType::FLUID;
        p.color = vec3(0.0f, 0.5f, 1.0f);
    }
    particles.insert(particles.end(), boundaries.begin(), boundaries.end());
    particles.insert(particles.end(), fluid.begin(), fluid.end());
    return particles;
}

// src/sph_engine.cpp:489-514
std::vector<Particle> create_fluid_drop_setup(const vec3& center, float radius, float spacing,
                                              const SPHParameters& params) {
    std::vector<Particle> particles;
    const int n = static_cast<int>(2.0f * radius / spacing);
    const vec3 start = center - vec3(radius);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            for (int k = 0; k < n; ++k) {
                const vec3 position =
                    start + vec3(static_cast<float>(i) * spacing, static_cast<float>(j) * spacing,
                                 static_cast<float>(k) * spacing);
                const vec3 to_center = position - center;
                if (dot(to_center, to_center) <= radius * radius) {
                    Particle p(position, params.particle_mass, ParticleType::FLUID);
                    p.color = vec3(0.0f, 0.7f, 1.0f);
                    particles.push_back(p);
                }
            }
        }
    }
    return particles;
}

// src/sph_engine.cpp:516-552
std::vector<Particle> create_granular_flow_setup(const vec3& pile_size, const vec3& domain_size,
                                                 float spacing, const SPHParameters& params) {
    std::vector<Particle> particles;
    std::vector<Particle> boundaries = create_boundary_box(
        vec3(0.0f, domain_size.y / 2.0f, 0.0f), domain_size, spacing, params.particle_mass);
    for (auto& p : boundaries) {
        p.type = ParticleType::BOUNDARY;
        p.color = vec3(0.4f, 0.4f, 0.4f);
    }
    std::vector<Particle> granular = create_fluid_block(
        vec3(0.0f, pile_size.y / 2.0f + spacing, 0.0f), pile_size, spacing, params.particle_mass);
    for (auto& p : granular) {
        p.type = ParticleType::SOLID;
        p.color = vec3(0.8f, 0.6f, 0.2f);
    }
    particles.insert(particles.end(), boundaries.begin(), boundaries.end());
    particles.insert(particles.end(), granular.begin(), granular.end());
    return particles;
}

// O(N^2) overlap scan (src/sph_engine.cpp:554-569)
bool validate_particle_setup(const std::vector<Particle>& particles) {
    if (particles.empty()) return false;
    for (size_t i = 0; i < particles.size(); ++i) {
        for (size_t j = i + 1; j < particles.size(); ++j) {
            const vec3 diff = particles[i].position - particles[j].position;
            if (dot(diff, diff) < 1e-8f) {
                std::cerr << "Warning: Overlapping particles detected\n";
                return false;
            }
        }
    }
    return true;
}

// Fixed estimate, as the reference returns (src/sph_engine.cpp:571-578).  The measured mean
// neighbour count is available from sphk_get_stats (total_neighbors / total_neighbor_queries).
float compute_average_neighbors(const SPHEngine& engine) {
    if (engine.get_particles().size() == 0) return 0.0f;
    return 50.0f;
}

}  // namespace utils

}  // namespace sph
