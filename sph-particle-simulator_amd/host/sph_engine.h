// sph_engine.h -- SPH engine with the reference's class API, stepping on the GPU.
//
// Public surface of the reference's src/sph_engine.h:14-204 (SPHParameters, SPHEngine,
// utils::), implemented over the C ABI of libsphk (include/sphk.h).  The particle state
// stays in GPU memory; the host ParticleSystem returned by get_particles() is a mirror that
// is refreshed from the device on demand.  Differences from the reference:
//  - PerformanceStats is public (the reference's bindings name the private nested type).
//  - SPHParameters has one extra field, `formulation` (0 = reference formulas, default).
//  - step() enqueues the step on the engine's stream and returns; every read-back waits.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "particle.h"

struct sphk_engine;

namespace sph {

// Same fields and defaults as the reference (src/sph_engine.h:14-34)
struct SPHParameters {
    float rest_density = 1000.0f;
    float gas_constant = 2000.0f;
    float viscosity = 0.001f;
    float smoothing_length = 0.02f;
    float particle_mass = 0.001f;
    float timestep = 0.001f;
    float gravity = -9.81f;
    float damping = 0.99f;
    float CFL_factor = 0.4f;

    struct Boundaries {
        float xmin = -1.0f, xmax = 1.0f;
        float ymin = -1.0f, ymax = 1.0f;
        float zmin = -1.0f, zmax = 1.0f;
    } bounds;

    float neighbor_search_radius = 0.04f;

    // 0: the reference's formulas (kernel 1/(pi h^3), a = F/m).
    // 1: consistent SPH (kernel 3/(2 pi h^3), a = F/rho_i); same passes otherwise.
    int formulation = 0;
};

// Thrown when the GPU library reports an error the reference would not have (no device,
// a domain beyond the dense cell grid, HIP runtime failures).
class SPHError : public std::runtime_error {
   public:
    SPHError(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

class SPHEngine {
   public:
    // Phase timings (s) and neighbour statistics (src/sph_engine.h:51-59)
    struct PerformanceStats {
        double total_time = 0.0;
        double neighbor_search_time = 0.0;
        double density_computation_time = 0.0;
        double force_computation_time = 0.0;
        double integration_time = 0.0;
        size_t max_neighbors = 0;
        size_t total_neighbor_queries = 0;
    };

    explicit SPHEngine(size_t max_particles = 1000000);
    ~SPHEngine();
    SPHEngine(const SPHEngine&) = delete;
    SPHEngine& operator=(const SPHEngine&) = delete;

    // Initialization (src/sph_engine.cpp:20-81)
    void initialize(const SPHParameters& params);
    void initialize_dam_break();
    void initialize_fluid_drop();
    void initialize_granular_flow();

    // Particle management (src/sph_engine.cpp:83-91)
    void add_particles(const std::vector<Particle>& particles);
    void clear_particles();

    // Simulation control (src/sph_engine.cpp:93-150); dt <= 0 selects the CFL timestep
    void step(float dt = 0.0f);
    void run_steps(size_t num_steps, bool adaptive_timestep = true);

    // Accessors
    const ParticleSystem& get_particles() const;
    const SPHParameters& get_parameters() const { return params_; }
    float get_current_time() const;
    size_t get_step_count() const { return step_count_; }

    // Parameter modification (src/sph_engine.cpp:152-172)
    void set_parameters(const SPHParameters& params);
    void set_gravity(float gravity);
    void set_viscosity(float viscosity);
    void set_smoothing_length(float h);
    void set_boundaries(float xmin, float xmax, float ymin, float ymax, float zmin, float zmax);

    // Performance monitoring
    const PerformanceStats& get_performance_stats() const;
    void reset_performance_stats();

    // Utility functions (src/sph_engine.cpp:178-200)
    void compute_conservation_errors(float& mass_error, float& energy_error) const;
    float get_total_mass() const;
    float get_total_energy() const;

    // Data export (copies in particle order; densities/pressures are capacity-length,
    // src/sph_engine.h:133-136)
    std::vector<vec3> get_positions() const;
    std::vector<vec3> get_velocities() const;
    std::vector<float> get_densities() const;
    std::vector<float> get_pressures() const;

    void validate_simulation() const;
    bool is_initialized() const { return initialized_; }

    // Additions of this build (no reference counterpart)
    sphk_engine* device_engine() const { return engine_; }
    size_t capacity() const { return capacity_; }
    void synchronize() const;

   private:
    void push_params();
    void upload_from(size_t first);
    void replace_particles(const std::vector<Particle>& ps);
    void refresh_mirror() const;
    void check(int rc) const;

    sphk_engine* engine_ = nullptr;
    size_t capacity_;
    SPHParameters params_;
    size_t step_count_ = 0;
    bool initialized_ = false;
    mutable ParticleSystem particles_;  // host mirror
    mutable bool mirror_stale_ = false;
    mutable PerformanceStats perf_stats_;
};

enum class IntegrationMethod { EULER, VERLET, LEAPFROG };

namespace utils {

// Scene builders (src/sph_engine.cpp:450-552)
std::vector<Particle> create_dam_break_setup(const vec3& dam_size, const vec3& fluid_size,
                                             float spacing, const SPHParameters& params);
std::vector<Particle> create_fluid_drop_setup(const vec3& center, float radius, float spacing,
                                              const SPHParameters& params);
std::vector<Particle> create_granular_flow_setup(const vec3& pile_size, const vec3& domain_size,
                                                 float spacing, const SPHParameters& params);

// Validation (src/sph_engine.cpp:554-578)
bool validate_particle_setup(const std::vector<Particle>& particles);
float compute_average_neighbors(const SPHEngine& engine);

}  // namespace utils

}  // namespace sph
