// particle.h -- particle record, host particle container and scene builders.
//
// Same types, fields, defaults and functions as the reference's src/particle.h:9-105.
// On this build the particle state lives in GPU memory while the engine steps; a
// ParticleSystem held by SPHEngine is the host mirror that get_particles() refreshes.
#pragma once

#include <vector>

#include "vec3.h"

namespace sph {

// Particle types for multi-phase simulations (src/particle.h:10-14)
enum class ParticleType { FLUID, BOUNDARY, SOLID };

// Particle record, 76 bytes like the reference's AoS Particle (src/particle.h:17-49)
struct Particle {
    vec3 position;
    vec3 velocity;
    vec3 acceleration;
    float density;
    float pressure;
    float mass;
    ParticleType type;
    float temperature;
    float viscosity;
    vec3 color;
    int id;

    Particle()
        : position(0.0f), velocity(0.0f), acceleration(0.0f), density(0.0f), pressure(0.0f),
          mass(1.0f), type(ParticleType::FLUID), temperature(293.15f), viscosity(0.001f),
          color(0.0f, 0.5f, 1.0f), id(-1) {}

    Particle(const vec3& pos, float mass_ = 1.0f, ParticleType type_ = ParticleType::FLUID)
        : position(pos), velocity(0.0f), acceleration(0.0f), density(0.0f), pressure(0.0f),
          mass(mass_), type(type_), temperature(293.15f), viscosity(0.001f),
          color(0.0f, 0.5f, 1.0f), id(-1) {}
};

// Host container with a capacity limit (src/particle.h:52-100, src/particle.cpp:9-153)
class ParticleSystem {
   private:
    std::vector<Particle> particles_;
    size_t capacity_;

   public:
    explicit ParticleSystem(size_t capacity = 1000000);

    void reserve(size_t capacity);
    void resize(size_t size);
    void clear();

    Particle& operator[](size_t index) { return particles_[index]; }
    const Particle& operator[](size_t index) const { return particles_[index]; }
    size_t size() const { return particles_.size(); }
    size_t capacity() const { return capacity_; }
    bool empty() const { return particles_.empty(); }

    auto begin() { return particles_.begin(); }
    auto end() { return particles_.end(); }
    auto begin() const { return particles_.cbegin(); }
    auto end() const { return particles_.cend(); }

    void add_particle(const Particle& particle);
    void add_particles(const std::vector<Particle>& new_particles);

    void remove_particle(size_t index);
    void remove_particles(const std::vector<size_t>& indices);

    std::vector<vec3> get_positions() const;
    std::vector<vec3> get_velocities() const;
    std::vector<float> get_densities() const;
    std::vector<float> get_pressures() const;

    void set_mass(float mass);
    void set_viscosity(float viscosity);
    void set_temperature(float temperature);

    void apply_boundary_conditions(float xmin, float xmax, float ymin, float ymax, float zmin,
                                   float zmax);
};

vec3 generate_random_position(float xmin, float xmax, float ymin, float ymax, float zmin,
                              float zmax);
std::vector<Particle> create_fluid_block(const vec3& center, const vec3& size, float spacing,
                                         float mass = 1.0f);
std::vector<Particle> create_boundary_box(const vec3& center, const vec3& size, float spacing,
                                          float mass = 1.0f);

}  // namespace sph
