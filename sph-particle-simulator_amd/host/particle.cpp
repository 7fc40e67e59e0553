// particle.cpp -- host particle container and scene builders.
// Behaviour follows the reference's src/particle.cpp (line citations per function).
#include "particle.h"

#include <algorithm>
#include <functional>
#include <iostream>
#include <random>

namespace sph {

ParticleSystem::ParticleSystem(size_t capacity) : capacity_(capacity) {
    particles_.reserve(capacity_);
}

void ParticleSystem::reserve(size_t capacity) {
    capacity_ = capacity;
    particles_.reserve(capacity_);
}

void ParticleSystem::resize(size_t size) { particles_.resize(size); }

void ParticleSystem::clear() { particles_.clear(); }

// Drops particles beyond capacity with a warning; id = index (src/particle.cpp:26-33)
void ParticleSystem::add_particle(const Particle& particle) {
    if (particles_.size() < capacity_) {
        particles_.push_back(particle);
        particles_.back().id = static_cast<int>(particles_.size() - 1);
    } else {
        std::cerr << "Warning: Particle capacity exceeded\n";
    }
}

void ParticleSystem::add_particles(const std::vector<Particle>& new_particles) {
    for (const auto& p : new_particles) add_particle(p);
}

// Removal semantics of src/particle.cpp:41-66.  remove_particle renumbers the particles from
// `index` on to their new positions; remove_particles applies its indices largest first, each
// to the container as it stands (one at or past the current end is skipped, a repeated one
// removes the particle that has moved into its place), then renumbers every particle.
void ParticleSystem::remove_particle(size_t index) {
    if (index >= particles_.size()) return;
    std::move(particles_.begin() + static_cast<std::ptrdiff_t>(index) + 1, particles_.end(),
              particles_.begin() + static_cast<std::ptrdiff_t>(index));
    particles_.pop_back();
    for (size_t i = index; i < particles_.size(); ++i) particles_[i].id = static_cast<int>(i);
}

void ParticleSystem::remove_particles(const std::vector<size_t>& indices) {
    std::vector<size_t> todo(indices);
    std::sort(todo.begin(), todo.end(), std::greater<size_t>());
    // Only positions at or after an index have moved when it is applied (the larger ones went
    // first), so position i then holds the first particle at original index >= i not yet gone.
    std::vector<char> gone(particles_.size(), 0);
    size_t live = particles_.size();
    for (size_t i : todo) {
        if (i >= live) continue;
        size_t k = i;
        while (gone[k]) ++k;
        gone[k] = 1;
        --live;
    }
    size_t w = 0;
    for (size_t k = 0; k < particles_.size(); ++k) {
        if (gone[k]) continue;
        if (w != k) particles_[w] = std::move(particles_[k]);
        particles_[w].id = static_cast<int>(w);
        ++w;
    }
    particles_.resize(w);
}

std::vector<vec3> ParticleSystem::get_positions() const {
    std::vector<vec3> out;
    out.reserve(particles_.size());
    for (const auto& p : particles_) out.push_back(p.position);
    return out;
}

std::vector<vec3> ParticleSystem::get_velocities() const {
    std::vector<vec3> out;
    out.reserve(particles_.size());
    for (const auto& p : particles_) out.push_back(p.velocity);
    return out;
}

std::vector<float> ParticleSystem::get_densities() const {
    std::vector<float> out;
    out.reserve(particles_.size());
    for (const auto& p : particles_) out.push_back(p.density);
    return out;
}

std::vector<float> ParticleSystem::get_pressures() const {
    std::vector<float> out;
    out.reserve(particles_.size());
    for (const auto& p : particles_) out.push_back(p.pressure);
    return out;
}

namespace {
// one field of every particle set to `v` (the set_* calls of src/particle.cpp:104-120)
void assign_field(std::vector<Particle>& ps, float Particle::*field, float v) {
    std::for_each(ps.begin(), ps.end(), [field, v](Particle& p) { p.*field = v; });
}
}  // namespace

void ParticleSystem::set_mass(float mass) { assign_field(particles_, &Particle::mass, mass); }

void ParticleSystem::set_viscosity(float viscosity) {
    assign_field(particles_, &Particle::viscosity, viscosity);
}

void ParticleSystem::set_temperature(float temperature) {
    assign_field(particles_, &Particle::temperature, temperature);
}

// Host-side wall clamp of the mirror (src/particle.cpp:122-153).  The engine's own clamp
// runs on the GPU inside each step.
void ParticleSystem::apply_boundary_conditions(float xmin, float xmax, float ymin, float ymax,
                                               float zmin, float zmax) {
    const float damping = 0.8f;
    const float lo[3] = {xmin, ymin, zmin};
    const float hi[3] = {xmax, ymax, zmax};
    for (auto& p : particles_) {
        for (int d = 0; d < 3; ++d) {
            if (p.position[d] < lo[d]) {
                p.position[d] = lo[d];
                p.velocity[d] *= -damping;
            } else if (p.position[d] > hi[d]) {
                p.position[d] = hi[d];
                p.velocity[d] *= -damping;
            }
        }
    }
}

// A point uniform in the box [xmin, xmax) x [ymin, ymax) x [zmin, zmax) (src/particle.cpp:156-165):
// one generator per thread, seeded once from the system's entropy source; coordinates drawn
// x, then y, then z.
vec3 generate_random_position(float xmin, float xmax, float ymin, float ymax, float zmin,
                              float zmax) {
    thread_local std::mt19937 engine{std::random_device{}()};
    auto draw = [](float lo, float hi) {
        return std::uniform_real_distribution<float>(lo, hi)(engine);
    };
    const float x = draw(xmin, xmax);
    const float y = draw(ymin, ymax);
    const float z = draw(zmin, zmax);
    return vec3(x, y, z);
}

// Lattice block, x-major loop order i, j, k (src/particle.cpp:166-188)
std::vector<Particle> create_fluid_block(const vec3& center, const vec3& size, float spacing,
                                         float mass) {
    std::vector<Particle> particles;
    const int nx = static_cast<int>(size.x / spacing);
    const int ny = static_cast<int>(size.y / spacing);
    const int nz = static_cast<int>(size.z / spacing);
    const vec3 start = center - size * 0.5f;
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            for (int k = 0; k < nz; ++k) {
                const vec3 position =
                    start + vec3(static_cast<float>(i) * spacing, static_cast<float>(j) * spacing,
                                 static_cast<float>(k) * spacing);
                Particle p(position, mass, ParticleType::FLUID);
                p.color = vec3(0.0f, 0.5f + 0.5f * (position.y - start.y) / size.y, 1.0f);
                particles.push_back(p);
            }
        }
    }
    return particles;
}

// Six lattice faces: bottom/top (z), left/right (x), front/back (y) (src/particle.cpp:190-229)
std::vector<Particle> create_boundary_box(const vec3& center, const vec3& size, float spacing,
                                          float mass) {
    std::vector<Particle> particles;
    const vec3 half = size * 0.5f;
    const vec3 start = center - half;
    auto add_face = [&](int a1, int a2, int a3, float value) {
        const int n1 = static_cast<int>(size[a1] / spacing);
        const int n2 = static_cast<int>(size[a2] / spacing);
        for (int i = 0; i < n1; ++i) {
            for (int j = 0; j < n2; ++j) {
                vec3 position = start;
                position[a1] = start[a1] + static_cast<float>(i) * spacing;
                position[a2] = start[a2] + static_cast<float>(j) * spacing;
                position[a3] = value;
                Particle p(position, mass, ParticleType::BOUNDARY);
                p.color = vec3(0.5f, 0.5f, 0.5f);
                particles.push_back(p);
            }
        }
    };
    add_face(0, 1, 2, start.z);
    add_face(0, 1, 2, start.z + size.z);
    add_face(1, 2, 0, start.x);
    add_face(1, 2, 0, start.x + size.x);
    add_face(0, 2, 1, start.y);
    add_face(0, 2, 1, start.y + size.y);
    return particles;
}

}  // namespace sph
