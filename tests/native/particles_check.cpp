// Host particle container (host/particle.cpp) under AddressSanitizer and
// UndefinedBehaviorSanitizer, built by tests/native/Makefile.  Removal is checked against the
// reference's semantics restated as a plain loop over a std::vector (src/particle.cpp:41-66:
// remove_particle renumbers from the index on; remove_particles erases largest index first,
// skips indices at or past the current end, then renumbers everything), on random index
// lists with repeats and out-of-range entries.
#include <algorithm>
#include <cstdio>
#include <functional>
#include <random>
#include <vector>

#include "particle.h"

using sph::Particle;
using sph::ParticleSystem;

static bool same(const ParticleSystem& ps, const std::vector<Particle>& want) {
    if (ps.size() != want.size()) return false;
    for (size_t i = 0; i < want.size(); ++i)
        if (ps[i].id != want[i].id || ps[i].mass != want[i].mass) return false;
    return true;
}

static void fill(ParticleSystem& ps, std::vector<Particle>& model, size_t n) {
    ps.clear();
    model.clear();
    for (size_t i = 0; i < n; ++i) {
        Particle p;
        p.mass = (float)i;  // identifies the particle
        ps.add_particle(p);
        p.id = (int)i;
        model.push_back(p);
    }
}

int main() {
    std::mt19937 rng(7);
    int failures = 0;
    for (int trial = 0; trial < 400; ++trial) {
        const size_t n = rng() % 40;
        ParticleSystem ps(1000);
        std::vector<Particle> model;
        fill(ps, model, n);
        // user-set ids before the removal point are left alone by remove_particle
        for (size_t i = 0; i < n; ++i) {
            ps[i].id = 1000 + (int)i;
            model[i].id = 1000 + (int)i;
        }
        const size_t one = rng() % (n + 3);
        ps.remove_particle(one);
        if (one < model.size()) {
            model.erase(model.begin() + (std::ptrdiff_t)one);
            for (size_t i = one; i < model.size(); ++i) model[i].id = (int)i;
        }
        if (!same(ps, model)) ++failures;

        fill(ps, model, n);
        std::vector<size_t> idx(rng() % 12);
        for (auto& v : idx) v = rng() % (n + 4);
        if (!idx.empty() && rng() % 2) idx.push_back(idx[0]);  // a repeat
        ps.remove_particles(idx);
        std::vector<size_t> s(idx);
        std::sort(s.begin(), s.end(), std::greater<size_t>());
        for (size_t v : s)
            if (v < model.size()) model.erase(model.begin() + (std::ptrdiff_t)v);
        for (size_t i = 0; i < model.size(); ++i) model[i].id = (int)i;
        if (!same(ps, model)) ++failures;
    }
    ParticleSystem ps(10);
    std::vector<Particle> model;
    fill(ps, model, 5);
    ps.set_mass(2.0f);
    ps.set_viscosity(3.0f);
    ps.set_temperature(4.0f);
    for (const Particle& p : ps)
        if (p.mass != 2.0f || p.viscosity != 3.0f || p.temperature != 4.0f) ++failures;
    for (int i = 0; i < 1000; ++i) {
        const sph::vec3 v = sph::generate_random_position(-1.0f, 2.0f, 0.5f, 0.75f, 3.0f, 3.5f);
        if (!(v.x >= -1.0f && v.x < 2.0f && v.y >= 0.5f && v.y < 0.75f && v.z >= 3.0f && v.z < 3.5f))
            ++failures;
    }
    if (failures) {
        std::fprintf(stderr, "%d particle-container checks failed\n", failures);
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
