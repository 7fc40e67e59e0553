// export.cpp -- CSV time series (examples/dam_break.cpp:97-101,161-166).
#include "export.h"

#include <algorithm>

namespace sph {
namespace utils {

CsvTimeSeries::CsvTimeSeries(const std::string& path) : out_(path) {
    if (out_.is_open())
        out_ << "time,particles,mass_error,energy,total_energy,avg_density,max_velocity\n";
}

void CsvTimeSeries::record(const SPHEngine& engine) {
    if (!out_.is_open()) return;
    float mass_error = 0.0f, energy_error = 0.0f;
    engine.compute_conservation_errors(mass_error, energy_error);
    float max_velocity = 0.0f;
    for (const vec3& v : engine.get_velocities()) max_velocity = std::max(max_velocity, length(v));
    const std::vector<float> densities = engine.get_densities();
    float avg_density = 0.0f;
    if (!densities.empty()) {
        for (float d : densities) avg_density += d;
        avg_density /= static_cast<float>(densities.size());
    }
    const double sim_time = engine.get_current_time();
    out_ << sim_time << "," << engine.get_particles().size() << "," << mass_error << ","
         << energy_error << "," << engine.get_total_energy() << "," << avg_density << ","
         << max_velocity << "\n";
}

}  // namespace utils
}  // namespace sph
