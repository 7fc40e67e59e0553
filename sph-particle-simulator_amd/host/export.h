// export.h -- CSV time series of a run, the reference's on-disk record of a simulation
// (examples/dam_break.cpp:97-101,161-166).
#pragma once

#include <fstream>
#include <string>

#include "sph_engine.h"

namespace sph {
namespace utils {

// Columns: time,particles,mass_error,energy,total_energy,avg_density,max_velocity, where
// `energy` holds the energy error and avg_density averages the capacity-length density
// vector, both as in the reference example.
class CsvTimeSeries {
   public:
    explicit CsvTimeSeries(const std::string& path);
    bool is_open() const { return out_.is_open(); }
    // Appends one row from the engine's current state (reads the state back from the GPU).
    void record(const SPHEngine& engine);

   private:
    std::ofstream out_;
};

}  // namespace utils
}  // namespace sph
