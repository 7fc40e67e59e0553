// sphk_engine.h -- host-side engine state of libsphk, shared by the C ABI sources.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "sphk_internal.h"

namespace sphk {

// Records the message returned by sphk_last_error (per thread) and returns code.
int set_error(int code, const char* fmt, ...);
const std::string& last_error_text();
void set_last_error_text(const std::string& s);

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return ::sphk::set_error(SPHK_E_HIP, "%s failed: %s", #expr,                    \
                                     hipGetErrorString(_e));                               \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

struct SurrogateState;  // learned macro-step model (sphk_surrogate_api.hip)
void surrogate_free(SurrogateState* s);

}  // namespace sphk

using sphk::Grid;
using sphk::SceneGrid;
using sphk::StepConsts;
using sphk::WalkBuffers;

// Host view of one scene: particle range, wall bounds, the box its finite positions may
// occupy, and its dense grid.
struct SceneHost {
    uint64_t begin = 0, end = 0;
    float bounds[6] = {-1.0f, 1.0f, -1.0f, 1.0f, -1.0f, 1.0f};  // xmin xmax ymin ymax zmin zmax
    float ext_lo[3] = {0, 0, 0}, ext_hi[3] = {0, 0, 0};
    bool ext_any = false;
    Grid grid;
};

// Everything a captured step depends on besides buffer contents.
struct GraphKey {
    StepConsts consts;
    uint32_t n;
    int32_t shape_density, shape_force;
    int32_t adaptive;
    float dt;
    uint64_t buffer_gen;
};

struct sphk_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t cap = 0;
    size_t n = 0;
    sphk_params params{};
    uint64_t step_count = 0;

    // home arrays (caller order)
    float4* pm = nullptr;
    float4* vel = nullptr;
    float4* acc = nullptr;
    float* rho = nullptr;
    float* prs = nullptr;
    // snapshot
    float4* snap_pm = nullptr;
    float4* snap_vel = nullptr;
    float4* snap_acc = nullptr;
    size_t snap_n = 0;
    uint64_t snap_steps = 0;
    std::vector<SceneHost> snap_scenes;
    bool snap_ext_unknown = false;
    bool has_snapshot = false;
    // work arrays
    uint32_t* keys = nullptr;
    uint32_t* keys_alt = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* vals_alt = nullptr;
    uint32_t* hist = nullptr;
    uint32_t* scan_scratch = nullptr;
    size_t scan_scratch_cap = 0;
    uint32_t* lookback = nullptr;  // state of the radix histograms' one-launch scan (zeroed)
    size_t lookback_cap = 0;
    WalkBuffers walk;
    int shape_density = 0;  // walk launch shapes (walk_shape_name), set by sphk_set_walk_shapes
    int shape_force = 0;
    bool shapes_explicit = false;  // false: ensure_grid picks the shapes by particle count
    // sorted copies, capacity + 1 rows (the last is the padding slot StepConsts::pad_slot):
    // spm (x, y, z, m), svr (vx, vy, vz, rho), and the force walk's neighbour rows written by
    // the density walk
    float4* spm = nullptr;
    float4* svr = nullptr;
    sphk::ForceRow* frows = nullptr;
    uint2* sq = nullptr;  // quantized coordinates of the candidate filter (x|y 16 bits, z 32)
    uint32_t* cell_start = nullptr;
    uint64_t cell_start_cap = 0;
    uint32_t* skey = nullptr;  // -> keys or keys_alt after the sort
    uint32_t* sid = nullptr;   // -> vals or vals_alt after the sort
    // device scalars: [0] dt, [1] time, [2] snapshot time
    float* fscal = nullptr;
    // [0] out-of-grid flag, [1] cfl scratch
    uint32_t* uscal = nullptr;
    // [0] max neighbours, [1] total neighbours
    unsigned long long* nstats = nullptr;

    // scenes and their grids
    std::vector<SceneHost> scenes;  // one implicit scene [0, n) unless set by sphk_set_scenes
    bool scenes_explicit = false;
    bool ext_unknown = false;  // positions may lie outside the scene extents (device bbox needed)
    bool grid_valid = false;
    std::vector<SceneGrid> table;  // host copy of the device scene table
    SceneGrid* d_scenes = nullptr;
    size_t d_scenes_cap = 0;
    // what d_scenes holds: the table last uploaded and the buffer generation it went to (a
    // restore that leaves the grids as they were re-uploads nothing)
    std::vector<SceneGrid> table_uploaded;
    uint64_t table_uploaded_gen = ~0ull;
    StepConsts consts{};
    bool sorted_valid = false;  // work arrays hold the cell-sorted image of the current state
    uint64_t buffer_gen = 0;    // bumped when a buffer used by the step is reallocated

    // captured step
    bool use_graph = false;
    bool graph_valid = false;
    GraphKey graph_key{};
    hipGraphExec_t graph_exec = nullptr;

    // timing: one event set per timed step, drained when stats are read
    struct EventSet {
        hipEvent_t ev[4];
    };
    bool timing = false;
    std::vector<EventSet> pending, free_sets;
    sphk_stats stats{};

    // learned macro-step model, created on first use
    sphk::SurrogateState* sur = nullptr;
};


namespace sphk {

int check_engine(const sphk_engine* e);
int set_device(const sphk_engine* e);
// Validates the scene table and (re)builds the cell grids when the step needs it.
int prepare_grid(sphk_engine* e);
// keys -> sort -> cell_start -> gather -> chunk list for the current state.
int build_sorted(sphk_engine* e, const StepTail* tail = nullptr);
// One plain step enqueued on the engine stream (ev: optional 4 timing events).
int enqueue_step(sphk_engine* e, float dt, bool adaptive, hipEvent_t* ev);
// Host bookkeeping after a step that moved particles (scene extents, grid validity).
void after_positions_moved(sphk_engine* e);

}  // namespace sphk
