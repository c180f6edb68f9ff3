// model.h -- icy::Model (simulation/model.h:32-57): owns the parameters and
// the GPU_Implementation5, and advances the simulation one frame
// (UpdateEveryNthStep substeps) per Step().
#pragma once

#include <fstream>
#include <memory>
#include <mutex>
#include <string>

#include "gpu_implementation5.h"
#include "parameters_sim.h"

namespace icy { class Model; }

class icy::Model {
public:
    Model();
    ~Model();
    void Reset();     // model.cpp:124-132

    void Prepare();   // model.cpp:134-139, invoked once, at simulation start
    bool Step();      // model.cpp:27-113
    void RequestAbort() { abortRequested = true; }

    void UnlockCycleMutex();   // model.cpp:116-120

    icy::SimParams prms;
    GPU_Implementation5 gpu;
    int max_points_transferred = 0, max_pt_deviation = 0;
    bool SyncTopologyRequired = true;

    std::mutex processing_current_cycle_data;  // locked until the host has taken the frame's results
    std::mutex accessing_point_data;

    // How a frame's substeps run.  false (default): one device-side call per
    // frame (clock, cadence and grid buffers kept on the device, replayed from
    // a hipGraph).  true: the reference's host loop, four calls per substep.
    bool hostDrivenSubsteps = false;
    // Re-bin and re-sort the points on the device at the start of each frame
    // (keeps the tiled kernels' point runs local; order within a cell stays
    // stable).  The reference sorts only on load and rebalance.
    bool resortEveryFrame = true;
    // Per-kernel device timings (hipEvents around each launch; disables the
    // graph replay).
    bool recordKernelTimings = false;
    // Directory of timings.csv and indenter_force.csv (model.cpp:8-11);
    // empty = no logs.  Opened on the first Step().
    std::string logDirectory = "logs";
    bool verbose = true;

private:
    bool abortRequested = false;
    std::unique_ptr<std::ofstream> log_timing, log_indenter_force;
    void open_logs();
};
