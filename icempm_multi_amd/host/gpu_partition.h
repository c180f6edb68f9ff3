// gpu_partition.h -- GPU_Partition (simulation/gpu_partition.h:84-165): one
// grid slab and its points on one device, over an icempm_partition handle.
// Every failing C-ABI call throws std::runtime_error, as the reference does.
#pragma once

#include <cstdint>
#include <vector>

#include "host_side_soa.h"
#include "parameters_sim.h"

class GPU_Partition {
public:
    GPU_Partition();
    ~GPU_Partition();
    GPU_Partition(GPU_Partition &&o) noexcept;
    GPU_Partition &operator=(GPU_Partition &&o) noexcept;
    GPU_Partition(const GPU_Partition &) = delete;
    GPU_Partition &operator=(const GPU_Partition &) = delete;

    // preparation
    void initialize(int device, int partition);                 // gpu_partition.cu:113-134
    void allocate(int n_points_capacity, int grid_x_capacity);  // gpu_partition.cu:137-197
    void transfer_points_from_soa_to_device(HostSideSOA &hssoa, int point_idx_offset);  // :49-68
    void update_constants();                                    // :200-208
    void transfer_from_device(HostSideSOA &hssoa, int point_idx_offset);                // :16-46

    // calculation
    void reset_grid();                         // :221-228
    void reset_indenter_force_accumulator();   // :212-218
    void p2g();                                // :230-242
    void receive_halos();                      // :296-318 (one slab: nothing to receive)
    void update_nodes(double simulation_time); // :244-257
    void g2p(const bool recordPQ, const bool enablePointTransfer);  // :259-293
    void receive_points(int nFromLeft, int nFromRight);             // :320-341 (one slab: no-op)

    // analysis
    void reset_timings();                            // :403-414
    void record_timings(const bool enablePointTransfer);  // :344-401
    void normalize_timings(int cycles);              // :416-426

    // helpers
    int getLeftBufferCount() { return 0; }
    int getRightBufferCount() { return 0; }
    int getMaxDeviationValue() { return 0; }

    // Device-side substep loop and re-sort (no counterpart in GPU_Partition:
    // the reference drives every substep from the host).
    void step(int n_substeps, int flags);
    void sort_points();
    void set_precision(int precision) { Precision = precision; }

    static icy::SimParams *prms;  // gpu_partition.h:119
    int PartitionID = 0;
    int Device = 0;
    int Precision = ICEMPM_PRECISION_FP64;

    size_t nPtsPitch = 0, nGridPitch = 0;  // capacities in elements
    int nPts_partition = 0;                // points, including disabled
    int nPts_disabled = 0;
    int GridX_offset = 0;
    int GridX_partition = 0;
    uint8_t error_code = 0;
    std::vector<double> host_side_indenter_force_accumulator;

    int max_pts_sent = 0, max_pt_deviation = 0;
    float timing_10_P2GAndHalo = 0, timing_20_acceptHalo = 0, timing_30_updateGrid = 0, timing_40_G2P = 0,
          timing_60_ptsSent = 0, timing_70_ptsAccepted = 0, timing_stepTotal = 0;

    icempm_partition *handle = nullptr;
    int uploads = 0;                       // transfer_points_from_soa_to_device calls (no reference counterpart)

private:
    void check(int rc, const char *what) const;
};
