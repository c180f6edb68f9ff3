// gpu_implementation5.h -- GPU_Implementation5 (simulation/gpu_implementation5.h:24-58):
// the partitions of a simulation and the host-side point buffer.  With one
// device the grid is one slab spanning GridXTotal, so the halo exchange and
// point migration calls have nothing to move.
#pragma once

#include <array>
#include <functional>
#include <vector>

#include "gpu_partition.h"
#include "host_side_soa.h"

namespace icy { class Model; }

class GPU_Implementation5 {
public:
    icy::Model *model = nullptr;
    std::vector<GPU_Partition> partitions;
    HostSideSOA hssoa;
    std::array<double, 2> indenter_force{0.0, 0.0};   // Eigen::Vector2d in the reference

    std::function<void()> transfer_completion_callback;

    void allocate_arrays();                   // gpu_implementation5.cpp:329-343
    void split_hssoa_into_partitions();       // :284-326
    void transfer_ponts_to_device();          // :346-357

    void initialize_and_enable_peer_access(); // :221-281
    void transfer_from_device();              // :360-436

    void synchronize();                       // :439-446
    void update_constants();                  // :448-451
    void reset_grid();                        // :18-28
    void reset_indenter_force_accumulator();  // :453-460

    void p2g();                               // :31-77
    void receive_halos();                     // :80-112
    void update_nodes(double simulation_time);            // :115-122
    void g2p(const bool recordPQ, const bool enablePointTransfer);  // :125-133
    void receive_points();                    // :135-206
    void record_timings(const bool enablePointTransfer);  // :209-218

    // Device-side frame: n substeps of reset/P2G/update/G2P in one call, the
    // recordPQ cadence and clock kept on the device (no reference counterpart).
    void step(int n_substeps, int flags);
    void sort_points();                       // device RemoveDisabledAndSort
    // Rebalance of one partition on the device (replaces the host
    // RemoveDisabledAndSort + split + re-upload of model.cpp:102-110): the
    // device drops disabled points and re-sorts, and the host copy, which holds
    // the frame just downloaded, is permuted the same way (4 bytes per point
    // read back, nothing uploaded).  Returns the points kept.
    int rebalance_on_device();
    int device_rebalances = 0;                // rebalances done by rebalance_on_device

    int precision = ICEMPM_PRECISION_FP64;    // precision form of new partitions
};
