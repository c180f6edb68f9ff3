// model.cpp -- icy::Model::Step frame driver (simulation/model.cpp:27-113).
#include "model.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <stdexcept>

icy::Model::Model()
{
    prms.Reset();
    gpu.model = this;
    GPU_Partition::prms = &this->prms;
}

icy::Model::~Model() = default;

void icy::Model::open_logs()
{
    if (logDirectory.empty() || log_timing) return;
    std::filesystem::create_directories(logDirectory);
    log_timing = std::make_unique<std::ofstream>(logDirectory + "/timings.csv", std::ios::trunc);
    log_indenter_force = std::make_unique<std::ofstream>(logDirectory + "/indenter_force.csv", std::ios::trunc);
    log_indenter_force->precision(17);
}

bool icy::Model::Step()
{
    open_logs();
    double simulation_time = prms.SimulationTime;
    if (verbose)
        std::printf("step %d (%d) started; sim_time %.3g; host pts %u; cap %u\n", prms.SimulationStep,
                    prms.AnimationFrameNumber(), simulation_time, gpu.hssoa.size, gpu.hssoa.capacity);

    int count_unupdated_steps = 0;
    gpu.reset_indenter_force_accumulator();
    const auto t0 = std::chrono::steady_clock::now();

    if (hostDrivenSubsteps) {
        do {
            simulation_time += prms.InitialTimeStep;
            prms.indenter_x = prms.indenter_x_initial + simulation_time * prms.IndVelocity;

            gpu.reset_grid();
            gpu.p2g();
            gpu.receive_halos();
            gpu.update_nodes(simulation_time);
            const bool isZeroStep = (prms.SimulationStep + count_unupdated_steps) % prms.UpdateEveryNthStep == 0;
            gpu.g2p(isZeroStep, false);
            gpu.record_timings(false);
            count_unupdated_steps++;
        } while ((prms.SimulationStep + count_unupdated_steps) % prms.UpdateEveryNthStep != 0);
    } else {
        // substeps up to the next frame boundary; the device advances t by the
        // same repeated addition and sets indenter_x = x0 + t*V per substep
        const int nsub = prms.UpdateEveryNthStep - prms.SimulationStep % prms.UpdateEveryNthStep;
        int flags = ICEMPM_STEP_GRAPH;
        if (resortEveryFrame) flags |= ICEMPM_STEP_RESORT;
        if (recordKernelTimings) flags |= ICEMPM_STEP_TIMINGS;
        gpu.step(nsub, flags);
        for (int k = 0; k < nsub; k++) simulation_time += prms.InitialTimeStep;
        prms.indenter_x = prms.indenter_x_initial + simulation_time * prms.IndVelocity;
        count_unupdated_steps = nsub;
        gpu.record_timings(false);
    }

    processing_current_cycle_data.lock();   // held until the host has processed the previous frame
    accessing_point_data.lock();

    const bool self_unlock = !gpu.transfer_completion_callback;
    try {
        gpu.transfer_from_device();
    } catch (...) {
        accessing_point_data.unlock();
        processing_current_cycle_data.unlock();
        throw;
    }
    const double frame_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    max_pt_deviation = 0;
    for (GPU_Partition &p : gpu.partitions) {
        max_points_transferred = std::max(max_points_transferred, p.max_pts_sent);
        max_pt_deviation = std::max(max_pt_deviation, p.max_pt_deviation);
    }
    prms.SimulationTime = simulation_time;
    prms.SimulationStep += count_unupdated_steps;
    gpu.indenter_force[0] /= (double)count_unupdated_steps;
    gpu.indenter_force[1] /= (double)count_unupdated_steps;
    if (log_indenter_force)
        *log_indenter_force << prms.SimulationTime << ','
                            << std::hypot(gpu.indenter_force[0], gpu.indenter_force[1]) << ','
                            << gpu.indenter_force[0] << ',' << gpu.indenter_force[1] << '\n'
                            << std::flush;

    if (max_pt_deviation > prms.GridHaloSize / 2) prms.PointTransferFrequency++;

    bool rebalance = false;
    for (GPU_Partition &p : gpu.partitions) {
        p.normalize_timings(count_unupdated_steps);
        if (!recordKernelTimings && !hostDrivenSubsteps)
            p.timing_stepTotal = (float)(frame_ms * 1000.0 / count_unupdated_steps);
        const double freeSpacePercentage = (double)(p.nPtsPitch - p.nPts_partition) / p.nPts_partition;
        const double disabledPercentage = (double)p.nPts_disabled / p.nPtsPitch;
        if (freeSpacePercentage < prms.RebalanceThresholdFreeSpaceRemaining) rebalance = true;
        if (disabledPercentage > prms.RebalanceThresholdDisabledPercentage) rebalance = true;
        if (verbose)
            std::printf("%d-%d %8d %8d %7d | p2g %6.1f upd %6.1f g2p %6.1f | tot %7.1f ms/1000 steps\n",
                        p.PartitionID, p.Device, p.nPts_partition, (int)(p.nPtsPitch - p.nPts_partition),
                        p.nPts_disabled, p.timing_10_P2GAndHalo, p.timing_30_updateGrid, p.timing_40_G2P,
                        p.timing_stepTotal);
        if (log_timing)
            *log_timing << prms.AnimationFrameNumber() << ',' << p.PartitionID << ',' << p.nPts_partition << ','
                        << (p.nPtsPitch - p.nPts_partition) << ',' << p.nPts_disabled << ',' << p.max_pts_sent
                        << ',' << p.max_pt_deviation << ',' << p.timing_10_P2GAndHalo << ','
                        << p.timing_20_acceptHalo << ',' << p.timing_30_updateGrid << ',' << p.timing_40_G2P
                        << ',' << p.timing_60_ptsSent << ',' << p.timing_70_ptsAccepted << ','
                        << p.timing_stepTotal << '\n'
                        << std::flush;
    }

    if (rebalance) {
        if (gpu.partitions.size() == 1) {
            // one partition: compaction and sort on the device, the host copy
            // permuted to match (no re-upload of the 16 point lines)
            const unsigned before = gpu.hssoa.size;
            const int kept = gpu.rebalance_on_device();
            if (verbose)
                std::printf("rebalancing on device: %d points kept, %u disabled dropped, no re-upload\n", kept,
                            before - (unsigned)kept);
        } else {
            if (verbose) std::printf("squeezing and sorting HSSOA\n");
            gpu.hssoa.RemoveDisabledAndSort(prms.cellsize_inv, prms.GridY);
            gpu.split_hssoa_into_partitions();
            gpu.transfer_ponts_to_device();
        }
        SyncTopologyRequired = true;
        if (verbose) std::printf("rebalancing done\n");
    }
    accessing_point_data.unlock();
    // with no completion callback nobody else would release the frame
    if (self_unlock) processing_current_cycle_data.unlock();

    return prms.SimulationTime < prms.SimulationEndTime;
}

void icy::Model::UnlockCycleMutex() { processing_current_cycle_data.unlock(); }

void icy::Model::Reset()
{
    max_points_transferred = 0;
    prms.SimulationStep = 0;
    prms.SimulationTime = 0;
    SyncTopologyRequired = true;
}

void icy::Model::Prepare()
{
    abortRequested = false;
    gpu.update_constants();
}
