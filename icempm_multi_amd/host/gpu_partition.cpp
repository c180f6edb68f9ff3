// gpu_partition.cpp -- GPU_Partition over the C-ABI (simulation/gpu_partition.cu).
#include "gpu_partition.h"

#include <stdexcept>
#include <string>
#include <utility>

icy::SimParams *GPU_Partition::prms = nullptr;

GPU_Partition::GPU_Partition() = default;

GPU_Partition::~GPU_Partition()
{
    if (handle) icempm_partition_destroy(handle);
}

GPU_Partition::GPU_Partition(GPU_Partition &&o) noexcept { *this = std::move(o); }

GPU_Partition &GPU_Partition::operator=(GPU_Partition &&o) noexcept
{
    if (this == &o) return *this;
    if (handle) icempm_partition_destroy(handle);
    handle = o.handle;
    o.handle = nullptr;
    PartitionID = o.PartitionID;
    uploads = o.uploads;
    Device = o.Device;
    Precision = o.Precision;
    nPtsPitch = o.nPtsPitch;
    nGridPitch = o.nGridPitch;
    nPts_partition = o.nPts_partition;
    nPts_disabled = o.nPts_disabled;
    GridX_offset = o.GridX_offset;
    GridX_partition = o.GridX_partition;
    error_code = o.error_code;
    host_side_indenter_force_accumulator = std::move(o.host_side_indenter_force_accumulator);
    max_pts_sent = o.max_pts_sent;
    max_pt_deviation = o.max_pt_deviation;
    timing_10_P2GAndHalo = o.timing_10_P2GAndHalo;
    timing_20_acceptHalo = o.timing_20_acceptHalo;
    timing_30_updateGrid = o.timing_30_updateGrid;
    timing_40_G2P = o.timing_40_G2P;
    timing_60_ptsSent = o.timing_60_ptsSent;
    timing_70_ptsAccepted = o.timing_70_ptsAccepted;
    timing_stepTotal = o.timing_stepTotal;
    return *this;
}

void GPU_Partition::check(int rc, const char *what) const
{
    if (rc == ICEMPM_OK) return;
    throw std::runtime_error(std::string("GPU_Partition ") + std::to_string(PartitionID) + " " + what + ": " +
                             icempm_last_error());
}

void GPU_Partition::initialize(int device, int partition)
{
    if (handle) { icempm_partition_destroy(handle); handle = nullptr; }
    PartitionID = partition;
    Device = device;
    check(icempm_partition_create(device, partition, Precision, &handle), "initialize");
}

void GPU_Partition::allocate(int n_points_capacity, int grid_x_capacity)
{
    check(icempm_partition_allocate(handle, n_points_capacity, grid_x_capacity), "allocate");
    nPtsPitch = (size_t)n_points_capacity;
    nGridPitch = (size_t)prms->GridY * (size_t)(grid_x_capacity + 2 * prms->GridHaloSize);
}

void GPU_Partition::update_constants()
{
    error_code = 0;
    check(icempm_partition_update_constants(handle, prms, sizeof(icempm_simparams)), "update_constants");
    if (GridX_partition > 0)
        check(icempm_partition_set_grid_span(handle, GridX_offset, GridX_partition), "set_grid_span");
}

void GPU_Partition::transfer_points_from_soa_to_device(HostSideSOA &hssoa, int point_idx_offset)
{
    check(icempm_partition_update_constants(handle, prms, sizeof(icempm_simparams)), "update_constants");
    check(icempm_partition_transfer_points_from_soa_to_device(handle, hssoa.host_buffer, hssoa.capacity,
                                                              point_idx_offset, nPts_partition),
          "transfer_points_from_soa_to_device");
    nPts_disabled = 0;
    uploads++;
}

void GPU_Partition::transfer_from_device(HostSideSOA &hssoa, int point_idx_offset)
{
    check(icempm_partition_transfer_from_device(handle, hssoa.host_buffer, hssoa.capacity, point_idx_offset),
          "transfer_from_device");
    host_side_indenter_force_accumulator.assign((size_t)prms->n_indenter_subdivisions * 2, 0.0);
    check(icempm_partition_get_indenter_force(handle, host_side_indenter_force_accumulator.data(),
                                              (int)host_side_indenter_force_accumulator.size()),
          "indenter force");
    check(icempm_partition_get_error_code(handle, &error_code), "error code");
    int n = 0, nd = 0;
    check(icempm_partition_get_point_count(handle, &n, &nd), "point count");
    nPts_partition = n;
    nPts_disabled = nd;
}

void GPU_Partition::reset_grid() { check(icempm_partition_reset_grid(handle), "reset_grid"); }

void GPU_Partition::reset_indenter_force_accumulator()
{
    check(icempm_partition_reset_indenter_force_accumulator(handle), "reset_indenter_force_accumulator");
}

void GPU_Partition::p2g() { check(icempm_partition_p2g(handle), "p2g"); }

void GPU_Partition::receive_halos() {}

void GPU_Partition::update_nodes(double simulation_time)
{
    check(icempm_partition_update_nodes(handle, simulation_time, prms->indenter_x, prms->indenter_y),
          "update_nodes");
}

void GPU_Partition::g2p(const bool recordPQ, const bool)
{
    check(icempm_partition_g2p(handle, recordPQ ? 1 : 0), "g2p");
}

void GPU_Partition::receive_points(int, int) {}

void GPU_Partition::reset_timings()
{
    max_pts_sent = 0;
    max_pt_deviation = 0;
    timing_10_P2GAndHalo = timing_20_acceptHalo = timing_30_updateGrid = timing_40_G2P = 0;
    timing_60_ptsSent = timing_70_ptsAccepted = timing_stepTotal = 0;
}

// The engine sums device time per kernel class since the last
// reset_indenter_force_accumulator; this copies the sums into the fields.
void GPU_Partition::record_timings(const bool)
{
    float t[8] = {0};
    check(icempm_partition_get_timings(handle, t, 8), "record_timings");
    timing_10_P2GAndHalo = t[0];
    timing_20_acceptHalo = t[1];
    timing_30_updateGrid = t[2];
    timing_40_G2P = t[3];
    timing_60_ptsSent = t[4];
    timing_70_ptsAccepted = t[5];
    timing_stepTotal = t[6];
}

void GPU_Partition::normalize_timings(int cycles)
{
    const float coeff = 1000.f / (float)cycles;
    timing_10_P2GAndHalo *= coeff;
    timing_20_acceptHalo *= coeff;
    timing_30_updateGrid *= coeff;
    timing_40_G2P *= coeff;
    timing_60_ptsSent *= coeff;
    timing_70_ptsAccepted *= coeff;
    timing_stepTotal *= coeff;
}

void GPU_Partition::step(int n_substeps, int flags)
{
    check(icempm_partition_step(handle, n_substeps, flags), "step");
}

void GPU_Partition::sort_points()
{
    check(icempm_partition_sort_points(handle), "sort_points");
    int n = 0, nd = 0;
    check(icempm_partition_get_point_count(handle, &n, &nd), "point count");
    nPts_partition = n;
    nPts_disabled = nd;
}
