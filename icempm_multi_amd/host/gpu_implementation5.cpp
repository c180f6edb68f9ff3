// gpu_implementation5.cpp -- GPU_Implementation5 over GPU_Partition.
#include "gpu_implementation5.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "model.h"

void GPU_Implementation5::initialize_and_enable_peer_access()
{
    const icy::SimParams &prms = model->prms;
    if (prms.nPartitions != 1)
        throw std::runtime_error("this build runs one partition per process (nPartitions = " +
                                 std::to_string(prms.nPartitions) + ")");
    partitions.clear();
    partitions.resize(1);
    GPU_Partition &p = partitions[0];
    p.set_precision(precision);
    p.initialize(0, 0);
}

// One slab: all points belong to partition 0 and the grid span is the whole
// GridXTotal; the point count is the HostSideSOA size.
void GPU_Implementation5::split_hssoa_into_partitions()
{
    const icy::SimParams &prms = model->prms;
    GPU_Partition &p = partitions.at(0);
    p.nPts_disabled = 0;
    p.nPts_partition = (int)hssoa.size;
    p.GridX_offset = 0;
    p.GridX_partition = prms.GridXTotal;
    for (unsigned j = 0; j < hssoa.size; j++) hssoa.setPartition(j, 0);
}

void GPU_Implementation5::allocate_arrays()
{
    const icy::SimParams &prms = model->prms;
    const unsigned points_requested =
        (unsigned)((double)prms.nPtsTotal / partitions.size() * (1 + prms.ExtraSpaceForIncomingPoints));
    for (GPU_Partition &p : partitions)
        p.allocate((int)std::max<unsigned>(points_requested, (unsigned)p.nPts_partition), prms.GridXTotal);
}

void GPU_Implementation5::transfer_ponts_to_device()
{
    int points_uploaded = 0;
    for (GPU_Partition &p : partitions) {
        p.transfer_points_from_soa_to_device(hssoa, points_uploaded);
        points_uploaded += p.nPts_partition;
    }
}

void GPU_Implementation5::transfer_from_device()
{
    unsigned offset_pts = 0;
    for (GPU_Partition &p : partitions) {
        int n = 0, nd = 0;
        if (icempm_partition_get_point_count(p.handle, &n, &nd) != ICEMPM_OK)
            throw std::runtime_error(icempm_last_error());
        if (offset_pts + (unsigned)n > hssoa.capacity)
            throw std::runtime_error("transfer_from_device capacity exceeded");
        p.transfer_from_device(hssoa, (int)offset_pts);
        offset_pts += (unsigned)p.nPts_partition;
    }
    hssoa.size = offset_pts;

    indenter_force = {0.0, 0.0};
    for (GPU_Partition &p : partitions) {
        if (p.error_code) {
            std::fprintf(stderr, "P %d; error code %d\n", p.PartitionID, (int)p.error_code);
            throw std::runtime_error("error code");
        }
        for (int j = 0; j < model->prms.n_indenter_subdivisions; j++) {
            indenter_force[0] += p.host_side_indenter_force_accumulator[j * 2 + 0];
            indenter_force[1] += p.host_side_indenter_force_accumulator[j * 2 + 1];
        }
    }

    int count = 0;
    for (unsigned i = 0; i < hssoa.size; i++)
        if (!hssoa.getDisabledStatus(i)) count++;
    if (count != model->prms.nPtsTotal)
        throw std::runtime_error("transfer_from_device(): active point count mismatch");

    if (transfer_completion_callback) transfer_completion_callback();
}

void GPU_Implementation5::synchronize()
{
    for (GPU_Partition &p : partitions)
        if (icempm_partition_synchronize(p.handle) != ICEMPM_OK) throw std::runtime_error(icempm_last_error());
}

void GPU_Implementation5::update_constants()
{
    for (GPU_Partition &p : partitions) p.update_constants();
}

void GPU_Implementation5::reset_grid()
{
    for (GPU_Partition &p : partitions) p.reset_grid();
}

void GPU_Implementation5::reset_indenter_force_accumulator()
{
    for (GPU_Partition &p : partitions) {
        p.reset_indenter_force_accumulator();
        p.reset_timings();
    }
}

void GPU_Implementation5::p2g()
{
    for (GPU_Partition &p : partitions) p.p2g();
}

void GPU_Implementation5::receive_halos()
{
    for (GPU_Partition &p : partitions) p.receive_halos();
}

void GPU_Implementation5::update_nodes(double simulation_time)
{
    for (GPU_Partition &p : partitions) p.update_nodes(simulation_time);
}

void GPU_Implementation5::g2p(const bool recordPQ, const bool enablePointTransfer)
{
    for (GPU_Partition &p : partitions) p.g2p(recordPQ, enablePointTransfer);
}

void GPU_Implementation5::receive_points()
{
    for (GPU_Partition &p : partitions) p.receive_points(0, 0);
}

void GPU_Implementation5::record_timings(const bool enablePointTransfer)
{
    for (GPU_Partition &p : partitions) p.record_timings(enablePointTransfer);
}

void GPU_Implementation5::step(int n_substeps, int flags)
{
    for (GPU_Partition &p : partitions) p.step(n_substeps, flags);
}

void GPU_Implementation5::sort_points()
{
    for (GPU_Partition &p : partitions) p.sort_points();
}

int GPU_Implementation5::rebalance_on_device()
{
    if (partitions.size() != 1) throw std::runtime_error("rebalance_on_device: one partition only");
    GPU_Partition &p = partitions[0];
    int n0 = 0, nd0 = 0;
    if (icempm_partition_get_point_count(p.handle, &n0, &nd0) != ICEMPM_OK)
        throw std::runtime_error(icempm_last_error());
    if ((unsigned)n0 != hssoa.size)
        throw std::runtime_error("rebalance_on_device: the host copy does not hold the device's points");
    p.sort_points();
    int n = 0, nd = 0;
    if (icempm_partition_get_point_count(p.handle, &n, &nd) != ICEMPM_OK)
        throw std::runtime_error(icempm_last_error());
    std::vector<int32_t> perm((size_t)std::max(n, 1));
    if (icempm_partition_get_bins(p.handle, nullptr, perm.data(), n, nullptr, 0) != ICEMPM_OK)
        throw std::runtime_error(icempm_last_error());
    hssoa.ApplyPermutation(perm.data(), (unsigned)n);
    p.nPts_partition = n;
    p.nPts_disabled = 0;
    device_rebalances++;
    return n;
}
