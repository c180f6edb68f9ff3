// snapshotmanager.cpp -- icy::SnapshotManager with the HDF5 C API.
#include "snapshotmanager.h"

#include <cstdio>
#include <cstring>
#include <filesystem>
#include <stdexcept>
#include <vector>

#include "model.h"

#ifdef ICEMPM_HAVE_HDF5
#include "h5io.h"
#endif

void icy::SnapshotManager::LoadRawPointsFromArrays(const double *x, const double *y, const uint64_t *grains, int n,
                                                   float volume)
{
    if (n <= 0) throw std::runtime_error("LoadRawPoints: no points");
    icy::Model &M = *model;
    HostSideSOA &S = M.gpu.hssoa;
    S.Allocate((unsigned)(n * (1 + M.prms.ExtraSpaceForIncomingPoints)));
    S.size = (unsigned)n;
    std::memcpy(S.getPointerToPosX(), x, sizeof(double) * n);
    std::memcpy(S.getPointerToPosY(), y, sizeof(double) * n);
    std::memcpy(S.host_buffer, grains, sizeof(double) * n);
    int kept = 0;
    if (icempm_prepare_raw_points(&M.prms, S.host_buffer, S.capacity, n, volume, &kept) != ICEMPM_OK)
        throw std::runtime_error(std::string("LoadRawPoints: ") + icempm_last_error());
    S.size = (unsigned)kept;

    M.gpu.initialize_and_enable_peer_access();
    M.gpu.split_hssoa_into_partitions();
    M.gpu.allocate_arrays();
    M.gpu.transfer_ponts_to_device();
    M.Reset();
    M.Prepare();
}

#ifdef ICEMPM_HAVE_HDF5

namespace {

void h5check(int rc)
{
    if (rc != 0) throw std::runtime_error(h5io_last_error());
}

h5io_attrs attrs_of(const icy::SimParams &p, long long nPtsPitch)
{
    h5io_attrs a{};
    a.indenter_x = p.indenter_x;
    a.indenter_y = p.indenter_y;
    a.SimulationTime = p.SimulationTime;
    a.GridX = p.GridXTotal;
    a.GridY = p.GridY;
    a.nPts = p.nPtsTotal;
    a.nPtsPitch = nPtsPitch;
    a.UpdateEveryNthStep = p.UpdateEveryNthStep;
    a.n_indenter_subdivisions = p.n_indenter_subdivisions;
    a.cellsize = p.cellsize;
    a.IndDiameter = p.IndDiameter;
    a.InitialTimeStep = p.InitialTimeStep;
    a.SetupType = p.SetupType;
    a.Volume = p.Volume;
    return a;
}

std::string frame_path(const std::string &dir, const std::string &sub, const char *fmt, int frame)
{
    std::filesystem::create_directories(dir + "/" + sub);
    char name[32];
    std::snprintf(name, sizeof name, fmt, frame);
    return dir + "/" + sub + "/" + name;
}

}  // namespace

void icy::SnapshotManager::LoadRawPoints(std::string fileName)
{
    if (!std::filesystem::exists(fileName)) throw std::runtime_error("error reading raw points file - no file");
    long long n = 0;
    h5check(h5io_raw_count(fileName.c_str(), &n));
    if (n <= 0 || n > (1LL << 31)) throw std::runtime_error("LoadRawPoints: bad point count");
    std::vector<double> x(n), y(n);
    std::vector<uint64_t> g(n);
    float volume = 0;
    h5check(h5io_read_raw(fileName.c_str(), x.data(), y.data(), g.data(), n, &volume));
    LoadRawPointsFromArrays(x.data(), y.data(), g.data(), (int)n, volume);
}

void icy::SnapshotManager::SaveRawPoints(std::string fileName, const double *x, const double *y,
                                         const uint64_t *grains, int n, float volume)
{
    h5check(h5io_write_raw(fileName.c_str(), x, y, grains, n, volume));
}

void icy::SnapshotManager::SaveSnapshot(std::string outputDirectory)
{
    icy::Model &M = *model;
    const std::string path = frame_path(outputDirectory, directory_snapshots, "d%05d.h5", M.prms.AnimationFrameNumber());
    std::vector<double> force((size_t)M.prms.n_indenter_subdivisions * 2, 0.0);
    if (!M.gpu.partitions.empty() && !M.gpu.partitions[0].host_side_indenter_force_accumulator.empty())
        force = M.gpu.partitions[0].host_side_indenter_force_accumulator;
    const HostSideSOA &S = M.gpu.hssoa;
    const h5io_attrs a = attrs_of(M.prms, S.capacity);
    h5check(h5io_write_snapshot(path.c_str(), &M.prms, sizeof(icy::SimParams), S.host_buffer,
                                (size_t)S.capacity * icy::SimParams::nPtsArrays, S.size, force.data(), force.size(),
                                &a));
}

// Restores the parameters and points of a snapshot and uploads them.
void icy::SnapshotManager::ReadSnapshot(std::string fileName)
{
    if (!std::filesystem::exists(fileName)) throw std::runtime_error("ReadSnapshot: no file");
    icy::Model &M = *model;
    size_t pb = 0, nd = 0;
    h5check(h5io_snapshot_sizes(fileName.c_str(), &pb, &nd));
    if (pb != sizeof(icy::SimParams)) throw std::runtime_error("ReadSnapshot: SimParams size mismatch");
    if (nd % icy::SimParams::nPtsArrays) throw std::runtime_error("ReadSnapshot: Points size");
    icy::SimParams tmp;
    M.gpu.hssoa.Allocate((unsigned)(nd / icy::SimParams::nPtsArrays));
    unsigned size = 0;
    h5check(h5io_read_snapshot(fileName.c_str(), &tmp, sizeof tmp, M.gpu.hssoa.host_buffer, nd, &size));
    const double ParticleViewSize = M.prms.ParticleViewSize;
    M.prms = tmp;
    M.prms.ParticleViewSize = ParticleViewSize;
    M.gpu.hssoa.size = size;

    M.gpu.initialize_and_enable_peer_access();
    M.gpu.split_hssoa_into_partitions();
    M.gpu.allocate_arrays();
    M.gpu.transfer_ponts_to_device();
    M.Prepare();
}

void icy::SnapshotManager::SavePQ(std::string outputDirectory)
{
    icy::Model &M = *model;
    const std::string path = frame_path(outputDirectory, directory_pq, "pq%05d.h5", M.prms.AnimationFrameNumber());
    std::vector<double> intact, crushed;
    HostSideSOA &S = M.gpu.hssoa;
    const double *P = S.getPointerToLine((int)icy::SimParams::idx_P);
    const double *Q = S.getPointerToLine((int)icy::SimParams::idx_Q);
    for (unsigned i = 0; i < S.size; i++) {
        if (S.getDisabledStatus(i)) continue;
        auto &b = S.getCrushedStatus(i) ? crushed : intact;
        b.push_back(P[i]);
        b.push_back(Q[i]);
    }
    const h5io_attrs a = attrs_of(M.prms, S.capacity);
    h5check(h5io_write_pq(path.c_str(), intact.data(), intact.size() / 2, crushed.data(), crushed.size() / 2, &a));
}

#else  // no HDF5

static void no_hdf5() { throw std::runtime_error("built without HDF5 (ICEMPM_HAVE_HDF5)"); }
void icy::SnapshotManager::LoadRawPoints(std::string) { no_hdf5(); }
void icy::SnapshotManager::SaveRawPoints(std::string, const double *, const double *, const uint64_t *, int, float)
{
    no_hdf5();
}
void icy::SnapshotManager::SaveSnapshot(std::string) { no_hdf5(); }
void icy::SnapshotManager::ReadSnapshot(std::string) { no_hdf5(); }
void icy::SnapshotManager::SavePQ(std::string) { no_hdf5(); }

#endif
