// snapshotmanager.h -- icy::SnapshotManager (simulation/snapshotmanager.h:13-29).
//
// LoadRawPoints reads the input point cloud (HDF5 datasets x, y, llGrainIDs
// and the float32 attribute `volume` on llGrainIDs; snapshotmanager.cpp:29-121)
// and brings the model up.  SaveSnapshot/ReadSnapshot write and read the
// layout of the reference's (commented-out) snapshot code: a `Params` byte
// dataset holding SimParams, `Points` (16 capacity-pitched lines of doubles)
// and `Indenter_2D` (n_indenter_subdivisions x 2 forces); SavePQ writes
// PQ_intact / PQ_crushed (p, q) pairs (snapshotmanager.cpp:125-302).
// The HDF5 paths need the library built with ICEMPM_HAVE_HDF5; without it
// they throw.
#pragma once

#include <cstdint>
#include <string>

namespace icy { class SnapshotManager; class Model; }

class icy::SnapshotManager {
public:
    icy::Model *model = nullptr;

    void SaveSnapshot(std::string directory);
    void ReadSnapshot(std::string fileName);
    void LoadRawPoints(std::string fileName);
    void SavePQ(std::string directory);

    // The part of LoadRawPoints after the file read: n points with
    // coordinates x, y and utility words (grain id and flags), and the block
    // volume as stored in the file (float32).
    void LoadRawPointsFromArrays(const double *x, const double *y, const uint64_t *llGrainIDs, int n,
                                 float volume);
    // Writes the LoadRawPoints input layout (x, y, llGrainIDs + volume).
    static void SaveRawPoints(std::string fileName, const double *x, const double *y, const uint64_t *llGrainIDs,
                              int n, float volume);

    const std::string directory_snapshots = "snapshots";
    const std::string directory_pq = "pq";
};
