/* h5io.h -- the HDF5 reads and writes of SnapshotManager, in C so that the
 * library that links the image's libhdf5 needs nothing else from its prefix
 * (the prefix also holds an older libstdc++).  0 = success; h5io_last_error()
 * describes a failure. */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SaveParametersAsAttributes (snapshotmanager.cpp:232-266) */
typedef struct h5io_attrs {
    double indenter_x, indenter_y, SimulationTime;
    int GridX, GridY, nPts;
    long long nPtsPitch;
    int UpdateEveryNthStep, n_indenter_subdivisions;
    double cellsize, IndDiameter, InitialTimeStep;
    int SetupType;
    double Volume;
} h5io_attrs;

const char *h5io_last_error(void);

/* LoadRawPoints input: datasets x, y (double), llGrainIDs (uint64) and the
 * float attribute `volume` on llGrainIDs. */
int h5io_raw_count(const char *path, long long *n);
int h5io_read_raw(const char *path, double *x, double *y, uint64_t *grains, long long n, float *volume);
int h5io_write_raw(const char *path, const double *x, const double *y, const uint64_t *grains, long long n,
                   float volume);

/* Snapshot: Params (bytes), Points (n_doubles doubles, attribute `size`),
 * Indenter_2D (n_ind doubles). */
int h5io_write_snapshot(const char *path, const void *params, size_t params_bytes, const double *points,
                        size_t n_doubles, unsigned size, const double *indenter, size_t n_ind,
                        const h5io_attrs *attrs);
int h5io_snapshot_sizes(const char *path, size_t *params_bytes, size_t *n_doubles);
int h5io_read_snapshot(const char *path, void *params, size_t params_bytes, double *points, size_t n_doubles,
                       unsigned *size);

/* SavePQ: PQ_intact and PQ_crushed as (count x 2) doubles. */
int h5io_write_pq(const char *path, const double *intact, size_t n_intact, const double *crushed, size_t n_crushed,
                  const h5io_attrs *attrs);

#ifdef __cplusplus
}
#endif
