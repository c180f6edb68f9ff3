/*
 * icempm.h -- C-ABI of the MI355X-native 2D MPM ice stepper.
 *
 * The reference (igorg520b/iceMPM_multi) has no FFI: its drop-in boundary is the
 * C++ class surface of GPU_Partition (simulation/gpu_partition.h:79-165) and
 * GPU_Implementation5 (simulation/gpu_implementation5.h:25-59), driven by
 * icy::Model::Step (simulation/model.cpp:27-113).  Every entry point below names
 * the reference method it replaces.  Plain pointers and sizes only; every call
 * returns an int status (0 = ok, <0 = ICEMPM_ERR_*); device-side stencil faults
 * are reported through the error byte (70..73, 5) exactly as the reference's
 * gpu_error_indicator (simulation/kernels.cuh:20,94-97,506).
 *
 * Threading: one handle is used from one host thread at a time.  All work is
 * asynchronous on the handle's own HIP stream; the calls that return host data
 * (transfer_from_device, get_*) synchronise that stream first.
 */
#ifndef ICEMPM_H
#define ICEMPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Parameters: field-for-field mirror of icy::SimParams                       */
/* (simulation/parameters_sim.h:37-85).  The reference copies that struct by  */
/* sizeof into __constant__ memory (gpu_partition.cu:373), so the member       */
/* order and types are kept identical: the bytes of a reference SimParams can  */
/* be passed to icempm_partition_update_constants unchanged.                   */
/* ------------------------------------------------------------------------- */
typedef struct icempm_simparams {
    int n_indenter_subdivisions;
    int tpb_P2G, tpb_Upd, tpb_G2P;      /* launch hints; accepted, autotuner decides */

    int nPtsTotal;
    int GridXTotal, GridY, GridTotal;
    double GridXDimension;

    double InitialTimeStep, SimulationEndTime;
    int UpdateEveryNthStep;
    int SimulationStep;
    double SimulationTime;

    double Gravity, Density, WaterDensity, PoissonsRatio, YoungsModulus;
    double lambda, mu, kappa;

    double IceCompressiveStrength, IceTensileStrength, IceShearStrength, IceTensileStrength2;
    double NACC_beta, NACC_M, NACC_Msq;

    double DP_tan_phi, DP_threshold_p;

    double IndDiameter, IndRSq, IndVelocity, IndDepth;

    double cellsize, cellsize_inv, Dp_inv;
    double xmin, xmax, ymin, ymax;
    int nxmin, nxmax, nymin, nymax;

    double ParticleVolume, ParticleMass, WaterParticleMass, ParticleViewSize;

    double indenter_x, indenter_x_initial, indenter_y, indenter_y_initial;
    double Volume;
    int SetupType;
    double GrainVariability;

    int GridHaloSize;
    double ExtraSpaceForIncomingPoints;
    double PointsTransferBufferFraction;
    double RebalanceThresholdFreeSpaceRemaining;
    double RebalanceThresholdDisabledPercentage;
    int PointTransferFrequency;

    int nPartitions;
    int VectorCapacity_transfer;

    double dt_vol_Dpinv, dt_Gravity, vmax, vmax_squared;
    int gbOffset;
} icempm_simparams;

/* Host particle SoA line indices (parameters_sim.h:27-35): a HostSideSOA buffer
 * holds ICEMPM_N_PTS_ARRAYS lines of `capacity` doubles each. */
#define ICEMPM_IDX_UTILITY_DATA 0
#define ICEMPM_IDX_P            1
#define ICEMPM_IDX_Q            2
#define ICEMPM_IDX_JP_INV       3
#define ICEMPM_IDX_POSX         4
#define ICEMPM_IDX_VELX         6
#define ICEMPM_IDX_FE00         8
#define ICEMPM_IDX_BP00         12
#define ICEMPM_N_PTS_ARRAYS     16
#define ICEMPM_N_GRID_ARRAYS    3

/* utility_data bits (kernels.cuh:12-14, proxypoint2d.cpp:90-176) */
#define ICEMPM_STATUS_CRUSHED   0x10000LL
#define ICEMPM_STATUS_DISABLED  0x20000LL
#define ICEMPM_STATUS_LIQUID    0x40000LL

/* Precision forms */
#define ICEMPM_PRECISION_FP64   0   /* FP64 state and arithmetic (the reference's form) */
#define ICEMPM_PRECISION_MIXED  1   /* FP32 state, FP64 grid sums, FP64 islands */

/* Status codes */
#define ICEMPM_OK               0
#define ICEMPM_ERR_ARGUMENT    (-1)
#define ICEMPM_ERR_HIP         (-2)
#define ICEMPM_ERR_CAPACITY    (-3)
#define ICEMPM_ERR_STATE       (-4)
#define ICEMPM_ERR_NOMEM       (-5)
#define ICEMPM_ERR_DEVICE      (-6)   /* device error byte set; see get_error_code */

/* Device error byte values (kernels.cuh:94-97, 506) */
#define ICEMPM_DEVERR_LEFT      70
#define ICEMPM_DEVERR_BOTTOM    71
#define ICEMPM_DEVERR_RIGHT     72
#define ICEMPM_DEVERR_TOP       73
#define ICEMPM_DEVERR_NO_SPACE  5

/* Flags for the frame-level step calls */
#define ICEMPM_STEP_GRAPH       0x1   /* run the substeps as replays of hipGraphs of 1-64 substeps */
#define ICEMPM_STEP_RESORT      0x2   /* re-bin and re-sort points on device before the substeps */
#define ICEMPM_STEP_TIMINGS     0x4   /* record per-phase hipEvents (timing_* fields) */
#define ICEMPM_STEP_DIRECT_P2G  0x8   /* force the unbinned one-thread-per-point P2G */
#define ICEMPM_STEP_UNFUSED     0x10  /* keep G2P and the next P2G as separate launches */
#define ICEMPM_STEP_SAMPLED     0x20  /* with TIMINGS: record hipEvents around the launches of one
                                         substep in 32 only (substeps 0, 32, ... of each call; kernel
                                         stats then count those launches); with GRAPH|TIMINGS the
                                         events are recorded inside the graph */
#define ICEMPM_STEP_SPAN        0x40  /* with TIMINGS: one event pair around each run of a call's
                                         substep launches (a graph chunk's, or all of them with
                                         direct launches), counted as that many launches; the
                                         kernel stats then give the mean of back-to-back launches
                                         with no event between them (other phases as TIMINGS) */

/* Library identification and the last error text of the calling thread. */
const char *icempm_version(void);
const char *icempm_last_error(void);

/* ------------------------------------------------------------------------- */
/* Parameters and point preparation (host only, no device needed)             */
/* ------------------------------------------------------------------------- */
/* icy::SimParams::Reset (parameters_sim.cpp:5-60). */
int icempm_params_reset(icempm_simparams *prms);
/* icy::SimParams::ParseFile (parameters_sim.cpp:63-126) on the file's text; the
 * InputRawPoints value is copied to input_raw_points (n bytes, may be null). */
int icempm_params_parse_json(icempm_simparams *prms, const char *json_text, char *input_raw_points, size_t n);
/* ComputeCamClayParams2 + ComputeHelperVariables (parameters_sim.cpp:128-165). */
int icempm_params_derive(icempm_simparams *prms);
/* icy::SnapshotManager::LoadRawPoints after the HDF5 read (snapshotmanager.cpp:57-110):
 * x, y (lines 4-5) and llGrainIDs (line 0) of n_points points are in hssoa;
 * offsets the block, removes disabled points and sorts by cell key, sets
 * Fe = I and Jp_inv = 1, places the indenter and sets the particle volume from
 * the float32 `volume` attribute.  *n_kept = points left. */
int icempm_prepare_raw_points(icempm_simparams *prms, double *hssoa, unsigned hssoa_capacity, int n_points,
                              float volume, int *n_kept);

/* ------------------------------------------------------------------------- */
/* Partition: one grid + point set on one device (GPU_Partition)              */
/* ------------------------------------------------------------------------- */
typedef struct icempm_partition icempm_partition;

/* GPU_Partition::initialize (gpu_partition.cu:113-134): stream, events. */
int icempm_partition_create(int device, int partition_id, int precision, icempm_partition **out);
/* GPU_Partition::~GPU_Partition (gpu_partition.cu:88-111). */
int icempm_partition_destroy(icempm_partition *p);
/* GPU_Partition::allocate (gpu_partition.cu:137-197): points and grid capacity. */
int icempm_partition_allocate(icempm_partition *p, int n_points_capacity, int grid_x_capacity);
/* GPU_Partition::update_constants (gpu_partition.cu:200-208): SimParams bytes. */
int icempm_partition_update_constants(icempm_partition *p, const icempm_simparams *prms, size_t prms_bytes);
/* GPU_Partition::GridX_offset / GridX_partition fields (gpu_partition.h:124-125),
 * set by GPU_Implementation5::split_hssoa_into_partitions (gpu_implementation5.cpp:284-326). */
int icempm_partition_set_grid_span(icempm_partition *p, int gridx_offset, int gridx_partition);
/* GPU_Partition::transfer_points_from_soa_to_device (gpu_partition.cu:49-68):
 * n_points points starting at point_idx_offset of a HostSideSOA buffer whose
 * lines are hssoa_capacity doubles apart. */
int icempm_partition_transfer_points_from_soa_to_device(icempm_partition *p, const double *hssoa,
                                                        unsigned hssoa_capacity, int point_idx_offset,
                                                        int n_points);
/* GPU_Partition::transfer_from_device (gpu_partition.cu:16-46): points, indenter
 * force accumulator and error byte; synchronises the stream. */
int icempm_partition_transfer_from_device(icempm_partition *p, double *hssoa, unsigned hssoa_capacity,
                                          int point_idx_offset);
/* GPU_Partition::reset_grid (gpu_partition.cu:221-228). */
int icempm_partition_reset_grid(icempm_partition *p);
/* GPU_Partition::reset_indenter_force_accumulator (gpu_partition.cu:212-218). */
int icempm_partition_reset_indenter_force_accumulator(icempm_partition *p);
/* GPU_Partition::p2g (gpu_partition.cu:230-242) -> partition_kernel_p2g (kernels.cuh:46-122). */
int icempm_partition_p2g(icempm_partition *p);
/* GPU_Partition::update_nodes (gpu_partition.cu:244-257) -> kernels.cuh:164-301.
 * The indenter centre is an argument, as the reference reads prms->indenter_x/y
 * on the host at launch time (gpu_partition.cu:251). */
int icempm_partition_update_nodes(icempm_partition *p, double simulation_time, double indenter_x,
                                  double indenter_y);
/* GPU_Partition::g2p (gpu_partition.cu:259-293) -> kernels.cuh:305-414.
 * Point migration (kernels.cuh:416-461) is a multi-GPU feature and not built. */
int icempm_partition_g2p(icempm_partition *p, int recordPQ);
/* GPU_Implementation5::synchronize (gpu_implementation5.cpp:439-446). */
int icempm_partition_synchronize(icempm_partition *p);
/* Host copy of the indenter force accumulator (n_indenter_subdivisions x 2
 * doubles), as read in GPU_Implementation5::transfer_from_device (gpu_implementation5.cpp:392-396). */
int icempm_partition_get_indenter_force(icempm_partition *p, double *out, int n_doubles);
/* The error byte (gpu_error_indicator), read back after a sync; 0 = none. */
int icempm_partition_get_error_code(icempm_partition *p, uint8_t *code);
/* nPts_partition and nPts_disabled (gpu_partition.h:122-123).  The reference
 * counts in nPts_disabled the points handed to another partition
 * (gpu_implementation5.cpp:146), so with one partition it is always 0. */
int icempm_partition_get_point_count(icempm_partition *p, int *n_points, int *n_disabled);
/* GPU_Partition timing_* fields (gpu_partition.h:156-162) in ms summed since the
 * last reset_indenter_force_accumulator: out[0]=P2G, [1]=halo (0), [2]=update,
 * [3]=G2P, [4]=sent (0), [5]=accepted (0), [6]=total, [7]=sort. */
int icempm_partition_get_timings(icempm_partition *p, float *out, int n);

/* Device time per kernel class since the last reset_indenter_force_accumulator,
 * recorded with hipEvents around each launch when ICEMPM_STEP_TIMINGS is set
 * (around each run of back-to-back fused launches with ICEMPM_STEP_SPAN):
 * class 0 clock, 1 zero-fill, 2 P2G, 3 fused G2P+P2G (the grid update folded
 * in), 4 G2P (the last of a binned call also folds the grid update in), 5 grid
 * update (unbinned and unfused stepping, and the phase calls), 6 sort.
 * ms[k] = summed milliseconds, count[k] = launches. */
int icempm_partition_get_kernel_stats(icempm_partition *p, double *ms, long long *count, int n);

/* Frame-level substep loop of icy::Model::Step (model.cpp:38-54) run on the
 * device: n_substeps x (reset_grid, p2g, update_nodes(t), g2p(recordPQ)), with
 * t advanced by repeated addition of InitialTimeStep from prms.SimulationTime,
 * indenter_x = indenter_x_initial + t*IndVelocity, and recordPQ on substeps
 * where (SimulationStep + k) % UpdateEveryNthStep == 0.  SimulationTime and
 * SimulationStep of the partition's parameters advance by the substeps run. */
int icempm_partition_step(icempm_partition *p, int n_substeps, int flags);
/* Current SimulationTime / SimulationStep of the partition. */
int icempm_partition_get_clock(icempm_partition *p, double *simulation_time, int *simulation_step);

/* HostSideSOA::RemoveDisabledAndSort (host_side_soa.cpp:50-61) run on the
 * device: disabled points are removed and the rest ordered by the cell key
 * (int)(x/h+0.5)*GridY + (int)(y/h+0.5) (proxypoint2d.cpp:164-169), ties kept
 * in their previous order (a stable sort). */
int icempm_partition_sort_points(icempm_partition *p);
/* Binning results of the last sort, for checks: keys[n] per sorted point,
 * perm[n] = previous index of each sorted point, and the per-cell counts over
 * the grid's cell range (n_cells entries, cell = key + GridHaloSize*GridY). */
int icempm_partition_get_bins(icempm_partition *p, int32_t *keys, int32_t *perm, int n,
                              int32_t *cell_counts, int n_cells);

/* Copy of grid buffer `buf` (0, 1 or 2) as 3 lines of n_nodes doubles (mass,
 * momentum or velocity x, y), n_nodes = GridY*(GridXTotal + 2*GridHaloSize),
 * node (gx, gy) at (gx+GridHaloSize)*GridY + gy; both precision forms keep
 * the grid in FP64.  For inspection and tests.  No reference counterpart
 * (the reference keeps the grid internal). */
int icempm_partition_get_grid(icempm_partition *p, int buf, double *out, long long n_nodes);

/* Position map of the mixed form: a coordinate is stored as the unsigned
 * 32-bit xq with x/h = xq * 2^-frac_bits - origin_x (y: origin_y).  No
 * reference counterpart (the reference stores FP64 positions). */
int icempm_partition_get_position_map(icempm_partition *p, int *frac_bits, int *origin_x, int *origin_y);

/* Tile shape of the binned kernels: tx x ty cells per workgroup tile, LDS
 * windows widened by `margin` cells for points that drifted since the last
 * sort.  The autotuner sets it; callers may pin it. */
int icempm_partition_set_tile_shape(icempm_partition *p, int tx, int ty, int margin);
/* Lane-to-point map of the binned kernels (the autotuner sets it with the tile
 * shape): 0 = lane l of a wave takes point l of its group, 1 = points of 16
 * different cells per 16 lanes, 2 = points of 32 different cells per 32 lanes.
 * The results do not depend on it beyond the order of floating-point sums. */
int icempm_partition_set_lane_map(icempm_partition *p, int lane_spread);
/* Tile distribution of the binned kernels: 0 = static round robin over the
 * tile list, 1 = per-XCD atomic tile queues with stealing.  Set by the autotuner. */
int icempm_partition_set_tile_schedule(icempm_partition *p, int dynamic);

/* The autotuner's candidate tile shapes (no reference counterpart): fills the
 * first n of tx, ty, lane_spread (each may be null) and returns the number of
 * candidates, K. */
int icempm_tile_candidates(int *tx, int *ty, int *lane_spread, int n);

/* Project-written autotuner: times the substep launch of every candidate
 * tile shape, each with a window drift margin of 0 and of 1, on the current
 * point set after half a frame of substeps (the state drifts from its sort as
 * it does mid-frame), keeps the fastest and restores the state.  Results (ms
 * per substep launch; -1 for a shape that does not fit) go to out_ms if
 * non-null: the K candidates at margin 0, then at margin 1 (2K values).
 * *chosen = margin * K + candidate, the shape kept; the shape itself is read
 * with icempm_partition_get_tile_config. */
int icempm_partition_autotune(icempm_partition *p, int n_trials, float *out_ms, int n_out, int *chosen);
/* The tile shape, window drift margin and lane map in use (no reference counterpart). */
int icempm_partition_get_tile_config(icempm_partition *p, int *tx, int *ty, int *margin, int *lane_spread);
/* hipGraph cache of the stepping calls (no reference counterpart): graphs
 * captured since creation and graphs held now.  A call of n substeps with
 * ICEMPM_STEP_GRAPH replays graphs of 1, 2, 4, ..., 64 substeps, captured on
 * first use and kept across calls and re-sorts. */
int icempm_partition_get_graph_stats(icempm_partition *p, long long *captures, int *cached);

/* ------------------------------------------------------------------------- */
/* Batch: many independent scenes stepped by one persistent launch per phase   */
/* (no reference counterpart; the reference runs one scene per process).       */
/* ------------------------------------------------------------------------- */
typedef struct icempm_batch icempm_batch;

int icempm_batch_create(int device, int precision, icempm_batch **out);
int icempm_batch_destroy(icempm_batch *b);
/* Adds a scene: its parameters and n_points points from a HostSideSOA buffer
 * (n_points may be 0: the scene's clock still advances with the others). */
int icempm_batch_add_scene(icempm_batch *b, const icempm_simparams *prms, size_t prms_bytes,
                           const double *hssoa, unsigned hssoa_capacity, int n_points, int *scene_id);
/* Allocates device memory for all scenes, uploads them and bins them. */
int icempm_batch_commit(icempm_batch *b);
/* n_substeps of every scene (each with its own clock), ICEMPM_STEP_* flags. */
int icempm_batch_step(icempm_batch *b, int n_substeps, int flags);
int icempm_batch_transfer_from_device(icempm_batch *b, int scene_id, double *hssoa, unsigned hssoa_capacity);
int icempm_batch_get_indenter_force(icempm_batch *b, int scene_id, double *out, int n_doubles);
int icempm_batch_get_error_code(icempm_batch *b, uint8_t *code);
/* Per-scene inspection, as the partition functions of the same names (no
 * reference counterpart): points held (after the last compaction), the bins of
 * the last sort, a grid buffer, the mixed form's position map. */
int icempm_batch_get_scene_point_count(icempm_batch *b, int scene_id, int *n_points);
int icempm_batch_get_bins(icempm_batch *b, int scene_id, int32_t *keys, int32_t *perm, int n, int32_t *cell_counts,
                          int n_cells);
int icempm_batch_get_grid(icempm_batch *b, int scene_id, int buf, double *out, long long n_nodes);
int icempm_batch_get_position_map(icempm_batch *b, int scene_id, int *frac_bits, int *origin_x, int *origin_y);
int icempm_batch_get_scene_count(icempm_batch *b, int *n_scenes, long long *n_points_total);
/* The autotuner of icempm_partition_autotune over all scenes of the batch. */
int icempm_batch_autotune(icempm_batch *b, int n_trials, float *out_ms, int n_out, int *chosen);
/* As the partition functions of the same names. */
int icempm_batch_get_tile_config(icempm_batch *b, int *tx, int *ty, int *margin, int *lane_spread);
/* icempm_partition_set_tile_shape for every scene of the batch (no reference
 * counterpart): pins the shape the autotuner would choose. */
int icempm_batch_set_tile_shape(icempm_batch *b, int tx, int ty, int margin);
int icempm_batch_get_graph_stats(icempm_batch *b, long long *captures, int *cached);
/* Average device time per substep (ms) of the last icempm_batch_step call. */
int icempm_batch_last_step_ms(icempm_batch *b, float *ms_per_substep);
/* Device time per kernel class of the batch (classes as for partitions),
 * recorded when ICEMPM_STEP_TIMINGS is set; reset by icempm_batch_reset_kernel_stats. */
int icempm_batch_get_kernel_stats(icempm_batch *b, double *ms, long long *count, int n);
int icempm_batch_reset_kernel_stats(icempm_batch *b);

/* ---- scene generation ---------------------------------------------------
 * Poisson-disk sample of the box [x0, x1) x [y0, y1) with minimum distance r:
 * the point cloud of GrainProcessor2D::GenerateBlock/GenerateFloe
 * (generator/2d/grainprocessor2d.cpp:40-56, 250-262), which call
 * thinks::PoissonDiskSampling (generator/poisson_disk_sampling/
 * poisson_disk_sampling.h) with r = sqrt(0.656 * area / n).  Bridson's
 * algorithm with 30 candidates per active sample, SplitMix64 from `seed`
 * (deterministic).  Writes the points to x[], y[] (room for cap each) and
 * their count to *n; ICEMPM_ERR_CAPACITY if more than cap were placed. */
int icempm_poisson_disk(double r, double x0, double y0, double x1, double y1, uint64_t seed, double *x, double *y,
                        long long cap, long long *n);

/* ---- allocation-free substeps -------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc): n_substeps substeps of one scene on points the caller holds in
 * device memory, FP64, in the HostSideSOA layout (16 lines of `pitch`
 * doubles, parameters_sim.h:27-35), with the unbinned P2G and G2P of
 * icempm_partition_step(..., ICEMPM_STEP_DIRECT_P2G).  The clock starts at
 * prms->SimulationTime / SimulationStep and the indenter at
 * prms->indenter_x / indenter_y.  Every other buffer (grid, node flags,
 * clocks, error byte, force bins) is a static device array of the library,
 * about 24 MiB per GPU, held until the process ends; grids of more than
 * ICEMPM_NOALLOC_MAX_NODES nodes (halo columns included) are refused with
 * ICEMPM_ERR_CAPACITY.  The call allocates nothing and is asynchronous on
 * `stream` (a hipStream_t of `device`, or NULL for its null stream).  If
 * non-null, the device error byte goes to *d_err_out and the indenter force
 * bins (2 * n_indenter_subdivisions doubles, summed over the call's
 * substeps) to d_force_out, both device memory.  Calls that overlap on one
 * GPU share the static array and race. */
#define ICEMPM_NOALLOC_MAX_NODES (1 << 20)
int icempm_substeps_noalloc(int device, const icempm_simparams *prms, size_t prms_bytes, double *d_points,
                            unsigned pitch, int n_points, int n_substeps, void *stream, unsigned *d_err_out,
                            double *d_force_out);

#ifdef __cplusplus
}
#endif

#endif /* ICEMPM_H */
