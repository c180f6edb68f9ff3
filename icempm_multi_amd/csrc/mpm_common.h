// mpm_common.h -- POD types shared by the host runtime and the HIP kernels.
//
// Every launch works over a table of scenes.  A partition of the reference
// (GPU_Partition, gpu_partition.h:79-165) is a table of one scene; the batch
// form is a table of many.  Per-scene constants are derived once from the
// reference's SimParams (parameters_sim.cpp:128-165) by the host.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

namespace icempm {

constexpr int kNumPtsLines = 16;     // HostSideSOA lines (parameters_sim.h:27-35)
constexpr int kLineUtility = 0, kLineP = 1, kLineQ = 2, kLineJp = 3, kLinePosX = 4, kLineVelX = 6,
              kLineFe00 = 8, kLineBp00 = 12;
constexpr int64_t kCrushed = 0x10000, kDisabled = 0x20000, kLiquid = 0x40000;
constexpr int kMaxIndenterSub = 1024;
// Point lines are addressed by a 32-bit byte offset from the line base
// (index * 8 in the FP64 form), so the point space of a launch stays below 2^29.
constexpr long long kMaxPointSpace = 1LL << 29;
// Grid nodes are tracked in segments of 64 consecutive nodes (one wave): a
// per-segment flag says that a buffer may hold mass there, so the grid update
// reads and clears only flagged segments.
constexpr int kSegShift = 6, kSegNodes = 1 << kSegShift;
// Segments per grid-update work item (one wave; one flag per lane): short runs
// spread dense grids over many waves, long runs let a wave skip a large empty
// grid with one flag load per 64 segments.
inline int update_run_length(long long grid_nodes) { return grid_nodes > (8LL << 20) ? 64 : 16; }

// FP32 constants of the mixed form, converted by the host (the kernels keep
// them in scalar registers).
struct ConstsF {
    float h, hinv, dt, dt_vol_Dpinv, pmass, wmass, mu, kappa;
    float ics, its, iss, its2, beta, grain_var, s4;   // s4 = 4/h
    float dq_scale;                                   // dt * hinv * 2^frac_bits
    int fb, origin, origin_y;
    // grain-independent terms of the failure surface (kernels.cuh:560-619)
    float gv033, gv01;   // GrainVariability * 0.033 and * 0.1
    float q0c;           // 2 sqrt(-pmax pmin) / (pmax - pmin)
    float msq_c;         // 4 (1 + 2 beta) / (pmax - pmin)^2
    float b2, bpmax;     // 1 + 2 beta, beta * pmax
};

// The grain-dependent terms of the failure surface (kernels.cuh:560-576) for
// every remainder a 16-bit signed grain id leaves: entries 2r, 2r+1 hold mSq
// and q0 for (grain + 4) % 11 = r - 10 (r in [0, 20]), entry 2*21 + s holds
// pmin2 for (grain + 3) % 30 = s - 29 (s in [0, 58]).
constexpr int kFailTab11 = 21, kFailTab30 = 59;
constexpr int kFailTab = 2 * kFailTab11 + kFailTab30;

// Everything a kernel needs to know about one scene's physics and geometry.
struct SceneConsts {
    int gridX, gridY, halo, ncols;    // node column col = gx - xoff + halo, gx in [xoff-halo, xoff+gridX+halo)
    int xoff, gridXTotal;             // GridX_offset of the partition, GridXTotal of the domain
    int n_ind_sub, update_every;      // n_indenter_subdivisions, UpdateEveryNthStep
    int frac_bits, origin, origin_y, pad_c;  // mixed form: x/h = xq*2^-frac_bits - origin (y: origin_y)
    double h, hinv;
    double dt, dt_vol_Dpinv, dt_gravity, vmax, vmax_sq;
    double pmass, wmass;              // ParticleMass, WaterParticleMass
    double mu, kappa;
    double ics, its, iss, its2;       // Ice{Compressive,Tensile,Shear,Tensile2}Strength
    double beta, grain_var, tan_phi, dp_threshold_p;
    double ind_rsq, ind_vel, ind_x0, ind_y;
    // derived on the host with the kernels' own operations (bit-identical
    // results); read from scalar registers instead of recomputed per tile
    double s4;                        // 4 * hinv
    double kappa_half;                // kappa * 0.5
    double coeff1_mu;                 // sqrt(2) * mu (kernels.cuh:10, 555)
    double gv033, gv01;               // GrainVariability * 0.033, * 0.1
    double b2, bpmax;                 // 1 + 2 beta, beta * pmax
    double dp, dp2;                   // pmax - pmin, (pmax - pmin)^2
    double q0a;                       // 2 sqrt(-pmax pmin)
    double je_tension, sq_je_tension; // sqrt(-2 p_new / kappa + 1), its square root (p_new = -DP_threshold_p)
    ConstsF f;                        // mixed form
};

struct SceneDev {
    SceneConsts c;
    long long pt_off;       // first point of the scene in the SoA
    int n_pts;              // points of the scene (after the last compaction)
    int cap_pts;            // capacity of the scene's point range
    long long grid_off[3];  // byte offsets of the three accumulation grids from StepArgs::grid_mem
                            // (3 FP64 lines of gpitch elements, both forms)
    long long gpitch;       // elements per grid line (>= ncols*gridY)
    long long cell_off;     // first entry of the scene in the cell_start array
    int ncells;             // ncols*gridY
    int n_tiles_x, n_tiles_y;
    int pad0;
    long long force_off;    // first of 2*n_ind_sub doubles in StepArgs::force_mem
    long long seg_off;      // first segment flag of the scene in each half of StepArgs::dirty
    // grain-dependent terms of the FP64 failure surface (scene_fail_table)
    double fail_tab[kFailTab];
};

// Device clock of a scene, advanced on the device once per substep so that a
// captured graph can be replayed frame after frame (model.cpp:39-41,47).
struct SceneClock {
    double t;               // SimulationTime after the current substep
    double ind_x;           // indenter_x_initial + t*IndVelocity
    double ind_y;           // indenter_y
    int step;               // SimulationStep + substeps done before the current one (recordPQ test)
    int next_step;          // step of the next substep
};

// A tile of TX x TY cells of one scene: the unit of the binned P2G / fused G2P.
// (col0, row0) and dims (ncols | nrows << 16) give the bounding box of the
// tile's cells that held points at the last sort, inside the TX x TY block:
// the points are those of the box's cells, and the tile's LDS window is the
// box widened by the drift margin and the stencil's one node on every side.
// Tiles of a sparse scene (floe edges, gaps) thus stage fewer empty nodes.
struct TileItem {
    int scene;
    int col0, row0;
    int dims;
};

// Node window of a tile in LDS: origin (c0, r0) in grid nodes, w x h nodes,
// node i at column i / h and row i % h.
struct TileWindow {
    int c0, r0, w, h;
    float inv_h;   // 1/h, rounded

    __device__ __forceinline__ TileWindow(const TileItem &ti, int margin)
    {
        c0 = ti.col0 - 1 - margin;
        r0 = ti.row0 - 1 - margin;
        w = (ti.dims & 0xffff) + 2 + 2 * margin;
        h = (ti.dims >> 16) + 2 + 2 * margin;
        inv_h = 1.0f / (float)h;
    }
    // i / h for 0 <= i < w*h: (i + 1/2)/h is at least 1/(2h) away from an
    // integer, and the two roundings of the float product err by less than
    // (i + 1/2) * 2^-23.  Windows fit the 64 KiB of LDS a workgroup may hold
    // (at least 56 B per node: i < 1171, h < 391), so the error, under
    // 1.4e-4, stays below 1/(2h) > 1.2e-3 and the truncation is exact.
    __device__ __forceinline__ int col_of(int i) const { return (int)(((float)i + 0.5f) * inv_h); }
    __device__ __forceinline__ int nodes() const { return w * h; }
};

// A run of up to 64 node segments of one scene (up to 4096 nodes): the unit
// of work of one wave in the grid update.
struct NodeChunk {
    int scene;
    int seg0;     // first segment (node seg0*64)
    int nseg;     // segments in the run, 1..64
    int pad;
};

// Arguments common to all stepping kernels.
struct StepArgs {
    const SceneDev *scenes;
    char *grid_mem;           // base of every scene's grids (a kernel argument keeps it in the global space)
    double *force_mem;        // base of every scene's indenter force bins
    SceneClock *clocks;       // clock of the current substep, per scene
    SceneClock *clocks_next;  // the other clock slot: the grid update writes the next substep's clock here
    void *pts;                // SoA base: kNumPtsLines lines of `pitch` elements
    long long pitch;
    const int *cell_start;    // per scene ncells+1 global point indices (from the last sort)
    const TileItem *tiles;
    const int *n_tiles;       // device-side count of live tiles
    const NodeChunk *chunks;
    int n_chunks;
    int nscenes;
    unsigned *err;            // per-scene error byte (max over points)
    int tx, ty, margin;       // tile shape and drift margin of the LDS windows
    int win_w, win_h;         // tx+2+2*margin, ty+2+2*margin: the largest tile window (LDS size)
    int lane_spread;          // lane-to-point map of the tile kernels: 0 identity, 1 over 16 cells, 2 over 32 cells
    int dyn_tiles;            // 1: tiles pulled from per-XCD atomic queues, 0: static round robin
    int *tile_queue;          // 8 queue counters + 1 finished-block counter, zero between launches
    int *dirty;               // segment flags: [buffer 0 of all scenes][buffer 1 ...][buffer 2 ...]
    long long dirty_stride;   // segments per buffer (all scenes)
    // Points a kTileFromMomenta launch sets aside for its last workgroup (one
    // slot per point of the point space; defer_count[0] = entries, [1] = the
    // launch's finished workgroups; both zero between launches).
    int *defer_pts;
    int *defer_scene;
    int *defer_count;
};

// The clock of the next substep (model.cpp:39-41): t += dt by repeated
// addition, indenter_x = x0 + t*V, each operation rounded on its own so that
// the device clock equals the host's bit for bit.
__device__ __forceinline__ SceneClock next_clock(const SceneConsts &c, SceneClock k)
{
    k.t = __dadd_rn(k.t, c.dt);
    k.ind_x = __dadd_rn(c.ind_x0, __dmul_rn(k.t, c.ind_vel));
    k.ind_y = c.ind_y;
    k.step = k.next_step;
    k.next_step = k.step + 1;
    return k;
}

// A copy of workgroup-uniform data held in scalar registers: every dword goes
// through readfirstlane.  Read through a pointer instead, the constants would
// be reloaded from memory after every store or atomic the kernel issues (the
// compiler cannot rule out aliasing) and kept in vector registers.
template <class T>
__device__ __forceinline__ T uniform_copy(const T &src)
{
    static_assert(sizeof(T) % 4 == 0, "uniform_copy works on whole dwords");
    T out;
    const unsigned *s = reinterpret_cast<const unsigned *>(&src);
    unsigned *d = reinterpret_cast<unsigned *>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
    return out;
}

// Flags the segment of node idx of buffer buf of scene S as holding mass.
__device__ __forceinline__ void mark_dirty(const StepArgs &a, const SceneDev &S, int buf, long long idx)
{
    a.dirty[buf * a.dirty_stride + S.seg_off + (idx >> kSegShift)] = 1;
}

}  // namespace icempm
