// engine.h -- device-resident scenes, buffers and the substep pipeline.
//
// An Engine owns one HIP stream and a table of scenes on one device.  The
// C-ABI's partition is an Engine with one scene (GPU_Partition,
// gpu_partition.h:79-165); the batch is an Engine with many.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/icempm.h"
#include "kernels.h"
#include "mpm_common.h"

namespace icempm {


// The per-scene constants the kernels read, derived from the reference's
// SimParams (parameters_sim.cpp:128-165) plus the invariants of the point loop.
void scene_consts(const icempm_simparams &P, SceneConsts &c);
// The scene's failure-surface table (SceneDev::fail_tab) from its constants,
// with the reference's operations in its order (kernels.cuh:560-576, 598-619).
void scene_fail_table(const SceneConsts &c, double *tab);

enum KernelClass { kClsClock = 0, kClsZero, kClsP2G, kClsFused, kClsG2P, kClsUpdate, kClsSort, kNumCls };

// Tile candidates of the autotuner: {tx, ty, lane spread, dynamic queues}.
extern const int kTileCandidates[][4];
extern const int kNumTileCandidates;

// Grid buffers per scene.  A stepping call rotates through three: the tile
// launch of substep k reads buffer k (momenta), accumulates into k+1 and
// clears k+2 (mod 3).
constexpr int kGridBuffers = 3;
// FP64 lines per grid buffer: m, px, py (the node velocity in lines 1-2 after
// a separate grid update), in both precision forms.
constexpr int kGridLines = 3;

struct SceneHost {
    icempm_simparams prms{};
    bool have_prms = false;
    int n_pts = 0;            // points currently on the device (incl. disabled until a sort)
    int cap_pts = 0;
    long long pt_off = 0;
    long long cells_off = 0;
    long long grid_elems = 0; // gpitch
    long long seg_off = 0;    // first segment flag of the scene (per buffer)
    void *grid[3] = {nullptr, nullptr, nullptr};
    double *ind_force = nullptr;
    bool clean[3] = {true, true, true};   // the buffer holds zeros (and no segment flags)
    double sim_time = 0;      // SimulationTime
    int sim_step = 0;         // SimulationStep
};

class Engine {
public:
    // Units of a substep call (see emit_units in engine.hip).
    enum UnitKind { kUnitHead = 0, kUnitChain, kUnitTail, kUnitUnfused, kUnitDirect };

    Engine(int device, int precision);
    ~Engine();

    int device() const { return device_; }
    int precision() const { return precision_; }
    int elem_bytes() const { return precision_ == ICEMPM_PRECISION_MIXED ? 4 : 8; }
    hipStream_t stream() const { return stream_; }

    // set-up
    int add_scene(const icempm_simparams *prms, int cap_pts);       // returns scene id
    int set_params(int scene, const icempm_simparams &prms);
    int set_grid_span(int scene, int xoff, int gx);
    int commit();                                                    // (re)allocate device memory
    int ensure_committed();
    int upload(int scene, const double *hssoa, unsigned capacity, int offset, int n);
    int download(int scene, double *hssoa, unsigned capacity, int offset, int *n_out);

    // fine-grained phases of one scene on grid buffer 0 (the reference's call sequence)
    int reset_grid(int scene);
    int p2g(int scene);
    int update_nodes(int scene, double t, double ind_x, double ind_y);
    int g2p(int scene, int recordPQ);
    int reset_indenter_force();
    int synchronize();

    // frame-level stepping of all scenes
    int step(int nsub, int flags);
    int sort();
    int autotune(int n_trials, float *out_ms, int n_out, int *chosen);

    // queries
    int get_indenter_force(int scene, double *out, int n);
    int get_error_code(uint8_t *code);
    int get_bins(int scene, int32_t *keys, int32_t *perm, int n, int32_t *cell_counts, int n_cells);
    int get_kernel_stats(double *ms, long long *count, int n);
    int reset_kernel_stats();
    int get_grid(int scene, int buf, double *out, long long n_nodes);
    int get_position_map(int scene, int *fb, int *ox, int *oy);
    const SceneHost &scene(int i) const { return scenes_[i]; }
    int nscenes() const { return (int)scenes_.size(); }
    long long total_points() const;
    float last_step_ms() const { return last_step_ms_; }
    long long graph_captures() const { return n_captures_; }
    int cached_graphs() const { return (int)graphs_.size(); }
    void clock(int scene, double *t, int *step) const;

    int tx() const { return tx_; }
    int ty() const { return ty_; }
    int margin() const { return margin_; }
    int lane_spread() const { return lane_spread_; }
    int set_tile_shape(int tx, int ty, int margin);
    int get_tile_config(int *tx, int *ty, int *margin, int *lane_spread) const;
    void set_lane_spread(int v) { lane_spread_ = v < 0 ? 0 : (v > 2 ? 2 : v); invalidate_graphs(); }
    void set_dyn_tiles(int v) { dyn_tiles_ = v ? 1 : 0; invalidate_graphs(); }

private:
    int fail(int code, const std::string &msg);
    int hipfail(hipError_t e, const char *what);
    int build_scene_dev(int i, SceneDev &d) const;
    int push_scenes();
    int push_clocks();
    int build_tiles();
    int alloc_bins();
    StepArgs args() const;
    // timed: record hipEvents around the launches of every sample-th substep
    int launch_substeps(int nsub, int start_buf, bool fused, bool direct, bool timed, int sample = 1);
    struct GraphExec;
    // n units of one kind from grid buffer b and clock slot cs; unit j is timed
    // when timed and (k0 + j) % sample == 0.  With span, the n units of a chain
    // share one event pair around them, counted as n launches (back-to-back
    // launches, with no event between them).
    int emit_units(int kind, int n, int b, int cs, bool timed, int sample, int k0, bool span = false);
    int chunk_graph(int kind, int n, int b, int cs, int tmode, bool first_timed, GraphExec *&out);
    int run_units(int kind, int n, int &b, int &cs, bool use_graph, int tmode, int k0,
                  std::vector<GraphExec *> &replayed);
    int read_graph_events(const GraphExec &gx);
    int zero_grid(int scene, int buf);
    int start_buffer(bool fused, bool direct, int &b);
    void mark_buffers(bool fused, bool direct, int b, int nsub);
    int record(int cls, bool timed, hipError_t (*)(void *), void *);
    void timed_begin(bool timed);
    void timed_end(bool timed, int cls, int launches = 1);
    int collect_timings();
    void record(hipEvent_t e);
    void invalidate_graphs();

    int device_ = 0, precision_ = 0;
    hipStream_t stream_ = nullptr;
    std::vector<SceneHost> scenes_;
    bool committed_ = false;
    int n_cus_ = 256;

    // device memory
    void *pts_[2] = {nullptr, nullptr};
    long long pitch_ = 0;          // elements per SoA line
    long long n_pts_space_ = 0;    // point index space (sum of scene capacities)
    void *grid_mem_ = nullptr;
    double *ind_force_mem_ = nullptr;
    SceneDev *d_scenes_ = nullptr;
    SceneClock *d_clocks_ = nullptr;
    unsigned *d_err_ = nullptr;
    int *d_tile_queue_ = nullptr;
    int *d_dirty_ = nullptr;   // kGridBuffers * n_segs_ segment flags
    int *d_defer_ = nullptr;   // deferred-point list: 2 * n_pts_space_ ints and 64 counters
    int clock_slot_ = 0;       // slot of d_clocks_ (2 x nscenes) holding the current clock
    // workgroups of a grid update over n runs: a run per wave, persistent beyond 8 per CU
    int tile_blocks(int mode);      // workgroups of a tile launch: the resident count
    int resident_[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // per tile mode; reset when the tile shape changes
    int update_blocks(int n_runs) const { return std::max(1, std::min((n_runs + 3) / 4, n_cus_ * 8)); }
    long long n_segs_ = 0;
    NodeChunk *d_chunks_ = nullptr;
    int n_chunks_ = 0;

    // binning
    BinWork bins_{};
    void *bin_mem_ = nullptr;
    bool bins_valid_ = false;      // cell_start/tiles describe the current point order
    void *tile_mem_ = nullptr;
    TileItem *d_tiles_ = nullptr;
    int *d_n_tiles_ = nullptr;
    long long max_tiles_ = 0;
    int n_live_tiles_ = 0;
    int tx_ = 8, ty_ = 32, margin_ = 1;
    int lane_spread_ = 1;
    int dyn_tiles_ = 0;
    std::vector<SceneDev> host_dev_;

    // timings
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool_;
    std::vector<int> ev_cls_;
    std::vector<int> ev_w_;     // launches each event pair spans
    size_t ev_used_ = 0;
    double cls_ms_[kNumCls] = {0};
    long long cls_cnt_[kNumCls] = {0};
    float last_step_ms_ = 0;
    hipEvent_t ev_a_ = nullptr, ev_b_ = nullptr;
    bool capturing_ = false;                 // a graph capture is in progress
    static constexpr int kTimingSample = 32;  // ICEMPM_STEP_SAMPLED: one substep in 32 is timed

    // Graphs of chunks of 1, 2, 4, ..., kMaxChunk units, keyed by (unit kind,
    // units, grid buffer, clock slot, point buffer parity, timing mode, first
    // unit timed).  A timed graph owns the hipEvents its capture recorded
    // (event record nodes); they are read after each replay.
    struct GraphExec {
        hipGraphExec_t exec = nullptr;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
        std::vector<int> cls;
        std::vector<int> w;
    };
    typedef std::array<int, 7> GraphKey;
    static constexpr int kMaxChunk = 64;      // units per graph: at most 192 kernel nodes (3 per unfused unit)
    static constexpr size_t kMaxGraphs = 256; // the cache is cleared between calls beyond this
    // Graphs one call can add: head and tail 1 each; the chain's 64-unit
    // chunks start from each of the 3 grid buffers, and its shorter chunks
    // (32, 16, ..., 1 units) occur once each (9 + 2 = 11; unfused and
    // unbinned runs: 7).
    static constexpr size_t kMaxGraphsPerCall = 16;
    std::map<GraphKey, GraphExec> graphs_;
    int pts_flip_ = 0;                        // which buffer of the pair pts_[0] is (the sort swaps them)
    long long n_captures_ = 0;                // graphs captured since creation
};

}  // namespace icempm
