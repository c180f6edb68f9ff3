// kernels.h -- host-side launchers of the HIP kernels (one per .hip file).
#pragma once

#include <hip/hip_runtime.h>
#include "mpm_common.h"

namespace icempm {

// kernels_fp64.hip / kernels_mixed.hip share these signatures.
// Tile kernels: mode bit 1 = G2P from grid[rbuf], bit 2 = P2G into grid[wbuf].
// Bit 4 folds the grid update in: G2P stages node velocities computed from the
// momenta of grid[rbuf] (instead of reading velocities a grid update wrote),
// and the launch also clears the flagged segments of grid[zbuf], adds the
// indenter force of the nodes inside the indenter and advances the clocks.
enum TileMode {
    kTileG2P = 1,
    kTileP2G = 2,
    kTileFused = 3,
    kTileFromMomenta = 4,
    kTileTail = kTileG2P | kTileFromMomenta,     // last G2P of a call
    kTileChain = kTileFused | kTileFromMomenta,  // grid update + G2P(n) + P2G(n+1)
};

namespace fp64 {
// record_mode: 1/0 = write P, Q or not; -1 = decide from the scene clock.
hipError_t launch_tiles(const StepArgs &a, int mode, int rbuf, int wbuf, int zbuf, int record_mode, int n_blocks,
                        hipStream_t s);
hipError_t launch_p2g_direct(const StepArgs &a, int scene, int n_pts, int wbuf, hipStream_t s);
hipError_t launch_g2p_direct(const StepArgs &a, int scene, int n_pts, int rbuf, int record_pq, hipStream_t s);
// Grid update of both precision forms: their grids share one layout (3 FP64
// lines per buffer) and the update works in FP64.
hipError_t launch_update(const StepArgs &a, int buf, int zero_other, int advance_clock, int n_blocks, hipStream_t s);
size_t tile_lds_bytes(const StepArgs &a, int mode);
// Workgroups of the tile kernel for this mode and window that one CU holds at
// once (registers and LDS); 0 if the runtime cannot tell.
int tiles_resident_per_cu(const StepArgs &a, int mode);
}  // namespace fp64

namespace mixed {
// record_mode: 1/0 = write P, Q or not; -1 = decide from the scene clock.
hipError_t launch_tiles(const StepArgs &a, int mode, int rbuf, int wbuf, int zbuf, int record_mode, int n_blocks,
                        hipStream_t s);
hipError_t launch_p2g_direct(const StepArgs &a, int scene, int n_pts, int wbuf, hipStream_t s);
hipError_t launch_g2p_direct(const StepArgs &a, int scene, int n_pts, int rbuf, int record_pq, hipStream_t s);
size_t tile_lds_bytes(const StepArgs &a, int mode);
// Workgroups of the tile kernel for this mode and window that one CU holds at
// once (registers and LDS); 0 if the runtime cannot tell.
int tiles_resident_per_cu(const StepArgs &a, int mode);
}  // namespace mixed

// kernels_common.hip
// Clock of every scene: t += dt; ind_x = x0 + t*V; step counter (model.cpp:39-41).
hipError_t launch_fill_zero(void *ptr, size_t bytes, hipStream_t s);

// binning.hip -- device RemoveDisabledAndSort and tile tables.
struct BinWork {
    int *keys;          // per point: bin (cell) index within its scene, -1 if disabled
    int *cell_count;    // concatenated per-scene cell histograms (+1 slot per scene)
    int *cell_fill;     // scatter cursors
    int *cell_start;    // exclusive scan (global point indices)
    int *perm;          // sorted position -> previous point index
    int *scan_tmp;      // block sums of the scan
    int *scene_count;   // per scene: kept points after the sort
    int *sorted_keys;   // per sorted point: its reference cell key
    long long n_cells_total;
    long long n_pts_total;   // size of the point index space
};
// elem_bytes: 8 (FP64 SoA) or 4 (mixed SoA).
hipError_t launch_sort(const StepArgs &a, const SceneDev *host_scenes, int nscenes, BinWork &w, int elem_bytes,
                       void *pts_src, void *pts_dst, hipStream_t s);
hipError_t launch_build_tiles(const StepArgs &a, const SceneDev *host_scenes, int nscenes, const BinWork &w,
                              TileItem *tiles, int *n_tiles, long long max_tiles, hipStream_t s);

}  // namespace icempm
