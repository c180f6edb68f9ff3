// kernels_fp64.hip -- FP64 binned P2G / G2P tile kernels for gfx950 (CDNA4).
//
// Replaces partition_kernel_p2g and partition_kernel_g2p
// (simulation/kernels.cuh:46-122, 305-414); the grid update between them is in
// kernels_fp64_update.hip.
//
// Layout: points are the reference's SoA (16 lines of `pitch` doubles); each
// scene's grid is 3 lines (mass, momentum->velocity x, y) of `gpitch`
// doubles with node (col, row) at col*gridY + row, col = gx + halo.
//
// The binned kernels work on tiles of TX x TY cells: a workgroup stages the
// grid velocities of the tile's node window in LDS (G2P gathers), and sums
// the P2G contributions of the tile's points in an LDS window (ds_add_f64)
// that is flushed with one global atomic per touched node and line.  The
// point runs of a tile come from the cell_start table of the last sort; a
// point that drifted out of the window since then falls back to global memory.
//
// Products and sums may contract to FMA (as the reference's nvcc build does by
// default); the FP64 oracle rounds each operation, so the two agree to a few
// ulps per operation plus the reordering of the atomic sums.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "point_fp64.h"

namespace icempm {
namespace fp64 {

// ---------------------------------------------------------------------------
// Binned tile kernel: G2P from grid[rbuf] and/or P2G into grid[wbuf].
// 3 waves per SIMD (up to 168 VGPRs; the fused kernel fits with no spills).
// What keeps it there: the scene invariants of the point loop (4/h, kappa/2,
// sqrt(2) mu, the failure-surface and tension-branch terms) come from the host
// in scalar registers, and the water stress uses J^-5 = 1/(J^2 J^2 J) rather
// than pow(), whose polynomial constants the compiler would otherwise keep in
// vector registers across the loop (242 VGPRs, 2 waves per SIMD, before).
// ---------------------------------------------------------------------------
//
// WHC is the window height (ty + 2 + 2*margin) when fixed at compile time for
// the common tile shapes (0: read from the arguments).  A compile-time height
// turns the stencil's LDS offsets into instruction immediates.
template <int MODE, int WHC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) k_tiles(StepArgs a, int rbuf, int wbuf,
                                                                                        int record_mode)
{
    constexpr bool DO_G2P = (MODE & kTileG2P) != 0;
    constexpr bool DO_P2G = (MODE & kTileP2G) != 0;
    extern __shared__ double lds_raw[];
    __shared__ long long run_start[kMaxTX];
    __shared__ int run_pre[kMaxTX + 1];

    const int win_h = WHC > 0 ? WHC : a.win_h, win_w = a.win_w;
    const int WN = win_w * win_h;
    lds_d *lds = (lds_d *)lds_raw;
    lds_d *vw = lds;                            // 2*WN: (vx, vy) per window node
    // Two accumulation windows of 3*WN, (m, px, py) per node: even lanes add
    // into the first, odd lanes into the second.  The adjacent lanes of a
    // 16-lane row hold points 4 apart in the sorted order, which share a cell
    // (and so every LDS address) whenever the cell holds more than 4 points;
    // with a window each they never collide.  The flush adds the two.
    lds_d *acc = lds + (DO_G2P ? 2 * WN : 0);
    lds_d *const acc_lane = acc + ((threadIdx.x & 1) ? 3 * WN : 0);
    const int tid = threadIdx.x;
    const int ntiles = *a.n_tiles;
    const PtRef R{(double *)a.pts, a.pitch};

    // Tile distribution.  Static: block b takes tiles b', b'+G, ... (b' the
    // XCD-contiguous remap of b).  Dynamic: each XCD (blockIdx.x % 8, the
    // hardware's round-robin dispatch order) owns a contiguous eighth of the
    // tile list; its blocks pull tiles with an atomic counter and then help the
    // other XCDs' ranges.  Every block counts itself done once at the end; the
    // last one clears the counters for the next launch.
    __shared__ int s_tile;
    const bool dyn = a.dyn_tiles != 0;
    const int nq = (gridDim.x >= 8 && (gridDim.x & 7) == 0) ? 8 : 1;
    const int q0 = nq == 8 ? (int)(blockIdx.x & 7) : 0;
    int qk = 0;                      // queues visited so far (dynamic)
    int w_static = xcd_block();
    for (;;) {
        int w;
        if (dyn) {
            if (tid == 0) {
                int t = -1;
                for (; qk < nq; qk++) {
                    const int q = (q0 + qk) % nq;
                    const int hi = (int)((long long)ntiles * (q + 1) / nq);
                    const int got = (int)((long long)ntiles * q / nq) + atomicAdd(a.tile_queue + q, 1);
                    if (got < hi) { t = got; break; }
                }
                s_tile = t;
            }
            __syncthreads();
            w = __builtin_amdgcn_readfirstlane(s_tile);   // uniform: keeps tile data in scalar registers
            if (w < 0) break;
        } else {
            w = w_static;
            w_static += gridDim.x;
            if (w >= ntiles) break;
        }
        const TileItem ti = a.tiles[w];
        const SceneDev &S = a.scenes[ti.scene];
        const SceneConsts c = uniform_copy(S.c);
        const int gridY = c.gridY;
        const int wc0 = ti.col0 - 1 - a.margin, wr0 = ti.row0 - 1 - a.margin;
        const glb_d *gr = (const glb_d *)(a.grid_mem + S.grid_off[rbuf]);
        glb_d *gw = (glb_d *)(a.grid_mem + S.grid_off[wbuf]);
        const long long gp = S.gpitch;
        int *const dflag = a.dirty + wbuf * a.dirty_stride + uniform_copy(S.seg_off);   // segment flags of wbuf
        bool record = false;
        if (DO_G2P)
            record = record_mode >= 0 ? record_mode != 0 : (a.clocks[ti.scene].step % c.update_every) == 0;

        // point runs of the tile's columns
        if (tid < a.tx) {
            const int col = ti.col0 + tid;
            long long s0 = 0, s1 = 0;
            if (col < c.ncols) {
                const long long base = S.cell_off + (long long)col * gridY;
                s0 = a.cell_start[base + ti.row0];
                s1 = a.cell_start[base + min(ti.row0 + a.ty, gridY)];
            }
            run_start[tid] = s0;
            run_pre[tid + 1] = (int)(s1 - s0);
        }
        // stage velocities / clear accumulators
        for (int i = tid; i < WN; i += kBlock) {
            if (DO_G2P) {
                const int col = wc0 + i / win_h, row = wr0 + i % win_h;
                double vx = 0, vy = 0;
                if (col >= 0 && col < c.ncols && row >= 0 && row < gridY) {
                    const long long gi = (long long)col * gridY + row;
                    vx = gr[gp + gi];
                    vy = gr[2 * gp + gi];
                }
                vw[2 * i] = vx;
                vw[2 * i + 1] = vy;
            }
            if (DO_P2G) {
#pragma unroll
                for (int q = 0; q < 3; q++) acc[3 * i + q] = acc[3 * WN + 3 * i + q] = 0;
            }
        }
        __syncthreads();
        if (tid == 0) {
            run_pre[0] = 0;
            for (int k = 0; k < a.tx; k++) run_pre[k + 1] += run_pre[k];
        }
        __syncthreads();
        const int total = run_pre[a.tx];

        // points of the tile, strided over the workgroup: consecutive lanes take
        // consecutive points of a run, so every SoA access is coalesced
        unsigned err = 0;
        int cidx = 0;
        // lane_spread: the lanes of one atomic instruction then hold points of
        // different cells (about 4 points per cell), so their LDS targets differ.
        //   0: lane l takes point l of its wave's 64-point group;
        //   1: point (l%16)*4 + l/16 of the group (16 cells per group);
        //   2: the wave takes a 128-point group in two passes, lane l taking
        //      point (l%32)*4 + l/32, plus 2 in the second pass (32 cells).
        static_assert(kBlock == 256, "lane spread 2 assumes 4 waves of 64");
        const int lane = tid & 63, wave = tid >> 6;
        const int n_it = a.lane_spread == 2 ? 2 * ((total + 2 * kBlock - 1) / (2 * kBlock))
                                            : (total + kBlock - 1) / kBlock;
        // point k of the tile -> SoA index; false past the end
        auto locate = [&](int it, long long &pt) -> bool {
            int k;
            if (a.lane_spread == 2)
                k = (it >> 1) * (2 * kBlock) + wave * 128 + ((lane & 31) << 2) + (lane >> 5) + 2 * (it & 1);
            else if (a.lane_spread == 1)
                k = it * kBlock + (wave << 6) + ((lane & 15) << 2) + (lane >> 4);
            else
                k = it * kBlock + tid;
            if (k >= total) return false;
            if (k < run_pre[cidx]) cidx = 0;
            while (k >= run_pre[cidx + 1]) cidx++;
            pt = run_start[cidx] + (k - run_pre[cidx]);
            return true;
        };
        // The state of the next point is loaded while the current one is
        // computed.  Its index is located one iteration earlier still: the
        // lookup reads LDS, and LDS reads wait for this wave's earlier LDS
        // atomics, so it is done after the loads are issued.
        long long pt_n = 0, ud_n = 0, pt_nn = 0;
        double x_n = 0, y_n = 0, F_n[4] = {0, 0, 0, 0}, Jp_n = 0;
        bool ok_n = n_it > 0 && locate(0, pt_n);
        if (ok_n) {
            ud_n = as_ll(R.at(kLineUtility, pt_n));
            load_state(R, pt_n, x_n, y_n, F_n, Jp_n);
        }
        bool ok_nn = n_it > 1 && locate(1, pt_nn);
#pragma unroll 1
        for (int it = 0; it < n_it; it++) {
            const bool ok = ok_n;
            const long long pt = pt_n;
            long long ud = ud_n;
            double x = x_n, y = y_n, F[4] = {F_n[0], F_n[1], F_n[2], F_n[3]}, Jp = Jp_n;
            ok_n = ok_nn;
            pt_n = pt_nn;
            if (ok_n) {
                ud_n = as_ll(R.at(kLineUtility, pt_n));
                load_state(R, pt_n, x_n, y_n, F_n, Jp_n);
            }
            ok_nn = it + 2 < n_it && locate(it + 2, pt_nn);
            if (!ok) continue;
            if (ud & kDisabled) continue;
            double v[2], B[4];
            StressTerms terms;   // stress terms of F, passed from G2P to P2G
            terms.valid = false;
            if (DO_G2P) {
                double pp, qq;
                unsigned e;
                if (!g2p_point(c, gr, gp, vw, wc0, wr0, win_w, win_h, x, y, v, B, F, Jp, ud, pp, qq, e, terms)) {
                    err = max(err, e);
                    continue;
                }
                if (DO_P2G) store_state_fused(R, pt, x, y, F, Jp, ud, record, pp, qq);
                else store_state(R, pt, x, y, v, F, B, Jp, ud, record, pp, qq);
            } else {
                v[0] = R.at(kLineVelX, pt);
                v[1] = R.at(kLineVelX + 1, pt);
#pragma unroll
                for (int m = 0; m < 4; m++) B[m] = R.at(kLineBp00 + m, pt);
            }
            if (DO_P2G) {
                P2GStencil st;
                const unsigned e = p2g_stencil(c, x, y, v, B, F, Jp, ud, terms, st);
                if (e) {
                    err = max(err, e);
                    continue;
                }
                const int lx = st.bxc - wc0, ly = st.by - wr0;
                if (lx >= 0 && lx + 2 < win_w && ly >= 0 && ly + 2 < win_h) {
                    lds_d *const p0 = acc_lane + 3 * (lx * win_h + ly);
                    st.nodes(c, [&](int i, int j, double dm, double dpx, double dpy) {
                        lds_d *const p = p0 + 3 * (i * win_h + j);
                        lds_add(p, dm);
                        lds_add(p + 1, dpx);
                        lds_add(p + 2, dpy);
                    });
                } else {
                    // the stencil leaves the window (a point that drifted since the sort)
                    st.nodes(c, [&](int i, int j, double dm, double dpx, double dpy) {
                        if (lx + i >= 0 && lx + i < win_w && ly + j >= 0 && ly + j < win_h) {
                            lds_d *const p = acc_lane + 3 * ((lx + i) * win_h + ly + j);
                            lds_add(p, dm);
                            lds_add(p + 1, dpx);
                            lds_add(p + 2, dpy);
                        } else {
                            const long long gi = (long long)(st.bxc + i) * gridY + st.by + j;
                            glb_add(gw + gi, dm);
                            glb_add(gw + gp + gi, dpx);
                            glb_add(gw + 2 * gp + gi, dpy);
                            dflag[gi >> kSegShift] = 1;
                        }
                    });
                }
            }
        }
        if (err) atomicMax(&a.err[ti.scene], err);
        if (DO_P2G) {
            __syncthreads();
            for (int i = tid; i < WN; i += kBlock) {
                const double m = acc[3 * i] + acc[3 * WN + 3 * i];
                if (m == 0) continue;
                const int col = wc0 + i / win_h, row = wr0 + i % win_h;
                if (col < 0 || col >= c.ncols || row < 0 || row >= gridY) continue;
                const long long gi = (long long)col * gridY + row;
                glb_add(gw + gi, m);
                glb_add(gw + gp + gi, acc[3 * i + 1] + acc[3 * WN + 3 * i + 1]);
                glb_add(gw + 2 * gp + gi, acc[3 * i + 2] + acc[3 * WN + 3 * i + 2]);
                dflag[gi >> kSegShift] = 1;
            }
        }
        __syncthreads();
    }
    if (dyn && tid == 0) {
        if (atomicAdd(a.tile_queue + 8, 1) == (int)gridDim.x - 1)
            for (int k = 0; k < 9; k++) a.tile_queue[k] = 0;
    }
}

// ---------------------------------------------------------------------------
size_t tile_lds_bytes(const StepArgs &a, int mode)
{
    const size_t WN = (size_t)a.win_w * a.win_h;
    size_t n = 0;
    if (mode & kTileG2P) n += 2 * WN;
    if (mode & kTileP2G) n += 6 * WN;   // two accumulation windows
    return n * sizeof(double);
}

// window heights compiled in: ty = 16, 32, 64 with a drift margin of 1 or 0
template <int MODE>
static hipError_t launch_tiles_wh(const StepArgs &a, int rbuf, int wbuf, int record_mode, int n_blocks, size_t lds,
                                  hipStream_t s)
{
    switch (a.win_h) {
    case 20: hipLaunchKernelGGL((k_tiles<MODE, 20>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode); break;
    case 36: hipLaunchKernelGGL((k_tiles<MODE, 36>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode); break;
    case 68: hipLaunchKernelGGL((k_tiles<MODE, 68>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode); break;
    case 18: hipLaunchKernelGGL((k_tiles<MODE, 18>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode); break;
    case 34: hipLaunchKernelGGL((k_tiles<MODE, 34>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode); break;
    case 66: hipLaunchKernelGGL((k_tiles<MODE, 66>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode); break;
    default: hipLaunchKernelGGL((k_tiles<MODE, 0>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, record_mode); break;
    }
    return hipGetLastError();
}

hipError_t launch_tiles(const StepArgs &a, int mode, int rbuf, int wbuf, int record_mode, int n_blocks,
                        hipStream_t s)
{
    if (a.tx > kMaxTX || n_blocks <= 0) return hipErrorInvalidValue;
    const size_t lds = tile_lds_bytes(a, mode);
    switch (mode) {
    case kTileG2P: return launch_tiles_wh<kTileG2P>(a, rbuf, wbuf, record_mode, n_blocks, lds, s);
    case kTileP2G: return launch_tiles_wh<kTileP2G>(a, rbuf, wbuf, record_mode, n_blocks, lds, s);
    case kTileFused: return launch_tiles_wh<kTileFused>(a, rbuf, wbuf, record_mode, n_blocks, lds, s);
    default: return hipErrorInvalidValue;
    }
}

template <int MODE>
static int resident_wh(const StepArgs &a, size_t lds)
{
    int n = 0;
    hipError_t e;
    switch (a.win_h) {
    case 20: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<MODE, 20>, kBlock, lds); break;
    case 36: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<MODE, 36>, kBlock, lds); break;
    case 68: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<MODE, 68>, kBlock, lds); break;
    case 18: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<MODE, 18>, kBlock, lds); break;
    case 34: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<MODE, 34>, kBlock, lds); break;
    case 66: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<MODE, 66>, kBlock, lds); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<MODE, 0>, kBlock, lds); break;
    }
    return e == hipSuccess ? n : 0;
}

int tiles_resident_per_cu(const StepArgs &a, int mode)
{
    const size_t lds = tile_lds_bytes(a, mode);
    switch (mode) {
    case kTileG2P: return resident_wh<kTileG2P>(a, lds);
    case kTileP2G: return resident_wh<kTileP2G>(a, lds);
    case kTileFused: return resident_wh<kTileFused>(a, lds);
    default: return 0;
    }
}

}  // namespace fp64
}  // namespace icempm
