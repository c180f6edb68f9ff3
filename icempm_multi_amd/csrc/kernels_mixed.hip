// kernels_mixed.hip -- mixed-precision binned tile kernels (the grid update
// is in kernels_mixed_update.hip).
//
// Same pipeline as kernels_fp64.hip, with FP32 point state (points read and
// written: 32 + 56 bytes per substep instead of 64 + 112) and compensated grid
// sums: within a tile the FP32 contributions are summed in FP64 LDS
// accumulators (ds_add_f64, exact for the few hundred terms a node gets per
// tile); across tiles each node value is split into an FP32 (hi, lo) pair and
// added with an atomic two-sum, so the stored hi + lo keeps ~48 bits.  The
// grid update forms hi + lo in FP64, divides momentum by mass in FP64 and
// stores the velocity in FP32.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "point_mixed.h"

namespace icempm {
namespace mixed {

// 4 waves per SIMD (<= 128 VGPRs).  WHC: the window height when fixed at
// compile time (0: read from the arguments), as in the FP64 form.
template <int MODE, int WHC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_tiles(StepArgs a, int rbuf, int wbuf, int record_mode)
{
    constexpr bool DO_G2P = (MODE & kTileG2P) != 0;
    constexpr bool DO_P2G = (MODE & kTileP2G) != 0;
    extern __shared__ double lds_raw[];
    __shared__ long long run_start[kMaxTX];
    __shared__ int run_pre[kMaxTX + 1];

    const int win_h = WHC > 0 ? WHC : a.win_h, win_w = a.win_w;
    const int WN = win_w * win_h;
    // Two FP64 accumulation windows of 3*WN, (m, px, py) per node, one for
    // even and one for odd lanes as in the FP64 form, then the FP32 velocity
    // window.
    lds_d *acc = (lds_d *)lds_raw;
    lds_d *const acc_lane = acc + ((threadIdx.x & 1) ? 3 * WN : 0);
    lds_f *vw = (lds_f *)(acc + (DO_P2G ? 6 * WN : 0));   // 2*WN: (vx, vy) per window node
    const int tid = threadIdx.x;
    const int ntiles = *a.n_tiles;
    const PtRef R{(unsigned *)a.pts, a.pitch};

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
        const ConstsF cf = to_float(c);
        const int gridY = c.gridY;
        const int wc0 = ti.col0 - 1 - a.margin, wr0 = ti.row0 - 1 - a.margin;
        const glb_f *gr = (const glb_f *)(a.grid_mem + S.grid_off[rbuf]);
        glb_f *gw = (glb_f *)(a.grid_mem + S.grid_off[wbuf]);
        const long long gp = S.gpitch;
        int *const dflag = a.dirty + wbuf * a.dirty_stride + uniform_copy(S.seg_off);   // segment flags of wbuf
        bool record = false;
        const int step = a.clocks[ti.scene].step;
        if (DO_G2P) record = record_mode >= 0 ? record_mode != 0 : (step % c.update_every) == 0;

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
        for (int i = tid; i < WN; i += kBlock) {
            if (DO_G2P) {
                const int col = wc0 + i / win_h, row = wr0 + i % win_h;
                float vx = 0, vy = 0;
                if (col >= 0 && col < c.ncols && row >= 0 && row < gridY) {
                    const long long gi = (long long)col * gridY + row;
                    vx = gr[kLinePxHi * gp + gi];
                    vy = gr[kLinePyHi * gp + gi];
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
        auto locate = [&](int it, unsigned &pt) -> bool {
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
        // computed; its index is located one iteration earlier still (as in
        // the FP64 form: the LDS lookup waits for earlier LDS atomics).
        unsigned pt_n = 0, pt_nn = 0;   // SoA indices (points per launch < 2^29)
        unsigned ud_n = 0, xq_n = 0, yq_n = 0;
        float D_n[4] = {0, 0, 0, 0}, Jm1_n = 0;
        bool ok_n = n_it > 0 && locate(0, pt_n);
        if (ok_n) {
            ud_n = R.u(kLineUtility, pt_n);
            load_state(R, pt_n, xq_n, yq_n, D_n, Jm1_n);
        }
        bool ok_nn = n_it > 1 && locate(1, pt_nn);
#pragma unroll 1
        for (int it = 0; it < n_it; it++) {
            const bool ok = ok_n;
            const unsigned pt = pt_n;
            unsigned ud = ud_n, xq = xq_n, yq = yq_n;
            float D[4] = {D_n[0], D_n[1], D_n[2], D_n[3]}, Jm1 = Jm1_n, v[2], C[4];
            ok_n = ok_nn;
            pt_n = pt_nn;
            if (ok_n) {
                ud_n = R.u(kLineUtility, pt_n);
                load_state(R, pt_n, xq_n, yq_n, D_n, Jm1_n);
            }
            ok_nn = it + 2 < n_it && locate(it + 2, pt_nn);
            if (!ok) continue;
            if (ud & (unsigned)kDisabled) continue;
            StressTerms terms;   // stress terms of D, passed from G2P to P2G
            terms.valid = false;
            if (DO_G2P) {
                float pp, qq;
                unsigned e;
                if (!g2p_point(c, cf, gr, gp, vw, wc0, wr0, win_w, win_h, xq, yq, v, C, D, Jm1, ud, pp, qq,
                               e, (unsigned)pt, (unsigned)step, terms)) {
                    err = max(err, e);
                    continue;
                }
                if (DO_P2G) store_state_fused(R, pt, xq, yq, D, Jm1, ud, record, pp, qq);
                else store_state(R, pt, xq, yq, v, D, C, Jm1, ud, record, pp, qq);
            } else {
                v[0] = R.f(kLineVelX, pt);
                v[1] = R.f(kLineVelX + 1, pt);
#pragma unroll
                for (int m = 0; m < 4; m++) C[m] = R.f(kLineBp00 + m, pt);
            }
            if (DO_P2G) {
                P2GStencil st;
                const unsigned e = p2g_stencil(c, cf, xq, yq, v, C, D, Jm1, ud, terms, st);
                if (e) {
                    err = max(err, e);
                    continue;
                }
                const int lx = st.bxc - wc0, ly = st.by - wr0;
                if (lx >= 0 && lx + 2 < win_w && ly >= 0 && ly + 2 < win_h) {
                    lds_d *const p0 = acc_lane + 3 * (lx * win_h + ly);
                    st.nodes(cf, [&](int i, int j, double dm, double dpx, double dpy) {
                        lds_d *const p = p0 + 3 * (i * win_h + j);
                        lds_add(p, dm);
                        lds_add(p + 1, dpx);
                        lds_add(p + 2, dpy);
                    });
                } else {
                    // the stencil leaves the window (a point that drifted since the sort)
                    st.nodes(cf, [&](int i, int j, double dm, double dpx, double dpy) {
                        if (lx + i >= 0 && lx + i < win_w && ly + j >= 0 && ly + j < win_h) {
                            lds_d *const p = acc_lane + 3 * ((lx + i) * win_h + ly + j);
                            lds_add(p, dm);
                            lds_add(p + 1, dpx);
                            lds_add(p + 2, dpy);
                        } else {
                            const long long gi = (long long)(st.bxc + i) * gridY + st.by + j;
                            twosum_add_d(gw + kLineMHi * gp + gi, gw + kLineMLo * gp + gi, dm);
                            twosum_add_d(gw + kLinePxHi * gp + gi, gw + kLinePxLo * gp + gi, dpx);
                            twosum_add_d(gw + kLinePyHi * gp + gi, gw + kLinePyLo * gp + gi, dpy);
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
                twosum_add_d(gw + kLineMHi * gp + gi, gw + kLineMLo * gp + gi, m);
                twosum_add_d(gw + kLinePxHi * gp + gi, gw + kLinePxLo * gp + gi, acc[3 * i + 1] + acc[3 * WN + 3 * i + 1]);
                twosum_add_d(gw + kLinePyHi * gp + gi, gw + kLinePyLo * gp + gi, acc[3 * i + 2] + acc[3 * WN + 3 * i + 2]);
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

size_t tile_lds_bytes(const StepArgs &a, int mode)
{
    const size_t WN = (size_t)a.win_w * a.win_h;
    size_t n = 0;
    if (mode & kTileP2G) n += 6 * WN * sizeof(double);   // two accumulation windows
    if (mode & kTileG2P) n += 2 * WN * sizeof(float);
    return n;
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

}  // namespace mixed
}  // namespace icempm
