// kernels_mixed.hip -- mixed-precision binned tile kernels (the separate grid
// update is the FP64 form's, kernels_fp64_update.hip).
//
// Same pipeline as kernels_fp64.hip, with FP32 point state (points read and
// written: 32 + 56 bytes per substep instead of 64 + 112) and the FP64 grid:
// within a tile the FP32 contributions are summed in FP64 LDS accumulators
// (ds_add_f64), and each node's sums are added into the grid with FP64
// atomics.  Node velocities are formed in FP64 (momentum over mass) and kept
// in FP32 in the tile's LDS window.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "point_mixed.h"

namespace icempm {
namespace mixed {

// Parts of the grid update that the tile kernels take over (kTileFromMomenta),
// as in the FP64 form (the grid is the same: 3 FP64 lines per buffer).
__device__ __forceinline__ void clear_flagged(const StepArgs &a, int zb)
{
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6), n_waves = gridDim.x * (kBlock / 64);
    for (int r = wave; r < a.n_chunks; r += n_waves) {
        const NodeChunk run = uniform_copy(a.chunks[r]);
        const SceneDev &S = a.scenes[run.scene];
        glb_d *g = (glb_d *)(a.grid_mem + S.grid_off[zb]);
        const long long gp = S.gpitch;
        int *fl = a.dirty + zb * a.dirty_stride + S.seg_off + run.seg0;
        const int f = lane < run.nseg ? fl[lane] : 0;
        unsigned long long mb = __ballot(f != 0);
        const long long node0 = (long long)run.seg0 << kSegShift;
        while (mb) {
            const int sg = __builtin_ctzll(mb);
            mb &= mb - 1;
            const long long idx = node0 + ((long long)sg << kSegShift) + lane;
#pragma unroll
            for (int l = 0; l < 3; l++) g[l * gp + idx] = 0;
        }
        if (f) fl[lane] = 0;
    }
}

// The indenter's reaction of the substep (kernels.cuh:265-287) from the nodes
// inside the indenter circle, one workgroup per scene (see the FP64 form).
__device__ __forceinline__ void indenter_force_nodes(const StepArgs &a, int sc, int rb)
{
    const SceneDev &S = a.scenes[sc];
    const SceneConsts c = uniform_copy(S.c);
    if (!(c.ind_rsq > 0)) return;
    const SceneClock clk = next_clock(c, uniform_copy(a.clocks[sc]));
    const double r = sqrt(c.ind_rsq);
    const int gx0 = max((int)floor((clk.ind_x - r) * c.hinv) - 1, c.xoff - c.halo);
    const int gx1 = min((int)ceil((clk.ind_x + r) * c.hinv) + 1, c.xoff - c.halo + c.ncols - 1);
    const int gy0 = max((int)floor((clk.ind_y - r) * c.hinv) - 1, 0);
    const int gy1 = min((int)ceil((clk.ind_y + r) * c.hinv) + 1, c.gridY - 1);
    if (gx1 < gx0 || gy1 < gy0) return;
    const int ny = gy1 - gy0 + 1;
    const long long nn = (long long)(gx1 - gx0 + 1) * ny;
    const glb_d *g = (const glb_d *)(a.grid_mem + S.grid_off[rb]);
    const long long gp = S.gpitch;
    double *force = a.force_mem + S.force_off;
    for (long long k = threadIdx.x; k < nn; k += kBlock) {
        const int gix = gx0 + (int)(k / ny), giy = gy0 + (int)(k % ny);
        const long long idx = (long long)(gix - c.xoff + c.halo) * c.gridY + giy;
        const double mass = g[kLineM * gp + idx];
        if (mass == 0) continue;
        const double px = g[kLinePx * gp + idx], py = g[kLinePy * gp + idx];
        double vx, vy;
        fp64::node_velocity(c, clk, force, gix, giy, mass, px, py, vx, vy);
    }
}

// Entry k of the deferred-point list, as in the FP64 form (G2P with node
// velocities formed from the global momenta; P2G with FP64 global atomics).
__device__ __forceinline__ void deferred_point(const StepArgs &a, bool p2g, int rb, int wb, int k)
{
    const PtRef R{(unsigned *)a.pts, a.pitch};
    const unsigned pt = (unsigned)a.defer_pts[k];
    const int sc = a.defer_scene[k];
    const SceneDev &S = a.scenes[sc];
    const SceneConsts &c = S.c;
    const ConstsF &cf = c.f;
    const SceneClock clk = next_clock(c, a.clocks[sc]);
    const bool record = (clk.step % c.update_every) == 0;
    unsigned ud = R.u(kLineUtility, pt), xq, yq;
    float D[4], Jm1, v[2], C[4], pp, qq;
    load_state(R, pt, xq, yq, D, Jm1);
    unsigned e;
    StressTerms terms;
    terms.valid = false;
    const glb_d *gr = (const glb_d *)(a.grid_mem + S.grid_off[rb]);
    if (!g2p_point(c, cf, gr, S.gpitch, &clk, nullptr, 0, 0, 0, 0, 0, false, xq, yq, v, C, D, Jm1, ud, pp, qq, e, pt,
                   (unsigned)clk.step, terms)) {
        atomicMax(&a.err[sc], e);
        return;
    }
    if (!p2g) {
        store_state(R, pt, xq, yq, v, D, C, Jm1, ud, record, pp, qq);
        return;
    }
    store_state_fused(R, pt, xq, yq, D, Jm1, ud, record, pp, qq);
    glb_d *gw = (glb_d *)(a.grid_mem + S.grid_off[wb]);
    const long long gp = S.gpitch;
    const int gridY = c.gridY;
    P2GStencil st;
    const unsigned e2 = p2g_stencil(c, cf, xq, yq, v, C, D, Jm1, ud, terms, st);
    if (e2) {
        atomicMax(&a.err[sc], e2);
        return;
    }
    st.nodes(cf, [&](int i, int j, double dm, double dpx, double dpy) {
        const long long gi = (long long)(st.bxc + i) * gridY + st.by + j;
        grid_add(gw, gp, gi, dm, dpx, dpy);
        mark_dirty(a, S, wb, gi);
    });
}

// 4 waves per SIMD (<= 128 VGPRs).  WHC: the LDS row stride (the largest
// window's height) when fixed at compile time (0: read from the arguments),
// as in the FP64 form; so are the tile windows and the folded grid update of
// kTileFromMomenta.
template <int MODE, int WHC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_tiles(StepArgs a, int rbuf, int wbuf, int zbuf, int record_mode)
{
    constexpr bool DO_G2P = (MODE & kTileG2P) != 0;
    constexpr bool DO_P2G = (MODE & kTileP2G) != 0;
    constexpr bool FROM_M = (MODE & kTileFromMomenta) != 0;
    extern __shared__ double lds_raw[];
    __shared__ long long run_start[kMaxTX];
    __shared__ int run_pre[kMaxTX + 1];

    // The LDS regions hold the largest window, a.win_w columns of WS nodes; a
    // tile's own window (TileWindow) takes its first columns and rows.
    const int WS = WHC > 0 ? WHC : a.win_h;
    const int WNmax = a.win_w * WS;
    // Two FP64 accumulation windows of 3*WNmax, (m, px, py) per node, one for
    // even and one for odd lanes as in the FP64 form, then the FP32 velocity
    // window.
    lds_d *acc = (lds_d *)lds_raw;
    lds_d *const acc_lane = acc + ((threadIdx.x & 1) ? 3 * WNmax : 0);
    lds_f *vw = (lds_f *)(acc + (DO_P2G ? 6 * WNmax : 0));   // 2*WNmax: (vx, vy) per window node
    const int tid = threadIdx.x;
    const int ntiles = *a.n_tiles;
    const PtRef R{(unsigned *)a.pts, a.pitch};

    if (FROM_M) {
        clear_flagged(a, zbuf);
        for (int sc = blockIdx.x; sc < a.nscenes; sc += gridDim.x) indenter_force_nodes(a, sc, rbuf);
        // the advanced clocks go to the other slot, which no block reads here
        if (blockIdx.x == 0)
            for (int sc = tid; sc < a.nscenes; sc += kBlock)
                a.clocks_next[sc] = next_clock(a.scenes[sc].c, a.clocks[sc]);
    }

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
    __shared__ int s_defer;          // this workgroup set points aside
    if (tid == 0) s_defer = 0;
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
        const TileWindow W(ti, a.margin);
        const int wc0 = W.c0, wr0 = W.r0, win_w = W.w, win_h = W.h, WN = W.nodes();
        const int ncol = ti.dims & 0xffff, nrow = ti.dims >> 16;
        const glb_d *gr = (const glb_d *)(a.grid_mem + S.grid_off[rbuf]);
        glb_d *gw = (glb_d *)(a.grid_mem + S.grid_off[wbuf]);
        const long long gp = S.gpitch;
        int *const dflag = a.dirty + wbuf * a.dirty_stride + uniform_copy(S.seg_off);   // segment flags of wbuf
        // the substep's clock: advanced here from the current slot, or already
        // advanced by the grid update
        SceneClock clk = uniform_copy(a.clocks[ti.scene]);
        if (FROM_M) clk = next_clock(c, clk);
        bool record = false;
        const int step = clk.step;
        if (DO_G2P) record = record_mode >= 0 ? record_mode != 0 : (step % c.update_every) == 0;

        // point runs of the columns of the tile's occupied box
        if (tid < ncol) {
            const int col = ti.col0 + tid;
            long long s0 = 0, s1 = 0;
            if (col < c.ncols) {
                const long long base = S.cell_off + (long long)col * gridY;
                s0 = a.cell_start[base + ti.row0];
                s1 = a.cell_start[base + min(ti.row0 + nrow, gridY)];
            }
            run_start[tid] = s0;
            run_pre[tid + 1] = (int)(s1 - s0);
        }
        if (DO_G2P && FROM_M) {
            // velocities formed from the momenta: four window nodes per
            // thread and pass, the loads of all four issued before any is used
            constexpr int kU = 4;
            for (int i0 = tid; i0 < WN; i0 += kU * kBlock) {
                double g3[kU][3];
                int qc[kU], qr[kU];
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    const int i = i0 + u * kBlock;
                    qc[u] = W.col_of(i);
                    qr[u] = i - qc[u] * win_h;
                    const int col = wc0 + qc[u], row = wr0 + qr[u];
                    const bool in = i < WN && col >= 0 && col < c.ncols && row >= 0 && row < gridY;
                    const long long gi = (long long)col * gridY + row;
#pragma unroll
                    for (int l = 0; l < 3; l++) g3[u][l] = in ? (double)gr[l * gp + gi] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    const int i = i0 + u * kBlock;
                    if (i >= WN) break;
                    const int col = wc0 + qc[u], row = wr0 + qr[u];
                    float vx = 0, vy = 0;
                    if (g3[u][kLineM] != 0) {
                        double dvx, dvy;
                        fp64::node_velocity(c, clk, nullptr, col - c.halo + c.xoff, row, g3[u][kLineM],
                                            g3[u][kLinePx], g3[u][kLinePy], dvx, dvy);
                        vx = (float)dvx;
                        vy = (float)dvy;
                    }
                    const int li = qc[u] * WS + qr[u];
                    vw[2 * li] = vx;
                    vw[2 * li + 1] = vy;
                }
            }
        }
        for (int i = tid; i < WN; i += kBlock) {
            const int qc = W.col_of(i), qr = i - qc * win_h, li = qc * WS + qr;
            if (DO_G2P && !FROM_M) {
                const int col = wc0 + qc, row = wr0 + qr;
                float vx = 0, vy = 0;
                if (col >= 0 && col < c.ncols && row >= 0 && row < gridY) {
                    const long long gi = (long long)col * gridY + row;
                    vx = (float)gr[kLinePx * gp + gi];
                    vy = (float)gr[kLinePy * gp + gi];
                }
                vw[2 * li] = vx;
                vw[2 * li + 1] = vy;
            }
            if (DO_P2G) {
#pragma unroll
                for (int q = 0; q < 3; q++) acc[3 * li + q] = acc[3 * WNmax + 3 * li + q] = 0;
            }
        }
        __syncthreads();
        if (tid == 0) {
            run_pre[0] = 0;
            for (int k = 0; k < ncol; k++) run_pre[k + 1] += run_pre[k];
        }
        __syncthreads();
        const int total = run_pre[ncol];

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
                // a point whose G2P stencil left the window goes to the
                // deferred list, handled by the launch's last workgroup
                if (!g2p_point(c, cf, gr, gp, nullptr, vw, wc0, wr0, win_w, win_h, WS, FROM_M, xq, yq, v, C, D, Jm1, ud,
                               pp, qq, e, (unsigned)pt, (unsigned)step, terms)) {
                    if (e) {
                        err = max(err, e);
                    } else {
                        const int slot = atomicAdd(a.defer_count, 1);
                        s_defer = 1;
                        a.defer_pts[slot] = (int)pt;
                        a.defer_scene[slot] = ti.scene;
                    }
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
                    lds_d *const p0 = acc_lane + 3 * (lx * WS + ly);
                    st.nodes(cf, [&](int i, int j, double dm, double dpx, double dpy) {
                        lds_d *const p = p0 + 3 * (i * WS + j);
                        lds_add(p, dm);
                        lds_add(p + 1, dpx);
                        lds_add(p + 2, dpy);
                    });
                } else {
                    // the stencil leaves the window (a point that drifted since the sort)
                    st.nodes(cf, [&](int i, int j, double dm, double dpx, double dpy) {
                        if (lx + i >= 0 && lx + i < win_w && ly + j >= 0 && ly + j < win_h) {
                            lds_d *const p = acc_lane + 3 * ((lx + i) * WS + ly + j);
                            lds_add(p, dm);
                            lds_add(p + 1, dpx);
                            lds_add(p + 2, dpy);
                        } else {
                            const long long gi = (long long)(st.bxc + i) * gridY + st.by + j;
                            grid_add(gw, gp, gi, dm, dpx, dpy);
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
                const int qc = W.col_of(i), qr = i - qc * win_h, li = qc * WS + qr;
                const double m = acc[3 * li] + acc[3 * WNmax + 3 * li];
                if (m == 0) continue;
                const int col = wc0 + qc, row = wr0 + qr;
                if (col < 0 || col >= c.ncols || row < 0 || row >= gridY) continue;
                const long long gi = (long long)col * gridY + row;
                grid_add(gw, gp, gi, m, acc[3 * li + 1] + acc[3 * WNmax + 3 * li + 1],
                         acc[3 * li + 2] + acc[3 * WNmax + 3 * li + 2]);
                dflag[gi >> kSegShift] = 1;
            }
        }
        __syncthreads();
    }
    if (dyn && tid == 0) {
        if (atomicAdd(a.tile_queue + 8, 1) == (int)gridDim.x - 1)
            for (int k = 0; k < 9; k++) a.tile_queue[k] = 0;
    }
    if (FROM_M) {
        // The points this launch set aside are handled by the last workgroup
        // to finish (their G2P reads grid[rbuf], complete before the launch;
        // their P2G adds into grid[wbuf] with atomics).  A workgroup that set
        // some aside publishes the list entries with a device-scope fence
        // before it counts itself done; the others (all of them in a step
        // where no point left its window) need none.
        __shared__ int s_last;
        __syncthreads();
        if (tid == 0) {
            if (s_defer) __threadfence();
            s_last = atomicAdd(a.defer_count + 1, 1) == (int)gridDim.x - 1;
        }
        __syncthreads();
        if (s_last) {
            const int n = __hip_atomic_load(a.defer_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (n > 0) {
                __threadfence();   // the entries the other workgroups published
                for (int k = tid; k < n; k += kBlock) deferred_point(a, DO_P2G, rbuf, wbuf, k);
            }
            __syncthreads();
            if (tid == 0) {
                a.defer_count[0] = 0;
                a.defer_count[1] = 0;
            }
        }
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
static hipError_t launch_tiles_wh(const StepArgs &a, int rbuf, int wbuf, int zbuf, int record_mode, int n_blocks,
                                  size_t lds, hipStream_t s)
{
    switch (a.win_h) {
    case 20: hipLaunchKernelGGL((k_tiles<MODE, 20>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, zbuf, record_mode); break;
    case 36: hipLaunchKernelGGL((k_tiles<MODE, 36>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, zbuf, record_mode); break;
    case 68: hipLaunchKernelGGL((k_tiles<MODE, 68>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, zbuf, record_mode); break;
    case 18: hipLaunchKernelGGL((k_tiles<MODE, 18>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, zbuf, record_mode); break;
    case 34: hipLaunchKernelGGL((k_tiles<MODE, 34>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, zbuf, record_mode); break;
    case 66: hipLaunchKernelGGL((k_tiles<MODE, 66>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, zbuf, record_mode); break;
    default: hipLaunchKernelGGL((k_tiles<MODE, 0>), dim3(n_blocks), dim3(kBlock), lds, s, a, rbuf, wbuf, zbuf, record_mode); break;
    }
    return hipGetLastError();
}

hipError_t launch_tiles(const StepArgs &a, int mode, int rbuf, int wbuf, int zbuf, int record_mode, int n_blocks,
                        hipStream_t s)
{
    if (a.tx > kMaxTX || n_blocks <= 0) return hipErrorInvalidValue;
    const size_t lds = tile_lds_bytes(a, mode);
    switch (mode) {
    case kTileG2P: return launch_tiles_wh<kTileG2P>(a, rbuf, wbuf, zbuf, record_mode, n_blocks, lds, s);
    case kTileP2G: return launch_tiles_wh<kTileP2G>(a, rbuf, wbuf, zbuf, record_mode, n_blocks, lds, s);
    case kTileTail: return launch_tiles_wh<kTileTail>(a, rbuf, wbuf, zbuf, record_mode, n_blocks, lds, s);
    case kTileChain: return launch_tiles_wh<kTileChain>(a, rbuf, wbuf, zbuf, record_mode, n_blocks, lds, s);
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
    case kTileTail: return resident_wh<kTileTail>(a, lds);
    case kTileChain: return resident_wh<kTileChain>(a, lds);
    default: return 0;
    }
}

}  // namespace mixed
}  // namespace icempm
